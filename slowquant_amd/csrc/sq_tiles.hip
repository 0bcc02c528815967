// Fused brick-column passes for gate layouts made of adjacent-pair gates
// (tUPS / QNP brick walls: sa_single(p,p+1), double(2p,2p+1,2p+2,2p+3)).
//
// A "stage" is a run of consecutive gates acting on a set of disjoint orbital
// pairs (p, p+1).  Under all gates of a stage a determinant can only move
// within the set of strings that keep, for every pair, the electron count
// c_j = n_p + n_{p+1}, and keep every unpaired orbital: per spin, strings
// group into hypercubes {p, p+1}^m over the m singly occupied pairs.  A tile
// = (one alpha group) x (a run of whole beta groups) is closed under the
// stage, so one launch loads each tile into LDS once, applies every gate of
// the stage there, and writes it back: one HBM pass per stage instead of one
// per rotation.  Tiles gather their columns from the canonical order through
// a position -> column table, so state vectors never change layout.
//
// Gates on different pairs commute, so the stage runs pair by pair.  For one
// pair j the gates only mix the <= 4 members of an "orbit" {alpha bit of j}
// x {beta bit of j} (other indices fixed); a thread owns an orbit, keeps its
// members in registers and applies all of pair j's gates (and, in the
// reverse sweep, their gradient dots) without synchronising; one barrier per
// pair.  Signs come from the pair-local bits only: the environment parity
// cancels for number-conserving spin parts.
//
// Per entry (one rotation string) and orbit type (alpha part of the pair
// empty / doubly occupied / single, same for beta) the host precomputes a
// 12-bit code: for each of the 6 member pairs {0,1},{2,3},{0,2},{1,3},{0,3},
// {1,2} whether the string rotates it and with which sign, so the device
// applies rotations with register indices fixed at compile time.
//
// Kernels:
//   k_stage_fwd     forward gates of a stage on one vector; dagger runs U^+;
//                   optional derivative insertion T_g after gate g
//   k_stage_rev     reverse sweep on (psi, lambda): per-entry partial dots
//                   <lambda|T psi>, then un-apply, for every gate of the stage
//   k_tile_grad     grad[g] = 2 * sum over tiles of the gate's partial dots

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>

#include "sq_internal.h"

namespace sq {

constexpr int kMaxPairs = 6;        // hypercubes of at most 64 strings per spin
constexpr int kTileMax = 4096;      // doubles per vector per tile
constexpr int kTileMaxCols = 1024;  // columns per tile
constexpr int kTB = 256;            // threads per tile block
constexpr int kMaxStageGates = 48;  // stage entries (sa_single counts twice)
constexpr int kMaxPairEntries = 8;  // entries per pair within a stage
constexpr int kTypes = 9;           // orbit types (alpha kind x beta kind)

struct LocalOps {
  int8_t n;
  int8_t orb[4];  // 0 = p, 1 = p+1
  int8_t cre[4];
};

struct StageEntry {
  int32_t gate;   // program gate index
  int32_t pair;   // pair index within the stage
  double scale;   // angle multiplier (|coefficient|)
  uint32_t code[kTypes];
};

struct StageDev {
  int npairs;
  int nentries;
  int ntiles;
  int nsb;
  int pe[kMaxPairs + 1];     // entries of pair j: [pe[j], pe[j+1]) (entries sorted by pair)
  int64_t partial_off;       // this stage's block of the partial-dot buffer
  const int32_t* arows;      // concatenated alpha rows (canonical alpha index), groups sorted by hypercube index
  const int32_t* agoff;      // [ngroups+1]
  const uint32_t* agocc;     // per alpha group: 2 bits per pair (local occupancy of non-single pairs)
  const uint32_t* agbit;     // per alpha group: 4 bits per pair: hypercube bit, 0xF if not single
  const uint64_t* bdesc;     // per beta stage position: bits(24) | occ(12) << 24 | offset in group << 36
  const int32_t* bcol;       // per beta stage position: canonical beta column
  const int4* tiles;         // (alpha group, col0, col1, -)
  const StageEntry* entries;
};

struct Stage {
  bool tiled = false;
  int g0 = 0, g1 = 0;  // program gates [g0, g1)
  int npairs = 0;
  int pp[kMaxPairs];
  std::vector<StageEntry> entries;
  StageDev dev{};
  std::vector<void*> allocs;
};

struct GateRed {
  int64_t off;  // partial_off of the gate's stage
  int32_t ntiles, stride, k0, k1, gate, pad;
};

struct TilePlan {
  std::vector<Stage> stages;
  double* d_partial = nullptr;
  int64_t partial_len = 0;
  GateRed* d_red = nullptr;
  int nred = 0;
  ~TilePlan() {
    for (auto& s : stages)
      for (void* p : s.allocs) (void)hipFree(p);
    if (d_partial) (void)hipFree(d_partial);
    if (d_red) (void)hipFree(d_red);
  }
};

// ---------------------------------------------------------------------------
// Host: stage construction.

static bool local_ops(const SpinOps& o, int p, LocalOps& out) {
  out.n = (int8_t)o.n;
  int nc = 0, na = 0;
  if (o.n > 4) return false;
  for (int j = 0; j < o.n; ++j) {
    int orb = o.orb[j];
    if (orb != p && orb != p + 1) return false;
    out.orb[j] = (int8_t)(orb - p);
    out.cre[j] = o.cre[j];
    (o.cre[j] ? nc : na)++;
  }
  return nc == na;
}

// Apply local ops to a 2-bit pair occupancy m (bit 0 = orbital p); false if zero.
static bool local_apply(const LocalOps& o, uint32_t& m, int& par) {
  for (int j = o.n - 1; j >= 0; --j) {
    uint32_t bit = 1u << o.orb[j];
    par += __builtin_popcount(m & (bit - 1u));
    if (o.cre[j]) {
      if (m & bit) return false;
      m |= bit;
    } else {
      if (!(m & bit)) return false;
      m &= ~bit;
    }
  }
  return true;
}

// Member pairs of an orbit (member o = alpha bit | beta bit << 1).
constexpr int kSlotA[6] = {0, 2, 0, 1, 0, 1};
constexpr int kSlotB[6] = {1, 3, 2, 3, 3, 2};

static int slot_of(int a, int b) {
  if (a > b) std::swap(a, b);
  for (int k = 0; k < 6; ++k)
    if (kSlotA[k] == a && kSlotB[k] == b) return k;
  return -1;
}

// Orbit-type codes of one rotation string with local spin parts la/lb.
static bool entry_codes(const LocalOps& la, const LocalOps& lb, double sgn, uint32_t* code) {
  for (int ta = 0; ta < 3; ++ta) {
    for (int tb = 0; tb < 3; ++tb) {
      uint32_t c = 0;
      for (int o = 0; o < 4; ++o) {
        const int ba = o & 1, bb = o >> 1;
        if ((ta != 2 && ba) || (tb != 2 && bb)) continue;
        uint32_t ma = ta == 2 ? (ba ? 2u : 1u) : (ta == 0 ? 0u : 3u);
        uint32_t mb = tb == 2 ? (bb ? 2u : 1u) : (tb == 0 ? 0u : 3u);
        const uint32_t ma0 = ma, mb0 = mb;
        int par = 0;
        if (!local_apply(la, ma, par) || !local_apply(lb, mb, par)) continue;
        if ((ta != 2 && ma != ma0) || (tb != 2 && mb != mb0)) return false;
        const int ba2 = ta == 2 ? (ma == 2u) : ba;
        const int bb2 = tb == 2 ? (mb == 2u) : bb;
        const int o2 = ba2 | (bb2 << 1);
        if (o2 == o) return false;
        const int sl = slot_of(o, o2);
        double sig = ((par & 1) ? -1.0 : 1.0) * sgn;
        if (o > o2) sig = -sig;  // orientation: slot rotations run from the lower member
        if (c & (3u << (2 * sl))) return false;
        c |= (1u | (sig < 0 ? 2u : 0u)) << (2 * sl);
      }
      code[ta * 3 + tb] = c;
    }
  }
  return true;
}

// Can gate g run inside a tile stage?  Fills its entries (pair index unset).
static bool tile_entries(const Program& P, int g, int& p_out, std::vector<StageEntry>& out) {
  const GateDev& gd = P.gates[g];
  int p = P.pair_p[g];
  if (p < 0) return false;
  if (gd.kind != GC_ROT && gd.kind != GC_SA_SINGLE) return false;
  int halves = gd.kind == GC_SA_SINGLE ? 2 : 1;
  out.clear();
  for (int h = 0; h < halves; ++h) {
    const Term& T = P.terms[gd.term_begin + 2 * h];
    StageEntry e{};
    e.gate = g;
    e.scale = halves == 2 ? 1.0 : gd.scale;
    LocalOps la, lb;
    if (!local_ops(T.a, p, la) || !local_ops(T.b, p, lb)) return false;
    if (!entry_codes(la, lb, T.coef < 0 ? -1.0 : 1.0, e.code)) return false;
    out.push_back(e);
  }
  p_out = p;
  return true;
}

struct GroupInfo {
  std::vector<int32_t> order;  // strings sorted by (key, h)
  std::vector<int32_t> start;  // group start per sorted position
  std::vector<int32_t> goff;   // group offsets
  std::vector<uint32_t> gocc, gbit;
  std::vector<uint32_t> occ, bit;  // per sorted position
};

static GroupInfo group_strings(const std::vector<uint32_t>& strs, int npairs, const int* pp) {
  uint32_t paired = 0;
  for (int j = 0; j < npairs; ++j) paired |= 3u << pp[j];
  const size_t n = strs.size();
  std::vector<std::pair<std::vector<uint32_t>, int32_t>> keyed(n);
  std::vector<uint32_t> hyper(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t s = strs[i];
    std::vector<uint32_t> key;
    uint32_t h = 0;
    int b = 0;
    for (int j = 0; j < npairs; ++j) {
      uint32_t loc = (s >> pp[j]) & 3u;
      int c = __builtin_popcount(loc);
      key.push_back((uint32_t)c);
      if (c == 1) {
        if (loc == 2u) h |= 1u << b;
        ++b;
      }
    }
    key.push_back(s & ~paired);
    hyper[i] = h;
    keyed[i] = {key, (int32_t)i};
  }
  std::vector<int32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) {
    if (keyed[x].first != keyed[y].first) return keyed[x].first < keyed[y].first;
    return hyper[x] < hyper[y];
  });
  GroupInfo G;
  G.order = idx;
  G.start.resize(n);
  G.occ.resize(n);
  G.bit.resize(n);
  for (size_t k = 0; k < n; ++k) {
    if (k == 0 || keyed[idx[k]].first != keyed[idx[k - 1]].first) {
      G.goff.push_back((int32_t)k);
      uint32_t s = strs[idx[k]];
      uint32_t gocc = 0, gbit = 0;
      int b = 0;
      for (int j = 0; j < npairs; ++j) {
        uint32_t loc = (s >> pp[j]) & 3u;
        if (__builtin_popcount(loc) == 1) {
          gbit |= (uint32_t)b << (4 * j);
          ++b;
        } else {
          gbit |= 0xFu << (4 * j);
          gocc |= loc << (2 * j);
        }
      }
      G.gocc.push_back(gocc);
      G.gbit.push_back(gbit);
    }
    G.start[k] = G.goff.back();
    uint32_t s = strs[idx[k]];
    uint32_t occ = 0;
    for (int j = 0; j < npairs; ++j) occ |= ((s >> pp[j]) & 3u) << (2 * j);
    G.occ[k] = occ;
    G.bit[k] = G.gbit.back();
  }
  G.goff.push_back((int32_t)n);
  return G;
}

template <class T>
static T* upload(std::vector<void*>& allocs, const std::vector<T>& v) {
  T* d = nullptr;
  SQ_HIP(hipMalloc(&d, sizeof(T) * std::max<size_t>(v.size(), 1)));
  allocs.push_back(d);
  if (!v.empty()) SQ_HIP(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return d;
}

static void finalize_stage(const Space& sp, Stage& S) {
  // entries grouped by pair, program order kept within a pair
  std::stable_sort(S.entries.begin(), S.entries.end(),
                   [](const StageEntry& a, const StageEntry& b) { return a.pair < b.pair; });
  for (int j = 0, k = 0; j <= S.npairs; ++j) {
    while (k < (int)S.entries.size() && S.entries[k].pair < j) ++k;
    S.dev.pe[j] = k;
  }
  GroupInfo A = group_strings(sp.h_astr, S.npairs, S.pp);
  GroupInfo B = group_strings(sp.h_bstr, S.npairs, S.pp);
  std::vector<int4> tiles;
  const int nag = (int)A.goff.size() - 1;
  const int nbg = (int)B.goff.size() - 1;
  for (int ga = 0; ga < nag; ++ga) {
    const int R = A.goff[ga + 1] - A.goff[ga];
    int c0 = 0;
    int gb = 0;
    while (gb < nbg) {
      int c1 = B.goff[gb + 1];
      ++gb;
      while (gb < nbg) {
        int nc1 = B.goff[gb + 1];
        if ((int64_t)R * (nc1 - c0) > kTileMax || nc1 - c0 > kTileMaxCols) break;
        c1 = nc1;
        ++gb;
      }
      SQ_CHECK((int64_t)R * (c1 - c0) <= kTileMax && c1 - c0 <= kTileMaxCols && R <= 64, SQ_ERR_ARG,
               "tile does not fit: hypercube groups too large");
      tiles.push_back(int4{ga, c0, c1, 0});
      c0 = c1;
    }
  }
  std::vector<uint64_t> bdesc(sp.nsb);
  for (int k = 0; k < sp.nsb; ++k)
    bdesc[k] = (uint64_t)B.bit[k] | ((uint64_t)B.occ[k] << 24) | ((uint64_t)(k - B.start[k]) << 36);
  S.dev.npairs = S.npairs;
  S.dev.nentries = (int)S.entries.size();
  S.dev.ntiles = (int)tiles.size();
  S.dev.nsb = sp.nsb;
  S.dev.arows = upload(S.allocs, A.order);
  S.dev.agoff = upload(S.allocs, A.goff);
  S.dev.agocc = upload(S.allocs, A.gocc);
  S.dev.agbit = upload(S.allocs, A.gbit);
  S.dev.bdesc = upload(S.allocs, bdesc);
  S.dev.bcol = upload(S.allocs, B.order);
  S.dev.tiles = upload(S.allocs, tiles);
  S.dev.entries = upload(S.allocs, S.entries);
}

TilePlan* build_tile_plan(const Space& sp, const Program& P) {
  auto* T = new TilePlan();
  try {
    const int n = (int)P.gates.size();
    int g = 0;
    std::vector<StageEntry> ent;
    while (g < n) {
      int p;
      if (!tile_entries(P, g, p, ent)) {
        Stage S;
        S.tiled = false;
        S.g0 = g;
        S.g1 = g + 1;
        T->stages.push_back(std::move(S));
        ++g;
        continue;
      }
      Stage S;
      S.tiled = true;
      S.g0 = g;
      S.npairs = 0;
      int per_pair[kMaxPairs] = {0};
      while (g < n) {
        if (!tile_entries(P, g, p, ent)) break;
        int j = -1;
        bool clash = false;
        for (int k = 0; k < S.npairs; ++k) {
          if (S.pp[k] == p) j = k;
          else if (std::abs(S.pp[k] - p) < 2) clash = true;
        }
        if (clash) break;
        if (j < 0 && S.npairs == kMaxPairs) break;
        if ((int)(S.entries.size() + ent.size()) > kMaxStageGates) break;
        if (j >= 0 && per_pair[j] + (int)ent.size() > kMaxPairEntries) break;
        if (j < 0) {
          j = S.npairs;
          S.pp[S.npairs++] = p;
        }
        per_pair[j] += (int)ent.size();
        for (auto& e : ent) {
          e.pair = j;
          S.entries.push_back(e);
        }
        ++g;
      }
      S.g1 = g;
      T->stages.push_back(std::move(S));
    }
    int64_t off = 0;
    std::vector<GateRed> red;
    for (auto& S : T->stages) {
      if (!S.tiled) continue;
      finalize_stage(sp, S);
      S.dev.partial_off = off;
      const int ne = S.dev.nentries;
      for (int k = 0; k < ne;) {
        int k1 = k + 1;
        while (k1 < ne && S.entries[k1].gate == S.entries[k].gate) ++k1;
        red.push_back(GateRed{off, S.dev.ntiles, ne, k, k1, S.entries[k].gate, 0});
        k = k1;
      }
      off += (int64_t)S.dev.ntiles * ne;
    }
    SQ_HIP(hipMalloc(&T->d_partial, sizeof(double) * std::max<int64_t>(off, 1)));
    T->partial_len = off;
    std::vector<void*> tmp;
    T->d_red = upload(tmp, red);
    T->nred = (int)red.size();
  } catch (...) {
    delete T;
    throw;
  }
  return T;
}

void destroy_tile_plan(TilePlan* T) { delete T; }

bool tile_plan_has_tiles(const TilePlan* T) {
  if (!T) return false;
  for (auto& S : T->stages)
    if (S.tiled) return true;
  return false;
}

// Tiled stages are used unless SQ_TILES=0 (column tables hold 16-bit indices).
const TilePlan* tiles_for(const Space& s, const Program& p) {
  if (!p.tiles_built) {
    p.tiles_built = true;
    const char* env = getenv("SQ_TILES");
    bool enabled = !(env && env[0] == '0');
    if (enabled && s.dim > 0 && s.nsb <= 65535) p.tiles = build_tile_plan(s, p);
    if (p.tiles && !tile_plan_has_tiles(p.tiles)) {
      destroy_tile_plan(p.tiles);
      p.tiles = nullptr;
    }
  }
  return p.tiles;
}

// ---------------------------------------------------------------------------
// Device kernels.

struct TileShared {
  int R, W, E, c0;
  float invW;
  uint32_t gocc, gbit;
  int32_t rows[64];
};

// Block-cooperative load of a tile's descriptors (row list, column table,
// per-entry cos/sin and codes) into LDS.
__device__ __forceinline__ void tile_setup(const StageDev& S, TileShared& t, uint64_t* cdesc, uint16_t* ccol,
                                           const double* __restrict__ theta, double sgn_th, double* cs_c,
                                           double* cs_s, uint32_t* codes) {
  const int4 T = S.tiles[blockIdx.x];
  const int r0 = S.agoff[T.x];
  const int R = S.agoff[T.x + 1] - r0;
  const int W = T.z - T.y;
  if (threadIdx.x == 0) {
    t.c0 = T.y;
    t.W = W;
    t.R = R;
    t.E = R * W;
    t.invW = 1.0f / (float)W;
    t.gocc = S.agocc[T.x];
    t.gbit = S.agbit[T.x];
  }
  for (int r = threadIdx.x; r < R; r += blockDim.x) t.rows[r] = S.arows[r0 + r];
  for (int c = threadIdx.x; c < W; c += blockDim.x) {
    cdesc[c] = S.bdesc[T.y + c];
    ccol[c] = (uint16_t)S.bcol[T.y + c];
  }
  for (int k = threadIdx.x; k < S.nentries; k += blockDim.x) {
    const StageEntry& en = S.entries[k];
    const double th = theta[en.gate];
    double s = 0.0, c = 1.0;
    if (fabs(th) >= 1e-28) sincos(sgn_th * en.scale * th, &s, &c);
    cs_c[k] = c;
    cs_s[k] = s;
  }
  for (int i = threadIdx.x; i < S.nentries * kTypes; i += blockDim.x) codes[i] = S.entries[i / kTypes].code[i % kTypes];
  __syncthreads();
}

// Orbit of pair j owned by work item i: member element offsets in the tile
// (-1 where a member does not exist) and the orbit type; false when item i is
// not an orbit representative.
struct Orbit {
  int m0, m1, m2, m3;
  int type;
};

__device__ __forceinline__ bool orbit_of(const TileShared& t, const uint64_t* cdesc, int j, int i, Orbit& ob) {
  const uint32_t abit = (t.gbit >> (4 * j)) & 0xFu;
  const bool asingle = abit != 0xFu;
  const int rr = (int)(((float)i + 0.5f) * t.invW);
  const int c = i - rr * t.W;
  const int r0 = asingle ? (((rr >> abit) << (abit + 1)) | (rr & ((1 << abit) - 1))) : rr;
  const uint64_t d = cdesc[c];
  const uint32_t bb = (uint32_t)(d >> (4 * j)) & 0xFu;
  const bool bsingle = bb != 0xFu;
  if (bsingle && ((d >> (36 + bb)) & 1u)) return false;
  const int ta = asingle ? 2 : (((t.gocc >> (2 * j)) & 3u) == 0 ? 0 : 1);
  const int tb = bsingle ? 2 : ((((uint32_t)(d >> (24 + 2 * j))) & 3u) == 0 ? 0 : 1);
  ob.type = ta * 3 + tb;
  const int W = t.W;
  ob.m0 = r0 * W + c;
  const int dr = asingle ? (W << abit) : 0;
  const int dc = bsingle ? (1 << bb) : 0;
  ob.m1 = asingle ? ob.m0 + dr : -1;
  ob.m2 = bsingle ? ob.m0 + dc : -1;
  ob.m3 = (asingle && bsingle) ? ob.m0 + dr + dc : -1;
  return true;
}

__device__ __forceinline__ void rot(double& a, double& b, double c, double s) {
  const double x = a, y = b;
  a = c * x - s * y;
  b = s * x + c * y;
}

// Rotation by (c, s) of every member pair that a code marks (sign per pair).
__device__ __forceinline__ void apply_code(uint32_t code, double c, double s, double (&v)[4]) {
#pragma unroll
  for (int sl = 0; sl < 6; ++sl) {
    const uint32_t f = (code >> (2 * sl)) & 3u;
    if (f & 1u) rot(v[kSlotA[sl]], v[kSlotB[sl]], c, (f & 2u) ? -s : s);
  }
}

// out = T v for one entry's code (T|a> = sg|b>, T|b> = -sg|a>), scaled.
__device__ __forceinline__ void t_code(uint32_t code, double scale, const double (&v)[4], double (&out)[4]) {
#pragma unroll
  for (int sl = 0; sl < 6; ++sl) {
    const uint32_t f = (code >> (2 * sl)) & 3u;
    if (f & 1u) {
      const double sg = (f & 2u) ? -scale : scale;
      out[kSlotB[sl]] += sg * v[kSlotA[sl]];
      out[kSlotA[sl]] -= sg * v[kSlotB[sl]];
    }
  }
}

// Contribution of one entry's code to <lam|T psi>.
__device__ __forceinline__ double dot_code(uint32_t code, const double (&p)[4], const double (&l)[4]) {
  double d = 0.0;
#pragma unroll
  for (int sl = 0; sl < 6; ++sl) {
    const uint32_t f = (code >> (2 * sl)) & 3u;
    if (f & 1u) {
      const double x = l[kSlotB[sl]] * p[kSlotA[sl]] - l[kSlotA[sl]] * p[kSlotB[sl]];
      d += (f & 2u) ? -x : x;
    }
  }
  return d;
}

__device__ __forceinline__ void load_orbit(const double* X, const Orbit& ob, double (&v)[4]) {
  v[0] = X[ob.m0];
  v[1] = ob.m1 >= 0 ? X[ob.m1] : 0.0;
  v[2] = ob.m2 >= 0 ? X[ob.m2] : 0.0;
  v[3] = ob.m3 >= 0 ? X[ob.m3] : 0.0;
}

__device__ __forceinline__ void store_orbit(double* X, const Orbit& ob, const double (&v)[4]) {
  X[ob.m0] = v[0];
  if (ob.m1 >= 0) X[ob.m1] = v[1];
  if (ob.m2 >= 0) X[ob.m2] = v[2];
  if (ob.m3 >= 0) X[ob.m3] = v[3];
}

// Gates of one stage on one vector (canonical order, in place).
// dagger: U^+ (pair entries reversed, -theta).  deriv_gate: insert T_g after gate g.
__global__ __launch_bounds__(kTB) void k_stage_fwd(StageDev S, const double* __restrict__ theta, int dagger,
                                                   double* __restrict__ psi, int deriv_gate) {
  __shared__ double X[kTileMax];
  __shared__ uint64_t cdesc[kTileMaxCols];
  __shared__ uint16_t ccol[kTileMaxCols];
  __shared__ double cs_c[kMaxStageGates], cs_s[kMaxStageGates];
  __shared__ uint32_t codes[kMaxStageGates * kTypes];
  __shared__ TileShared t;
  tile_setup(S, t, cdesc, ccol, theta, dagger ? -1.0 : 1.0, cs_c, cs_s, codes);
  const int W = t.W, E = t.E, nsb = S.nsb;
  const float invW = t.invW;
  for (int e = threadIdx.x; e < E; e += kTB) {
    const int r = (int)(((float)e + 0.5f) * invW);
    X[e] = psi[(size_t)t.rows[r] * nsb + ccol[e - r * W]];
  }
  __syncthreads();
  for (int j = 0; j < S.npairs; ++j) {
    const int k0 = S.pe[j], k1 = S.pe[j + 1];
    const int nent = k1 - k0;
    if (nent == 0) continue;
    const bool asingle = ((t.gbit >> (4 * j)) & 0xFu) != 0xFu;
    const int items = asingle ? E >> 1 : E;
    // derivative insertion: after the last entry of deriv_gate on this pair
    int dq0 = -1, dq1 = -1;
    if (deriv_gate >= 0) {
      for (int k = k0; k < k1; ++k)
        if (S.entries[k].gate == deriv_gate) {
          if (dq0 < 0) dq0 = k;
          dq1 = k + 1;
        }
    }
    for (int i = threadIdx.x; i < items; i += kTB) {
      Orbit ob;
      if (!orbit_of(t, cdesc, j, i, ob)) continue;
      double v[4];
      load_orbit(X, ob, v);
#pragma unroll
      for (int kk = 0; kk < kMaxPairEntries; ++kk) {
        if (kk < nent) {
          const int k = dagger ? k1 - 1 - kk : k0 + kk;
          apply_code(codes[k * kTypes + ob.type], cs_c[k], cs_s[k], v);
          if (k + 1 == dq1) {
            double o[4] = {0.0, 0.0, 0.0, 0.0};
            for (int q = dq0; q < dq1; ++q) t_code(codes[q * kTypes + ob.type], S.entries[q].scale, v, o);
#pragma unroll
            for (int m = 0; m < 4; ++m) v[m] = o[m];
          }
        }
      }
      store_orbit(X, ob, v);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < E; e += kTB) {
    const int r = (int)(((float)e + 0.5f) * invW);
    psi[(size_t)t.rows[r] * nsb + ccol[e - r * W]] = X[e];
  }
}

// Reverse sweep of one stage on (psi, lam), in place, with per-entry partial
// dots written to partial[partial_off + tile * nentries + k].
__global__ __launch_bounds__(kTB) void k_stage_rev(StageDev S, const double* __restrict__ theta,
                                                   double* __restrict__ psi, double* __restrict__ lam,
                                                   double* __restrict__ partial, int write_back) {
  __shared__ double X[kTileMax];
  __shared__ double L[kTileMax];
  __shared__ uint64_t cdesc[kTileMaxCols];
  __shared__ uint16_t ccol[kTileMaxCols];
  __shared__ double cs_c[kMaxStageGates], cs_s[kMaxStageGates];
  __shared__ uint32_t codes[kMaxStageGates * kTypes];
  __shared__ double red[kMaxStageGates][kTB / 64];
  __shared__ TileShared t;
  tile_setup(S, t, cdesc, ccol, theta, -1.0, cs_c, cs_s, codes);
  const int W = t.W, E = t.E, nsb = S.nsb;
  const float invW = t.invW;
  for (int e = threadIdx.x; e < E; e += kTB) {
    const int r = (int)(((float)e + 0.5f) * invW);
    const size_t gi = (size_t)t.rows[r] * nsb + ccol[e - r * W];
    X[e] = psi[gi];
    L[e] = lam[gi];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = 0; j < S.npairs; ++j) {
    const int k0 = S.pe[j], k1 = S.pe[j + 1];
    const int nent = k1 - k0;
    if (nent == 0) continue;
    const bool asingle = ((t.gbit >> (4 * j)) & 0xFu) != 0xFu;
    const int items = asingle ? E >> 1 : E;
    double dots[kMaxPairEntries];
#pragma unroll
    for (int kk = 0; kk < kMaxPairEntries; ++kk) dots[kk] = 0.0;
    for (int i = threadIdx.x; i < items; i += kTB) {
      Orbit ob;
      if (!orbit_of(t, cdesc, j, i, ob)) continue;
      double v[4], w[4];
      load_orbit(X, ob, v);
      load_orbit(L, ob, w);
#pragma unroll
      for (int kk = 0; kk < kMaxPairEntries; ++kk) {
        if (kk < nent) {
          const int k = k1 - 1 - kk;
          const uint32_t code = codes[k * kTypes + ob.type];
          dots[kk] += dot_code(code, v, w);
          apply_code(code, cs_c[k], cs_s[k], v);
          apply_code(code, cs_c[k], cs_s[k], w);
        }
      }
      if (write_back) {
        store_orbit(X, ob, v);
        store_orbit(L, ob, w);
      }
    }
#pragma unroll
    for (int kk = 0; kk < kMaxPairEntries; ++kk) {
      if (kk < nent) {
        double d = dots[kk];
        for (int off = 32; off > 0; off >>= 1) d += __shfl_down(d, off, 64);
        if (lane == 0) red[k1 - 1 - kk][wave] = d * S.entries[k1 - 1 - kk].scale;
      }
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < S.nentries; k += kTB) {
    double v = 0.0;
    for (int w = 0; w < kTB / 64; ++w) v += red[k][w];
    partial[S.partial_off + (size_t)blockIdx.x * S.nentries + k] = v;
  }
  if (!write_back) return;  // the first stage of the program: the un-applied states are not needed
  for (int e = threadIdx.x; e < E; e += kTB) {
    const int r = (int)(((float)e + 0.5f) * invW);
    const size_t gi = (size_t)t.rows[r] * nsb + ccol[e - r * W];
    psi[gi] = X[e];
    lam[gi] = L[e];
  }
}

// grad[gate] = 2 * sum over the stage's tiles of the gate's entry partials;
// one block per tiled gate.
__global__ void k_tile_grad(const double* __restrict__ partial, const GateRed* __restrict__ red,
                            double* __restrict__ grad) {
  __shared__ double rs[kTB / 64];
  const GateRed R = red[blockIdx.x];
  double v = 0.0;
  for (int i = threadIdx.x; i < R.ntiles; i += blockDim.x)
    for (int q = R.k0; q < R.k1; ++q) v += partial[R.off + (size_t)i * R.stride + q];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) rs[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += rs[w];
    grad[R.gate] = 2.0 * tot;
  }
}

// ---------------------------------------------------------------------------
// Host orchestration.

static double stage_bytes(const Space& s, int vectors) { return 16.0 * (double)s.dim * vectors; }

// Forward sweep (dagger: U^+, gates reversed with -theta) over the whole
// program, tiles where possible, in place on psi (canonical order).
// scratch >= dim doubles (derivative insertion on per-gate stages);
// gwork >= 2*dim doubles for gates that run through the per-gate kernels.
// deriv_gate >= 0 inserts T_g right after gate g (plain gradient; forward only).
void tiles_forward(const Space& s, const Program& p, const TilePlan& T, const double* theta, double* psi,
                   double* scratch, double* gwork, int deriv_gate, bool dagger, hipStream_t st) {
  const int ns = (int)T.stages.size();
  for (int ii = 0; ii < ns; ++ii) {
    const int i = dagger ? ns - 1 - ii : ii;
    const Stage& S = T.stages[i];
    if (!S.tiled) {
      apply_gate_impl(s, p, S.g0, theta, dagger, psi, gwork, st);
      if (S.g0 == deriv_gate) {
        grad_action_impl(s, p, S.g0, psi, scratch, st);
        copy_vec(scratch, psi, s.dim, st);
      }
      continue;
    }
    const int dg = (deriv_gate >= S.g0 && deriv_gate < S.g1) ? deriv_gate : -1;
    SQ_PROF(PC_TILE_FWD, st, stage_bytes(s, 1), 0.0,
            k_stage_fwd<<<S.dev.ntiles, kTB, 0, st>>>(S.dev, theta, dagger ? 1 : 0, psi, dg));
  }
}

// Reverse sweep on (psi, lam) with gradient dots into grad_dev[g] (device).
// gwork >= 2*dim; partial: reduction scratch of the per-gate kernels.
void tiles_reverse(const Space& s, const Program& p, const TilePlan& T, const double* theta, double* psi, double* lam,
                   double* sp_, double* sl_, double* gwork, double* partial, double* grad_dev, hipStream_t st) {
  (void)sp_;
  (void)sl_;
  for (int i = (int)T.stages.size() - 1; i >= 0; --i) {
    const Stage& S = T.stages[i];
    if (!S.tiled) {
      reverse_step_impl(s, p, S.g0, theta, psi, lam, grad_dev, gwork, partial, st);
      continue;
    }
    const int wb = i > 0 ? 1 : 0;
    SQ_PROF(PC_TILE_REV, st, stage_bytes(s, wb ? 2 : 1), 0.0,
            k_stage_rev<<<S.dev.ntiles, kTB, 0, st>>>(S.dev, theta, psi, lam, T.d_partial, wb));
  }
  if (T.nred > 0)
    SQ_PROF(PC_REDUCE, st, 8.0 * (double)T.partial_len, 0.0,
            k_tile_grad<<<T.nred, kTB, 0, st>>>(T.d_partial, T.d_red, grad_dev));
}

}  // namespace sq
