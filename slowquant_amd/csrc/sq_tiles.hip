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
// the stage there (pair rotations with signs from the pair-local bits only:
// the environment parity cancels for number-conserving spin parts), and
// writes it back: one HBM pass per stage instead of one per rotation.
//
// Alpha groups are addressed through row lists, so alpha needs no data
// movement.  Beta groups are contiguous only in the stage's own column order
// ("stage order"); a row-wise permutation pass converts between orders.
//
// Kernels:
//   k_perm_rows   out[a][k] = in[a][src[k]]           (LDS-staged rows)
//   k_stage_fwd   forward gates of a stage on one vector, optional
//                 derivative insertion T_g after gate g (plain gradient)
//   k_stage_rev   reverse sweep on (psi, lambda): per-gate partial dots
//                 <lambda|T psi>, then un-apply, for every gate of the stage
//   k_stage_reduce  grad[g] = 2 * sum over tiles of a gate's partials

#include <algorithm>
#include <cmath>
#include <map>
#include <cstdlib>
#include <numeric>

#include "sq_internal.h"

namespace sq {

constexpr int kMaxPairs = 6;       // hypercubes of at most 64 strings per spin
constexpr int kTileMax = 4096;     // doubles per vector per tile
constexpr int kTileMaxCols = 1024; // columns per tile
constexpr int kTB = 256;           // threads per tile block
constexpr int kMaxStageGates = 48; // stage entries (sa_single counts twice)

struct LocalOps {
  int8_t n;
  int8_t orb[4];  // 0 = p, 1 = p+1
  int8_t cre[4];
};

// Action of one spin part on the 2-bit local occupancy m of a pair (bit 0 =
// orbital p, bit 1 = orbital p+1): nibble m of the map holds
// valid << 3 | parity << 2 | image.
struct StageEntry {
  int32_t gate;   // program gate index
  int32_t pair;   // pair index within the stage
  uint32_t amap, bmap;
  double sgn;     // sign of the string coefficient
  double scale;   // |coefficient| (angle multiplier)
};

static uint32_t local_map(const LocalOps& o) {
  uint32_t map = 0;
  for (uint32_t m0 = 0; m0 < 4; ++m0) {
    uint32_t m = m0;
    int par = 0;
    bool ok = true;
    for (int j = o.n - 1; j >= 0 && ok; --j) {
      uint32_t bit = 1u << o.orb[j];
      par += __builtin_popcount(m & (bit - 1u));
      if (o.cre[j]) {
        if (m & bit) ok = false;
        m |= bit;
      } else {
        if (!(m & bit)) ok = false;
        m &= ~bit;
      }
    }
    if (ok) map |= (8u | ((uint32_t)(par & 1) << 2) | m) << (4 * m0);
  }
  return map;
}

struct StageDev {
  int npairs;
  int nentries;
  int ntiles;
  int nsb;
  const int32_t* arows;     // concatenated alpha rows (canonical alpha index)
  const int32_t* agoff;     // [ngroups+1]
  const uint32_t* agocc;    // per group: 2 bits per pair (fixed occupancy of non-single pairs)
  const uint32_t* agbit;    // per group: 4 bits per pair: hypercube bit, 0xF if not single
  const uint32_t* bocc;     // per beta position: 2 bits per pair (local occupancy)
  const uint32_t* bbit;     // per beta position: 4 bits per pair
  const int32_t* bstart;    // per beta position: start position of its group
  const int4* tiles;        // (alpha group, col0, col1, tile index)
  const StageEntry* entries;
};

struct Stage {
  bool tiled = false;
  int g0 = 0, g1 = 0;  // program gates [g0, g1)
  int npairs = 0;
  int pp[kMaxPairs];
  std::vector<StageEntry> entries;
  std::vector<int32_t> bperm, bpos;  // position -> canonical beta, canonical -> position
  // device
  StageDev dev{};
  std::vector<void*> allocs;
  int32_t* d_from_canon = nullptr;  // src table: stage order <- canonical
  int32_t* d_to_canon = nullptr;    // src table: canonical <- stage order
  int32_t* d_from_prev = nullptr;   // src table: this stage <- previous tiled stage order (or canonical)
  int32_t* d_to_prev = nullptr;     // src table: previous stage order <- this stage
  int prev_tiled = -1;              // index of previous tiled stage if directly adjacent
};

struct TilePlan {
  std::vector<Stage> stages;
  double* d_partial = nullptr;  // [max tiles * kMaxStageGates]
  int64_t partial_len = 0;
  ~TilePlan() {
    for (auto& s : stages)
      for (void* p : s.allocs) (void)hipFree(p);
    if (d_partial) (void)hipFree(d_partial);
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
    e.sgn = T.coef < 0 ? -1.0 : 1.0;
    e.scale = halves == 2 ? 1.0 : gd.scale;
    LocalOps la, lb;
    if (!local_ops(T.a, p, la) || !local_ops(T.b, p, lb)) return false;
    e.amap = local_map(la);
    e.bmap = local_map(lb);
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
      // group-level descriptors
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
static T* upload(Stage& S, const std::vector<T>& v) {
  T* d = nullptr;
  SQ_HIP(hipMalloc(&d, sizeof(T) * std::max<size_t>(v.size(), 1)));
  if (!v.empty()) SQ_HIP(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  S.allocs.push_back(d);
  return d;
}

static void finalize_stage(const Space& sp, Stage& S) {
  GroupInfo A = group_strings(sp.h_astr, S.npairs, S.pp);
  GroupInfo B = group_strings(sp.h_bstr, S.npairs, S.pp);
  S.bperm = B.order;
  S.bpos.assign(sp.nsb, 0);
  for (int k = 0; k < sp.nsb; ++k) S.bpos[S.bperm[k]] = k;
  // tiles: each alpha group x runs of whole beta groups
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
      SQ_CHECK((int64_t)R * (c1 - c0) <= kTileMax && c1 - c0 <= kTileMaxCols, SQ_ERR_ARG,
               "tile does not fit: hypercube groups too large");
      tiles.push_back(int4{ga, c0, c1, (int)tiles.size()});
      c0 = c1;
    }
  }
  std::vector<int32_t> arows(A.order.begin(), A.order.end());
  S.dev.npairs = S.npairs;
  S.dev.nentries = (int)S.entries.size();
  S.dev.ntiles = (int)tiles.size();
  S.dev.nsb = sp.nsb;
  S.dev.arows = upload(S, arows);
  S.dev.agoff = upload(S, A.goff);
  S.dev.agocc = upload(S, A.gocc);
  S.dev.agbit = upload(S, A.gbit);
  S.dev.bocc = upload(S, B.occ);
  S.dev.bbit = upload(S, B.bit);
  S.dev.bstart = upload(S, B.start);
  S.dev.tiles = upload(S, tiles);
  S.dev.entries = upload(S, S.entries);
  // permutation source tables for "out[k] = in[src[k]]"
  std::vector<int32_t> from_canon(sp.nsb), to_canon(sp.nsb);
  for (int k = 0; k < sp.nsb; ++k) from_canon[k] = S.bperm[k];  // stage pos k holds canonical bperm[k]
  for (int b = 0; b < sp.nsb; ++b) to_canon[b] = S.bpos[b];
  S.d_from_canon = upload(S, from_canon);
  S.d_to_canon = upload(S, to_canon);
}

static void link_stages(const Space& sp, TilePlan& T) {
  for (size_t i = 1; i < T.stages.size(); ++i) {
    Stage& cur = T.stages[i];
    Stage& prv = T.stages[i - 1];
    if (!cur.tiled || !prv.tiled) continue;
    cur.prev_tiled = (int)i - 1;
    std::vector<int32_t> from_prev(sp.nsb), to_prev(sp.nsb);
    // cur pos k holds canonical cur.bperm[k], found in prev order at prv.bpos[...]
    for (int k = 0; k < sp.nsb; ++k) from_prev[k] = prv.bpos[cur.bperm[k]];
    for (int k = 0; k < sp.nsb; ++k) to_prev[k] = cur.bpos[prv.bperm[k]];
    cur.d_from_prev = upload(cur, from_prev);
    cur.d_to_prev = upload(cur, to_prev);
  }
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
      while (g < n) {
        if (!tile_entries(P, g, p, ent)) break;
        int j = -1;
        bool clash = false;
        for (int k = 0; k < S.npairs; ++k) {
          if (S.pp[k] == p) j = k;
          else if (std::abs(S.pp[k] - p) < 2) clash = true;
        }
        if (clash) break;
        if (j < 0) {
          if (S.npairs == kMaxPairs) break;
          j = S.npairs;
          S.pp[S.npairs++] = p;
        }
        if ((int)(S.entries.size() + ent.size()) > kMaxStageGates) break;
        for (auto& e : ent) {
          e.pair = j;
          S.entries.push_back(e);
        }
        ++g;
      }
      S.g1 = g;
      T->stages.push_back(std::move(S));
    }
    int64_t maxp = 1;
    for (auto& S : T->stages) {
      if (!S.tiled) continue;
      finalize_stage(sp, S);
      maxp = std::max<int64_t>(maxp, (int64_t)S.dev.ntiles * kMaxStageGates);
    }
    link_stages(sp, *T);
    SQ_HIP(hipMalloc(&T->d_partial, sizeof(double) * maxp));
    T->partial_len = maxp;
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

// Tiled stages are used unless SQ_TILES=0; spaces whose beta rows exceed the
// permutation kernel's LDS row buffer fall back to per-gate kernels.
const TilePlan* tiles_for(const Space& s, const Program& p) {
  if (!p.tiles_built) {
    p.tiles_built = true;
    const char* env = getenv("SQ_TILES");
    bool enabled = !(env && env[0] == '0');
    if (enabled && s.nsb <= 8192 && s.dim > 0) p.tiles = build_tile_plan(s, p);
    if (p.tiles && !tile_plan_has_tiles(p.tiles)) {
      destroy_tile_plan(p.tiles);
      p.tiles = nullptr;
    }
  }
  return p.tiles;
}

// ---------------------------------------------------------------------------
// Device kernels.

__device__ __forceinline__ bool apply_local(uint32_t map, uint32_t& m, int& par) {
  const uint32_t nib = (map >> (4 * m)) & 0xFu;
  if (!(nib & 8u)) return false;
  par += (nib >> 2) & 1u;
  m = nib & 3u;
  return true;
}

// out[a][k] = in[a][src[k]] for every alpha row a; one block per row.
__global__ void k_perm_rows(const double* __restrict__ in, double* __restrict__ out, const int32_t* __restrict__ src,
                            int nsb) {
  extern __shared__ double row[];
  const size_t base = (size_t)blockIdx.x * nsb;
  for (int k = threadIdx.x; k < nsb; k += blockDim.x) row[k] = in[base + k];
  __syncthreads();
  for (int k = threadIdx.x; k < nsb; k += blockDim.x) out[base + k] = row[src[k]];
}

struct TileCtx {
  int R, W, E, c0;
  uint32_t gocc, gbit;
  int32_t rows[64];
};

// Block-cooperative load of a tile's descriptors into LDS.
__device__ __forceinline__ void init_tile(const StageDev& S, TileCtx& t, uint32_t* cocc, uint32_t* cbit,
                                          int16_t* cstart) {
  const int4 T = S.tiles[blockIdx.x];
  const int r0 = S.agoff[T.x];
  const int R = S.agoff[T.x + 1] - r0;
  const int W = T.z - T.y;
  if (threadIdx.x == 0) {
    t.c0 = T.y;
    t.W = W;
    t.R = R;
    t.E = R * W;
    t.gocc = S.agocc[T.x];
    t.gbit = S.agbit[T.x];
  }
  for (int r = threadIdx.x; r < R; r += blockDim.x) t.rows[r] = S.arows[r0 + r];
  for (int c = threadIdx.x; c < W; c += blockDim.x) {
    cocc[c] = S.bocc[T.y + c];
    cbit[c] = S.bbit[T.y + c];
    cstart[c] = (int16_t)(S.bstart[T.y + c] - T.y);
  }
  __syncthreads();
}

// Locate the rotation partner of tile element e under entry en; false when
// the entry's string does not act on e (e is not the lower member of a pair).
__device__ __forceinline__ bool partner(const StageEntry& en, const TileCtx& t, const uint32_t* cocc,
                                        const uint32_t* cbit, const int16_t* cstart, int e, int& ep, double& sigma) {
  const int r = e / t.W;
  const int cc = e - r * t.W;
  const int j = en.pair;
  const uint32_t abit = (t.gbit >> (4 * j)) & 0xFu;
  uint32_t la = abit == 0xFu ? (t.gocc >> (2 * j)) & 3u : (((r >> abit) & 1) ? 2u : 1u);
  const uint32_t bb = (cbit[cc] >> (4 * j)) & 0xFu;
  uint32_t lb = (cocc[cc] >> (2 * j)) & 3u;
  const uint32_t la0 = la, lb0 = lb;
  int par = 0;
  if (!apply_local(en.amap, la, par)) return false;
  if (!apply_local(en.bmap, lb, par)) return false;
  int rp = r, cp = cc;
  if (la != la0) {
    if (abit == 0xFu) return false;  // cannot happen for number-conserving pair maps
    rp = r ^ (1 << abit);
  }
  if (lb != lb0) {
    if (bb == 0xFu) return false;
    cp = cstart[cc] + ((cc - cstart[cc]) ^ (1 << bb));
  }
  ep = rp * t.W + cp;
  sigma = ((par & 1) ? -1.0 : 1.0) * en.sgn;
  return true;
}

// Gates of one stage on one vector (in stage column order, in place).
// dagger: reverse entry order with -theta.  deriv_gate: insert T_g after gate g.
__global__ __launch_bounds__(kTB) void k_stage_fwd(StageDev S, const double* __restrict__ theta, int dagger,
                                                   double* __restrict__ psi, int deriv_gate) {
  __shared__ double A[kTileMax];
  __shared__ double Bv[kTileMax];
  __shared__ uint32_t cocc[kTileMaxCols], cbit[kTileMaxCols];
  __shared__ int16_t cstart[kTileMaxCols];
  __shared__ TileCtx t;
  init_tile(S, t, cocc, cbit, cstart);
  const int W = t.W, E = t.E, c0 = t.c0, nsb = S.nsb;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    const int r = e / W;
    A[e] = psi[(size_t)t.rows[r] * nsb + c0 + (e - r * W)];
  }
  __syncthreads();
  double* cur = A;
  double* oth = Bv;
  const double sgn_th = dagger ? -1.0 : 1.0;
  for (int kk = 0; kk < S.nentries; ++kk) {
    const int k = dagger ? S.nentries - 1 - kk : kk;
    const StageEntry en = S.entries[k];
    const double th0 = theta[en.gate];
    if (fabs(th0) >= 1e-28) {
      double s, c;
      sincos(sgn_th * en.scale * th0, &s, &c);
      for (int e = threadIdx.x; e < E; e += blockDim.x) {
        int ep;
        double sg;
        if (!partner(en, t, cocc, cbit, cstart, e, ep, sg)) continue;
        const double pd = cur[e], pj = cur[ep];
        cur[e] = c * pd - sg * s * pj;
        cur[ep] = sg * s * pd + c * pj;
      }
      __syncthreads();
    }
    if (deriv_gate >= 0 && en.gate == deriv_gate) {
      const bool last_of_gate = (k + 1 == S.nentries) || (S.entries[k + 1].gate != en.gate);
      if (last_of_gate) {
        for (int e = threadIdx.x; e < E; e += blockDim.x) oth[e] = 0.0;
        __syncthreads();
        int k0 = k;
        while (k0 > 0 && S.entries[k0 - 1].gate == en.gate) --k0;
        for (int q = k0; q <= k; ++q) {
          const StageEntry eq = S.entries[q];
          for (int e = threadIdx.x; e < E; e += blockDim.x) {
            int ep;
            double sg;
            if (!partner(eq, t, cocc, cbit, cstart, e, ep, sg)) continue;
            // T|D> = sg |J>,  T|J> = -sg |D>
            oth[ep] += eq.scale * sg * cur[e];
            oth[e] -= eq.scale * sg * cur[ep];
          }
          __syncthreads();
        }
        double* tmp = cur;
        cur = oth;
        oth = tmp;
      }
    }
  }
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    const int r = e / W;
    psi[(size_t)t.rows[r] * nsb + c0 + (e - r * W)] = cur[e];
  }
}

// Reverse sweep of one stage on (psi, lam), in place, with per-entry partial dots.
__global__ __launch_bounds__(kTB) void k_stage_rev(StageDev S, const double* __restrict__ theta,
                                                   double* __restrict__ psi, double* __restrict__ lam,
                                                   double* __restrict__ partial, int write_back) {
  __shared__ double A[kTileMax];
  __shared__ double L[kTileMax];
  __shared__ uint32_t cocc[kTileMaxCols], cbit[kTileMaxCols];
  __shared__ int16_t cstart[kTileMaxCols];
  __shared__ TileCtx t;
  __shared__ double red[kMaxStageGates][kTB / 64];
  init_tile(S, t, cocc, cbit, cstart);
  const int W = t.W, E = t.E, c0 = t.c0, nsb = S.nsb;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    const int r = e / W;
    const size_t gi = (size_t)t.rows[r] * nsb + c0 + (e - r * W);
    A[e] = psi[gi];
    L[e] = lam[gi];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = S.nentries - 1; k >= 0; --k) {
    const StageEntry en = S.entries[k];
    const double th0 = theta[en.gate];
    double s, c;
    sincos(-en.scale * th0, &s, &c);
    const bool skip = fabs(th0) < 1e-28;
    double dot = 0.0;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      int ep;
      double sg;
      if (!partner(en, t, cocc, cbit, cstart, e, ep, sg)) continue;
      const double pd = A[e], pj = A[ep], ld = L[e], lj = L[ep];
      dot += en.scale * sg * (lj * pd - ld * pj);
      if (!skip) {
        A[e] = c * pd - sg * s * pj;
        A[ep] = sg * s * pd + c * pj;
        L[e] = c * ld - sg * s * lj;
        L[ep] = sg * s * ld + c * lj;
      }
    }
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
    if (lane == 0) red[k][wave] = dot;
    __syncthreads();
  }
  for (int k = threadIdx.x; k < S.nentries; k += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < kTB / 64; ++w) v += red[k][w];
    partial[(size_t)blockIdx.x * kMaxStageGates + k] = v;
  }
  if (!write_back) return;  // the first stage of the program: the un-applied states are not needed
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    const int r = e / W;
    const size_t gi = (size_t)t.rows[r] * nsb + c0 + (e - r * W);
    psi[gi] = A[e];
    lam[gi] = L[e];
  }
}

// grad[gate] = 2 * sum_tiles sum_{entries of gate} partial; one block per entry
// group (entries of one gate are consecutive).
__global__ void k_stage_reduce(const double* __restrict__ partial, int ntiles, const StageEntry* __restrict__ entries,
                               int nentries, double* __restrict__ grad) {
  __shared__ double red[kTB / 64];
  const int k = blockIdx.x;
  if (k > 0 && entries[k - 1].gate == entries[k].gate) return;  // handled by the first entry of the gate
  int k1 = k + 1;
  while (k1 < nentries && entries[k1].gate == entries[k].gate) ++k1;
  double v = 0.0;
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x)
    for (int q = k; q < k1; ++q) v += partial[(size_t)i * kMaxStageGates + q];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w];
    grad[entries[k].gate] = 2.0 * tot;
  }
}

// ---------------------------------------------------------------------------
// Host orchestration.

static void perm(const Space& s, const double* in, double* out, const int32_t* src, hipStream_t st) {
  SQ_PROF(PC_PERM, st, 16.0 * (double)s.dim, 0.0,
          k_perm_rows<<<s.nsa, 256, sizeof(double) * s.nsb, st>>>(in, out, src, s.nsb));
}

static double stage_bytes(const Space& s, int vectors) { return 16.0 * (double)s.dim * vectors; }

// Forward sweep (dagger: U^+, gates reversed with -theta) over the whole
// program with tiles where possible.  psi holds the canonical-order input and
// receives the output; scratch >= dim doubles; gwork >= 2*dim doubles for
// gates that run through the per-gate kernels.  deriv_gate >= 0 inserts T_g
// right after gate g (plain gradient; forward only).
void tiles_forward(const Space& s, const Program& p, const TilePlan& T, const double* theta, double* psi,
                   double* scratch, double* gwork, int deriv_gate, bool dagger, hipStream_t st) {
  double* bufs[2] = {psi, scratch};
  int cur = 0;     // buffer holding the state
  int order = -1;  // -1 canonical, else the tiled stage whose column order the state is in
  const int ns = (int)T.stages.size();
  for (int ii = 0; ii < ns; ++ii) {
    const int i = dagger ? ns - 1 - ii : ii;
    const Stage& S = T.stages[i];
    if (!S.tiled) {
      if (order >= 0) {
        perm(s, bufs[cur], bufs[1 - cur], T.stages[order].d_to_canon, st);
        cur = 1 - cur;
        order = -1;
      }
      apply_gate_impl(s, p, S.g0, theta, dagger, bufs[cur], gwork, st);
      if (S.g0 == deriv_gate) {
        grad_action_impl(s, p, S.g0, bufs[cur], bufs[1 - cur], st);
        cur = 1 - cur;
      }
      continue;
    }
    const int32_t* src = nullptr;
    if (order < 0) {
      src = S.d_from_canon;
    } else if (!dagger && S.prev_tiled == order) {
      src = S.d_from_prev;
    } else if (dagger && T.stages[order].prev_tiled == i) {
      src = T.stages[order].d_to_prev;
    } else {
      perm(s, bufs[cur], bufs[1 - cur], T.stages[order].d_to_canon, st);
      cur = 1 - cur;
      src = S.d_from_canon;
    }
    perm(s, bufs[cur], bufs[1 - cur], src, st);
    cur = 1 - cur;
    order = i;
    const int dg = (deriv_gate >= S.g0 && deriv_gate < S.g1) ? deriv_gate : -1;
    SQ_PROF(PC_TILE_FWD, st, stage_bytes(s, 1), 0.0,
            k_stage_fwd<<<S.dev.ntiles, kTB, 0, st>>>(S.dev, theta, dagger ? 1 : 0, bufs[cur], dg));
  }
  if (order >= 0) {
    perm(s, bufs[cur], bufs[1 - cur], T.stages[order].d_to_canon, st);
    cur = 1 - cur;
  }
  if (cur != 0) copy_vec(bufs[cur], psi, s.dim, st);
}

// Reverse sweep on (psi, lam) with gradient dots into grad_dev[g] (device).
// sp/sl: scratch buffers (>= dim each); gwork >= 2*dim; partial: reduction scratch.
void tiles_reverse(const Space& s, const Program& p, const TilePlan& T, const double* theta, double* psi, double* lam,
                   double* sp_, double* sl_, double* gwork, double* partial, double* grad_dev, hipStream_t st) {
  double* pb[2] = {psi, sp_};
  double* lb[2] = {lam, sl_};
  int cur = 0;
  int order = -1;
  for (int i = (int)T.stages.size() - 1; i >= 0; --i) {
    const Stage& S = T.stages[i];
    if (!S.tiled) {
      if (order >= 0) {
        perm(s, pb[cur], pb[1 - cur], T.stages[order].d_to_canon, st);
        perm(s, lb[cur], lb[1 - cur], T.stages[order].d_to_canon, st);
        cur = 1 - cur;
        order = -1;
      }
      reverse_step_impl(s, p, S.g0, theta, pb[cur], lb[cur], grad_dev, gwork, partial, st);
      continue;
    }
    // the state is in canonical order or in the order of stage i+1
    const int32_t* src = nullptr;
    if (order < 0) {
      src = S.d_from_canon;
    } else if (T.stages[order].prev_tiled == i) {
      src = T.stages[order].d_to_prev;
    } else {
      perm(s, pb[cur], pb[1 - cur], T.stages[order].d_to_canon, st);
      perm(s, lb[cur], lb[1 - cur], T.stages[order].d_to_canon, st);
      cur = 1 - cur;
      src = S.d_from_canon;
    }
    perm(s, pb[cur], pb[1 - cur], src, st);
    perm(s, lb[cur], lb[1 - cur], src, st);
    cur = 1 - cur;
    order = i;
    const int wb = i > 0 ? 1 : 0;
    SQ_PROF(PC_TILE_REV, st, stage_bytes(s, wb ? 2 : 1), 0.0,
            k_stage_rev<<<S.dev.ntiles, kTB, 0, st>>>(S.dev, theta, pb[cur], lb[cur], T.d_partial, wb));
    SQ_PROF(PC_REDUCE, st, 8.0 * S.dev.ntiles * S.dev.nentries, 0.0,
            k_stage_reduce<<<S.dev.nentries, kTB, 0, st>>>(T.d_partial, S.dev.ntiles, S.dev.entries, S.dev.nentries,
                                                           grad_dev));
  }
}

}  // namespace sq
