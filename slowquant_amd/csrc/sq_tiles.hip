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
// x {beta bit of j} (other indices fixed), and their whole product is one
// small matrix per orbit type; a thread owns an orbit and applies it.  Signs
// come from the pair-local bits only: the environment parity cancels for
// number-conserving spin parts.  One barrier per pair.
//
// Per rotation string and orbit type (alpha part of the pair empty / doubly
// occupied / single, same for beta) the host precomputes a 12-bit code: for
// each of the 6 member pairs {0,1},{2,3},{0,2},{1,3},{0,3},{1,2} whether the
// string rotates it and with which sign.
//
// Kernels:
//   k_stage_mats    per stage, pair and orbit type: the product of the pair's
//                   gates (optionally with T_g inserted) from theta
//   k_stage_fwd     one stage on one vector; dagger runs U^+
//   k_stage_rev     reverse stage on (psi, lambda): antisymmetric orbit
//                   products, then un-apply the pair matrices
//   k_stage_grad    grad[g] from the orbit products summed over tiles

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <numeric>

#include "sq_internal.h"

namespace sq {

constexpr int kMaxPairs = 6;        // hypercubes of at most 64 strings per spin
constexpr int kTileMax = 4096;      // doubles per vector per tile
constexpr int kTileMaxCols = 1024;  // columns per tile
constexpr int kDenseCols = 256;     // columns per tile of a split stage's main launch
constexpr int kTB = 256;            // threads per block of the reductions
constexpr int kNT = 512;            // threads per block of the tile kernels
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

// Per pair of a stage, the gates act on orbits of <= 4 amplitudes (member
// o = alpha bit | beta bit << 1).  Their combined action is one matrix per
// orbit type, stored compactly in kPairVals doubles:
//   [0..15]  type (alpha single, beta single): 4x4, row-major M[a][b]
//   [16..19] type (alpha single, beta empty):  2x2 on members {0, 1}
//   [20..23] type (alpha single, beta double): 2x2 on members {0, 1}
//   [24..27] type (alpha empty, beta single):  2x2 on members {0, 2}
//   [28..31] type (alpha double, beta single): 2x2 on members {0, 2}
// (types without a single pair are left alone by every gate).  A gate's
// gradient is sum over orbits of <lam|W T W^T|psi> with W the product of the
// pair's later gates; W T W^T is antisymmetric, so the reverse sweep only
// accumulates Q[a<b] = sum (lam_a psi_b - lam_b psi_a):
//   Q[0..5]  type (single, single): pairs (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)
//   Q[6..9]  the four two-member types, (member 0, second member)
constexpr int kPairVals = 32;
constexpr int kQVals = 10;
// Pair matrices carry one more value, [32]: the factor for orbits of types
// without a single pair (1, or 0 when a derivative T_g is inserted on the
// pair: T_g annihilates them); padded to 36.
constexpr int kMatVals = 36;

// Per pair j of a stage, a 5-bit field (bits 5j..5j+4) of a string's
// descriptor: kind (2 bits: 0 pair empty, 1 pair doubly occupied, 2 singly
// occupied with hypercube bit 0, 3 singly occupied with hypercube bit 1) and
// the hypercube bit index (3 bits).
// One dispatch slot: alpha rows arows[r0, r0+R), beta stage positions
// [c0, c1), the alpha group's pair descriptor, the tile's column run and that
// run's table offset, and the tile's index among the stage's tiles (its
// gradient partials).  R = 0 marks an empty slot.
struct TileRec {
  int32_t r0, R, c0, c1;
  uint32_t adesc;
  int32_t run, toff, tile;
};

// Orbit column lists of one run of beta stage positions [c0, c1), shared by
// every tile over that run.  Per pair j the run's table holds W = c1 - c0
// 16-bit entries at ctab[toff + j*W]: first the columns whose beta pair j is
// singly occupied with hypercube bit 0 (cnt[0][j] of them: 4-member orbits,
// or 2-member beta orbits when the alpha pair is not single), then the
// columns with beta pair j empty (cnt[1][j]), then doubly occupied
// (cnt[2][j]).  inv[t][j] = 1 / max(cnt[t][j], 1) and inv[3][j] =
// 1 / max(cnt[1][j] + cnt[2][j], 1) (float) split work items into (row,
// list entry) without a division in the kernels.  Entry = column within the run (10 bits) | hypercube bit of
// the pair << 10.  Columns holding the upper member of a single beta pair are
// reached from their lower member and listed nowhere.  The counts are 32-bit
// so that the tile kernels read them with uniform (scalar) loads.
struct ColRun {
  int32_t c0, c1, toff, pad;
  int32_t cnt[3][8];
  float inv[4][8];
};

struct StageDev {
  int npairs;
  int nsb;
  int nt;                    // tiles
  int nb;                    // dispatch slots (the tiles in XCD-aware order, and empty slots)
  int64_t poff;              // this stage's block of the gradient partials
  const double* mats;        // this stage's pair matrices [npairs][kMatVals]
  const int32_t* arows;      // concatenated alpha rows (canonical alpha index), groups sorted by hypercube index
  const int32_t* bcol;       // per beta stage position: canonical beta column
  const struct TileRec* tiles;  // [nb]: slot b's tile
  const ColRun* runs;        // column runs of the stage's tiles
  const uint16_t* ctab;      // their orbit column tables (runs start 16-byte aligned)
};

// Dynamic LDS of one tile, carved in this order (16-byte aligned parts):
// X[E] (| L[E] in the reverse sweep) | orbit column table [npairs][W]
// (16-bit) | canonical columns [W] | alpha rows [R].
__host__ __device__ constexpr size_t align16(size_t b) { return (b + 15) / 16 * 16; }
__host__ __device__ constexpr size_t tab_entries(int npairs, int W) { return ((size_t)npairs * W + 7) / 8 * 8; }
__host__ __device__ constexpr size_t tile_lds(int E, int W, int R, int npairs, int nvec) {
  return align16((size_t)E * 8 * nvec) + align16(tab_entries(npairs, W) * 2) + align16((size_t)W * 4) + (size_t)R * 4;
}

// Per tiled stage, what the matrix and gradient kernels need.
struct StageInfo {
  int npairs, nentries, ntiles, pad;
  int pe[kMaxPairs + 1];     // entries of pair j: [pe[j], pe[j+1]) (entries sorted by pair)
  int64_t poff;              // gradient partials: [npairs][kQVals][ntiles]
  const StageEntry* entries;
};

struct Stage {
  bool tiled = false;
  int g0 = 0, g1 = 0;  // program gates [g0, g1)
  int npairs = 0;
  int pp[kMaxPairs];
  std::vector<StageEntry> entries;
  int pe[kMaxPairs + 1];
  int ntiles = 0;
  int index = -1;              // position among the tiled stages
  StageDev dev{};              // launch descriptor
  size_t lds_fwd = 0, lds_rev = 0;  // dynamic LDS of the forward / reverse launches (largest tile)
  StageDev dev_big{};          // tiles above the target when they run as a side launch (nb = 0: none)
  size_t lds_fwd_big = 0, lds_rev_big = 0;
  const StageEntry* d_entries = nullptr;
  std::vector<void*> allocs;
};

struct TilePlan {
  std::vector<Stage> stages;
  int ntiled = 0;
  int nt = kNT;  // threads per block of the stage kernels (256 or 512)
  double* d_partial = nullptr;  // gradient partials of all tiled stages
  int64_t partial_len = 0;
  double* d_mats = nullptr;     // [ntiled][kMaxPairs][kMatVals]
  StageInfo* d_info = nullptr;  // [ntiled]
  hipStream_t side = nullptr;   // side launches of large tiles (default priority, non-blocking)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  ~TilePlan() {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    for (auto& s : stages)
      for (void* p : s.allocs) (void)hipFree(p);
    if (d_partial) (void)hipFree(d_partial);
    if (d_mats) (void)hipFree(d_mats);
    if (d_info) (void)hipFree(d_info);
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

// Elements per tile that runs of beta groups are filled up to (SQ_TILE_TARGET
// overrides); a single group pair above it still makes one tile.
static int tile_target(int64_t dim, int program_target) {
  if (program_target > 0) return std::min(std::max(program_target, 16), kTileMax);
  const char* env = getenv("SQ_TILE_TARGET");
  if (env) return std::min(std::max(atoi(env), 16), kTileMax);
  // 256..2048 elements, about 400 tiles or more per stage (measured fastest at
  // (12e,12o) and (14e,14o); the autotuner also tries 4096)
  int t = 256;
  while (t < 2048 && (int64_t)t * 2 * 384 <= dim) t *= 2;
  return t;
}

// 5-bit pair fields of a string given its local occupancies (2 bits per pair)
// and hypercube bit indices (4 bits per pair, 0xF if the pair is not single).
static uint32_t pair_desc(uint32_t occ, uint32_t bits, int npairs) {
  uint32_t d = 0;
  for (int j = 0; j < npairs; ++j) {
    const uint32_t loc = (occ >> (2 * j)) & 3u;
    const uint32_t b = (bits >> (4 * j)) & 0xFu;
    uint32_t f;
    if (b == 0xFu) f = loc == 0 ? 0u : 1u;
    else f = (loc == 2u ? 3u : 2u) | (b << 2);
    d |= f << (5 * j);
  }
  return d;
}

static void finalize_stage(const Space& sp, Stage& S, int program_target) {
  // entries grouped by pair, program order kept within a pair
  std::stable_sort(S.entries.begin(), S.entries.end(),
                   [](const StageEntry& a, const StageEntry& b) { return a.pair < b.pair; });
  int* pe = S.pe;
  for (int j = 0, k = 0; j <= kMaxPairs; ++j) {
    while (k < (int)S.entries.size() && S.entries[k].pair < j) ++k;
    pe[j] = k;
  }
  GroupInfo A = group_strings(sp.h_astr, S.npairs, S.pp);
  GroupInfo B = group_strings(sp.h_bstr, S.npairs, S.pp);
  // tiles: each alpha group x runs of whole beta groups, up to the target
  // size and maxcols columns; a beta group that alone exceeds the target gets
  // its own tile
  const int target = tile_target(sp.dim, program_target);
  const int nag = (int)A.goff.size() - 1;
  const int nbg = (int)B.goff.size() - 1;
  auto build_tiles = [&](int maxcols) {
    std::vector<int4> out;
    for (int ga = 0; ga < nag; ++ga) {
      const int R = A.goff[ga + 1] - A.goff[ga];
      SQ_CHECK(R <= 64, SQ_ERR_ARG, "tile does not fit: alpha hypercube too large");
      int gb = 0;
      while (gb < nbg) {
        const int c0 = B.goff[gb];
        int c1 = B.goff[gb + 1];
        ++gb;
        while (gb < nbg && (int64_t)R * (B.goff[gb + 1] - c0) <= target && B.goff[gb + 1] - c0 <= maxcols) {
          c1 = B.goff[gb + 1];
          ++gb;
        }
        const int64_t E = (int64_t)R * (c1 - c0);
        SQ_CHECK(E <= kTileMax && c1 - c0 <= kTileMaxCols, SQ_ERR_ARG, "tile does not fit: hypercube groups too large");
        out.push_back(int4{ga, c0, c1, (int)E});
      }
    }
    return out;
  };
  std::vector<int4> tiles = build_tiles(kTileMaxCols);
  // Occupancy.  Every block of a launch reserves the LDS of the launch's
  // largest tile.  When the stage has more tiles than the CUs hold at once at
  // that footprint, the stage runs as two launches side by side: the tiles
  // above the target (alpha and beta groups that are both large hypercubes,
  // up to twice the target) in a side launch, the others in the main launch,
  // rebuilt with at most kDenseCols columns so that their column tables stay
  // small too (SQ_TILE_SPLIT=0 keeps one launch of the tiles above).
  bool split = false;
  {
    size_t lmax = 0;
    for (const int4& t : tiles)
      lmax = std::max(lmax, tile_lds(t.w, t.z - t.y, A.goff[t.x + 1] - A.goff[t.x], S.npairs, 2));
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t resident = (int64_t)std::max(1, cus) * std::max<int64_t>(1, (int64_t)(160 * 1024) / (int64_t)lmax);
    const char* senv = getenv("SQ_TILE_SPLIT");
    split = !(senv && senv[0] == '0') && (int64_t)tiles.size() > resident;
    if (split) tiles = build_tiles(kDenseCols);
  }
  // largest tiles first: they start in the first dispatch round
  std::stable_sort(tiles.begin(), tiles.end(), [](const int4& a, const int4& b) { return a.w > b.w; });
  S.ntiles = (int)tiles.size();
  std::vector<uint32_t> adesc(nag), bdesc(sp.nsb);
  for (int ga = 0; ga < nag; ++ga) {
    // kind 2 (single, bit 0) for single pairs; the row's own bit is its hypercube index
    uint32_t d = 0;
    for (int j = 0; j < S.npairs; ++j) {
      const uint32_t b = (A.gbit[ga] >> (4 * j)) & 0xFu;
      const uint32_t loc = (A.gocc[ga] >> (2 * j)) & 3u;
      const uint32_t f = b == 0xFu ? (loc == 0 ? 0u : 1u) : (2u | (b << 2));
      d |= f << (5 * j);
    }
    adesc[ga] = d;
  }
  for (int k = 0; k < sp.nsb; ++k) bdesc[k] = pair_desc(B.occ[k], B.bit[k], S.npairs);
  // column runs (distinct [c0, c1)) and their orbit column tables
  std::map<std::pair<int, int>, int> run_of;
  std::vector<ColRun> runs;
  std::vector<uint16_t> ctab;
  auto run_index = [&](int c0, int c1) {
    auto it = run_of.find({c0, c1});
    if (it != run_of.end()) return it->second;
    ColRun r{};
    r.c0 = c0;
    r.c1 = c1;
    r.toff = (int)ctab.size();
    const int W = c1 - c0;
    for (int j = 0; j < S.npairs; ++j) {
      std::vector<uint16_t> lists[3];
      for (int k = c0; k < c1; ++k) {
        const uint32_t f = (bdesc[k] >> (5 * j)) & 31u;
        const uint32_t kind = f & 3u;
        const uint16_t e = (uint16_t)((uint32_t)(k - c0) | ((f >> 2) << 10));
        if (kind == 2u) lists[0].push_back(e);
        else if (kind == 0u) lists[1].push_back(e);
        else if (kind == 1u) lists[2].push_back(e);
      }
      for (int t = 0; t < 3; ++t) {
        r.cnt[t][j] = (int32_t)lists[t].size();
        r.inv[t][j] = 1.0f / (float)std::max<int32_t>(r.cnt[t][j], 1);
        ctab.insert(ctab.end(), lists[t].begin(), lists[t].end());
      }
      r.inv[3][j] = 1.0f / (float)std::max<int32_t>(r.cnt[1][j] + r.cnt[2][j], 1);
      ctab.resize((size_t)r.toff + (size_t)(j + 1) * W, 0);
    }
    ctab.resize((ctab.size() + 7) / 8 * 8, 0);  // next run 16-byte aligned
    runs.push_back(r);
    run_of[{c0, c1}] = (int)runs.size() - 1;
    return (int)runs.size() - 1;
  };
  std::vector<TileRec> recs(tiles.size());
  for (size_t t = 0; t < tiles.size(); ++t) {
    const int ga = tiles[t].x;
    const int run = run_index(tiles[t].y, tiles[t].z);
    recs[t] = TileRec{A.goff[ga], A.goff[ga + 1] - A.goff[ga], tiles[t].y, tiles[t].z, adesc[ga], run,
                      runs[run].toff, (int32_t)t};
  }
  StageDev d{};
  d.npairs = S.npairs;
  d.nsb = sp.nsb;
  d.arows = upload(S.allocs, A.order);
  d.bcol = upload(S.allocs, B.order);
  d.runs = upload(S.allocs, runs);
  d.ctab = upload(S.allocs, ctab);
  S.d_entries = upload(S.allocs, S.entries);
  d.nt = S.ntiles;
  S.dev = d;
  auto lds_of = [&](const TileRec& r, int nvec) {
    const int W = r.c1 - r.c0;
    return tile_lds(r.R * W, W, r.R, S.npairs, nvec);
  };
  // the tiles of the side launch (split stages only)
  std::vector<char> big(tiles.size(), 0);
  {
    int nbig = 0;
    if (split)
      for (size_t t = 0; t < tiles.size(); ++t) nbig += (big[t] = tiles[t].w > target) ? 1 : 0;
    if (nbig == (int)tiles.size()) std::fill(big.begin(), big.end(), 0);
  }
  // Dispatch order of one launch.  Workgroups go to the XCDs round robin
  // (slot b on XCD b % 8), each XCD with its own L2.  Tiles of one alpha
  // group read the same rows of the state, so whole groups are given to one
  // XCD (groups largest first, each to the XCD with the fewest elements so
  // far) and run there back to back; SQ_TILE_XCD=0 keeps the plain
  // largest-first order.  The slots hold the tile records themselves (empty
  // slots: R = 0).
  auto make_launch = [&](bool want_big, StageDev& out, size_t& lds_fwd, size_t& lds_rev) {
    constexpr int kXcds = 8;
    std::vector<int32_t> mine;
    for (size_t t = 0; t < tiles.size(); ++t)
      if ((big[t] != 0) == want_big) mine.push_back((int32_t)t);
    out = d;
    out.nb = 0;
    out.tiles = nullptr;
    lds_fwd = lds_rev = 0;
    if (mine.empty()) return;
    std::vector<int32_t> order;
    const char* xenv = getenv("SQ_TILE_XCD");
    if (xenv && xenv[0] == '0') {
      order = mine;
    } else {
      std::vector<std::vector<int32_t>> of_group(nag);
      std::vector<int64_t> gsize(nag, 0);
      for (int32_t t : mine) {
        of_group[tiles[t].x].push_back(t);
        gsize[tiles[t].x] += tiles[t].w;
      }
      std::vector<int> gorder(nag);
      std::iota(gorder.begin(), gorder.end(), 0);
      std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int b) { return gsize[a] > gsize[b]; });
      std::vector<std::vector<int32_t>> bins(kXcds);
      std::vector<int64_t> load(kXcds, 0);
      for (int ga : gorder) {
        if (of_group[ga].empty()) continue;
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += gsize[ga];
        bins[x].insert(bins[x].end(), of_group[ga].begin(), of_group[ga].end());
      }
      size_t len = 0;
      for (auto& b : bins) len = std::max(len, b.size());
      order.assign(len * kXcds, -1);
      for (int x = 0; x < kXcds; ++x)
        for (size_t k = 0; k < bins[x].size(); ++k) order[k * kXcds + x] = bins[x][k];
    }
    std::vector<TileRec> slots(order.size(), TileRec{0, 0, 0, 0, 0u, 0, 0, -1});
    for (size_t b = 0; b < order.size(); ++b)
      if (order[b] >= 0) slots[b] = recs[order[b]];
    out.nb = (int)slots.size();
    out.tiles = upload(S.allocs, slots);
    for (int32_t t : mine) {
      lds_fwd = std::max(lds_fwd, lds_of(recs[t], 1));
      lds_rev = std::max(lds_rev, lds_of(recs[t], 2));
    }
  };
  make_launch(false, S.dev, S.lds_fwd, S.lds_rev);
  make_launch(true, S.dev_big, S.lds_fwd_big, S.lds_rev_big);
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
    std::vector<StageInfo> info;
    for (auto& S : T->stages) {
      if (!S.tiled) continue;
      finalize_stage(sp, S, P.tile_target);
      S.index = (int)info.size();
      StageInfo I{};
      I.npairs = S.npairs;
      I.nentries = (int)S.entries.size();
      I.ntiles = S.ntiles;
      for (int j = 0; j <= kMaxPairs; ++j) I.pe[j] = S.pe[j];
      I.poff = off;
      I.entries = S.d_entries;
      info.push_back(I);
      S.dev.poff = off;
      S.dev_big.poff = off;
      off += (int64_t)S.ntiles * S.npairs * kQVals;
    }
    T->ntiled = (int)info.size();
    T->nt = P.tile_threads == 256 ? 256 : kNT;
    SQ_HIP(hipMalloc(&T->d_partial, sizeof(double) * std::max<int64_t>(off, 1)));
    T->partial_len = off;
    SQ_HIP(hipMalloc(&T->d_mats, sizeof(double) * std::max<int64_t>((int64_t)T->ntiled * kMaxPairs * kMatVals, 1)));
    for (auto& S : T->stages) {
      if (!S.tiled) continue;
      S.dev.mats = T->d_mats + (size_t)S.index * kMaxPairs * kMatVals;
      S.dev_big.mats = S.dev.mats;
    }
    std::vector<void*> tmp;
    T->d_info = upload(tmp, info);
    bool any_big = false;
    for (auto& S : T->stages) any_big = any_big || (S.tiled && S.dev_big.nb > 0);
    if (any_big) {
      SQ_HIP(hipStreamCreateWithFlags(&T->side, hipStreamNonBlocking));
      SQ_HIP(hipEventCreateWithFlags(&T->ev_fork, hipEventDisableTiming));
      SQ_HIP(hipEventCreateWithFlags(&T->ev_join, hipEventDisableTiming));
    }
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
//
// k_stage_mats   per tiled stage and pair, the product of the pair's gates
//                (optionally with T_g inserted after gate g) per orbit type
// k_stage_fwd    one stage on one vector: per pair, orbit matrix-vector
//                products (transposed for U^+)
// k_stage_rev    reverse stage on (psi, lam): per pair, accumulate the orbit
//                outer products, then un-apply the pair matrix on both
// k_stage_grad   per stage and pair, sum the outer products over tiles and
//                contract them with each gate's conjugated generator
//
// Tile kernels run kNT threads per block, one block per tile, one launch per
// stage; tiles are sorted largest first so that they start in the first
// dispatch round.  LDS is dynamic and carved per tile (tile_lds); the
// launch requests the largest tile's footprint.  Pair matrices are read from
// global memory with uniform (scalar) loads.

// ---- 4x4 orbit algebra (matrix kernels) ------------------------------------

// R M for the rotation of one entry code (disjoint member planes).
__device__ __forceinline__ void rot_rows(uint32_t code, double c, double s, double (&M)[16]) {
#pragma unroll
  for (int sl = 0; sl < 6; ++sl) {
    const uint32_t f = (code >> (2 * sl)) & 3u;
    if (!(f & 1u)) continue;
    const int a = kSlotA[sl], b = kSlotB[sl];
    const double sg = (f & 2u) ? -s : s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double x = M[a * 4 + q], y = M[b * 4 + q];
      M[a * 4 + q] = c * x - sg * y;
      M[b * 4 + q] = sg * x + c * y;
    }
  }
}

// M R^T (columns), the other half of a conjugation.
__device__ __forceinline__ void rot_cols(uint32_t code, double c, double s, double (&M)[16]) {
#pragma unroll
  for (int sl = 0; sl < 6; ++sl) {
    const uint32_t f = (code >> (2 * sl)) & 3u;
    if (!(f & 1u)) continue;
    const int a = kSlotA[sl], b = kSlotB[sl];
    const double sg = (f & 2u) ? -s : s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double x = M[q * 4 + a], y = M[q * 4 + b];
      M[q * 4 + a] = c * x - sg * y;
      M[q * 4 + b] = sg * x + c * y;
    }
  }
}

// out += scale * T M (rows), T|a> = sg|b>, T|b> = -sg|a> per active plane.
__device__ __forceinline__ void t_rows(uint32_t code, double scale, const double (&M)[16], double (&out)[16]) {
#pragma unroll
  for (int sl = 0; sl < 6; ++sl) {
    const uint32_t f = (code >> (2 * sl)) & 3u;
    if (!(f & 1u)) continue;
    const int a = kSlotA[sl], b = kSlotB[sl];
    const double sg = (f & 2u) ? -scale : scale;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      out[b * 4 + q] += sg * M[a * 4 + q];
      out[a * 4 + q] -= sg * M[b * 4 + q];
    }
  }
}

// Orbit-type classes of the compact layout: type index (alpha kind * 3 +
// beta kind), offset, second member.
__device__ __constant__ int c_cls_type[5] = {8, 6, 7, 2, 5};
__device__ __constant__ int c_cls_off[5] = {0, 16, 20, 24, 28};

// Stage entries [e0, e0 + n) staged in LDS with their rotation cosine/sine
// (one thread per entry), so the products below run on LDS values instead of
// chains of dependent global loads.
struct EntryLds {
  int gate[kMaxStageGates];
  double scale[kMaxStageGates], c[kMaxStageGates], s[kMaxStageGates];
  uint32_t code[kMaxStageGates][kTypes];
};

__device__ __forceinline__ void stage_entries(const StageEntry* __restrict__ entries, int e0, int n,
                                              const double* __restrict__ theta, EntryLds& E) {
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const StageEntry& en = entries[e0 + e];
    const double th = theta[en.gate];
    double s = 0.0, c = 1.0;
    if (fabs(th) >= 1e-28) sincos(en.scale * th, &s, &c);
    E.gate[e] = en.gate;
    E.scale[e] = en.scale;
    E.c[e] = c;
    E.s[e] = s;
#pragma unroll
    for (int t = 0; t < kTypes; ++t) E.code[e][t] = en.code[t];
  }
}

// One thread per (stage, pair, class): the pair's gate product in forward
// order, with T_g inserted after gate g when deriv_gate is on this pair.
__global__ __launch_bounds__(64) void k_stage_mats(const StageInfo* __restrict__ info, const double* __restrict__ theta, int deriv_gate,
                             double* __restrict__ mats, int64_t th_stride, int64_t mats_stride) {
  __shared__ EntryLds E;
  theta += blockIdx.y * th_stride;
  mats += blockIdx.y * mats_stride;
  const StageInfo* __restrict__ I = info + blockIdx.x;
  stage_entries(I->entries, 0, I->nentries, theta, E);
  __syncthreads();
  const int j = threadIdx.x / 5, cls = threadIdx.x % 5;
  if (j >= I->npairs) return;
  const int t = c_cls_type[cls];
  double M[16];
  #pragma unroll
  for (int q = 0; q < 16; ++q) M[q] = (q % 5 == 0) ? 1.0 : 0.0;
  const int k0 = I->pe[j], k1 = I->pe[j + 1];
  bool has_deriv = false;
  for (int k = k0; k < k1; ++k) {
    rot_rows(E.code[k][t], E.c[k], E.s[k], M);
    if (E.gate[k] == deriv_gate && (k + 1 == k1 || E.gate[k + 1] != deriv_gate)) {
      has_deriv = true;
      double o[16];
      #pragma unroll
      for (int q = 0; q < 16; ++q) o[q] = 0.0;
      for (int q = k; q >= k0 && E.gate[q] == deriv_gate; --q) t_rows(E.code[q][t], E.scale[q], M, o);
      #pragma unroll
      for (int q = 0; q < 16; ++q) M[q] = o[q];
    }
  }
  double* base = mats + ((size_t)blockIdx.x * kMaxPairs + j) * kMatVals;
  if (cls == 0) base[kPairVals] = has_deriv ? 0.0 : 1.0;
  double* out = base + c_cls_off[cls];
  if (cls == 0) {
#pragma unroll
    for (int q = 0; q < 16; ++q) out[q] = M[q];
  } else {
    const bool a = cls <= 2;  // second member 1 (alpha pair) or 2 (beta pair)
    out[0] = M[0];
    out[1] = a ? M[1] : M[2];
    out[2] = a ? M[4] : M[8];
    out[3] = a ? M[5] : M[10];
  }
}

// One block per (stage, pair): Q summed over the stage's tiles, then for every
// entry k of the pair dot_k = sum_{a<b} (W T_k W^T)[a][b] Q[a<b] per class, W
// the product of the pair's entries after k; grad[gate] = 2 sum over its
// entries of scale * dot.
constexpr int kQa[6] = {0, 0, 0, 1, 1, 2};
constexpr int kQb[6] = {1, 2, 3, 2, 3, 3};

__global__ __launch_bounds__(kTB) void k_stage_grad(const StageInfo* __restrict__ info, const double* __restrict__ theta,
                             const double* __restrict__ partial, double* __restrict__ grad, int64_t th_stride,
                             int64_t part_stride, int64_t grad_stride) {
  theta += blockIdx.y * th_stride;
  partial += blockIdx.y * part_stride;
  grad += blockIdx.y * grad_stride;
  __shared__ double Q[kQVals];
  __shared__ double dots[kMaxPairEntries * 5];
  __shared__ double red[kQVals * (kTB / 64)];
  __shared__ EntryLds E;
  const StageInfo I = info[blockIdx.x / kMaxPairs];
  const int j = blockIdx.x % kMaxPairs;
  if (j >= I.npairs) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k0 = I.pe[j], k1 = I.pe[j + 1];
  const int nent = k1 - k0;
  stage_entries(I.entries, k0, nent, theta, E);  // entries of this pair, indexed from 0
  {
    const double* src = partial + I.poff + (size_t)j * kQVals * I.ntiles;
    double acc[kQVals];
#pragma unroll
    for (int v = 0; v < kQVals; ++v) acc[v] = 0.0;
    for (int tile = threadIdx.x; tile < I.ntiles; tile += blockDim.x) {
#pragma unroll
      for (int v = 0; v < kQVals; ++v) acc[v] += src[(size_t)v * I.ntiles + tile];
    }
#pragma unroll
    for (int v = 0; v < kQVals; ++v) {
      double a = acc[v];
      for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
      if (lane == 0) red[v * (kTB / 64) + wave] = a;
    }
    __syncthreads();
    if (threadIdx.x < kQVals) {
      double tot = 0.0;
      for (int w = 0; w < kTB / 64; ++w) tot += red[threadIdx.x * (kTB / 64) + w];
      Q[threadIdx.x] = tot;
    }
    __syncthreads();
  }
  if (threadIdx.x < nent * 5) {
    const int k = threadIdx.x / 5, cls = threadIdx.x % 5;
    const int t = c_cls_type[cls];
    double M[16];
    #pragma unroll
    for (int q = 0; q < 16; ++q) M[q] = 0.0;
    double I4[16];
    #pragma unroll
    for (int q = 0; q < 16; ++q) I4[q] = (q % 5 == 0) ? 1.0 : 0.0;
    t_rows(E.code[k][t], 1.0, I4, M);  // M = T_k
    for (int e = k + 1; e < nent; ++e) {
      rot_rows(E.code[e][t], E.c[e], E.s[e], M);
      rot_cols(E.code[e][t], E.c[e], E.s[e], M);
    }
    double d = 0.0;
    if (cls == 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) d += M[kQa[q] * 4 + kQb[q]] * Q[q];
    } else {
      d = (cls <= 2 ? M[1] : M[2]) * Q[5 + cls];  // M[0][m1], m1 = 1 (alpha pair) or 2 (beta pair)
    }
    dots[threadIdx.x] = d * E.scale[k];
  }
  __syncthreads();
  // one thread per gate: the first entry of each run of equal gates sums the run
  if (threadIdx.x < nent && (threadIdx.x == 0 || E.gate[threadIdx.x - 1] != E.gate[threadIdx.x])) {
    const int g = E.gate[threadIdx.x];
    double tot = 0.0;
    for (int q = threadIdx.x; q < nent && E.gate[q] == g; ++q)
      for (int cls = 0; cls < 5; ++cls) tot += dots[q * 5 + cls];
    grad[g] = 2.0 * tot;
  }
}

// ---- tile kernels -----------------------------------------------------------

struct TileView {
  int R, W, E, c0;
  float invW;
  uint32_t adesc;
  int tile;  // index among the stage's tiles
  const ColRun* run;
};

struct TileLds {
  double* X;
  double* L;
  uint16_t* tab;  // the run's orbit column table, [npairs][W]
  uint32_t* ccol;
  int32_t* rows;
};

// LDS parts of one tile (layout of tile_lds).
__device__ __forceinline__ TileLds carve_lds(double* smem, int E, int W, int npairs, bool rev) {
  TileLds m;
  char* p = reinterpret_cast<char*>(smem);
  m.X = smem;
  m.L = rev ? m.X + E : nullptr;
  p += align16((size_t)E * 8 * (rev ? 2 : 1));
  m.tab = reinterpret_cast<uint16_t*>(p);
  p += align16(tab_entries(npairs, W) * 2);
  m.ccol = reinterpret_cast<uint32_t*>(p);
  p += align16((size_t)W * 4);
  m.rows = reinterpret_cast<int32_t*>(p);
  return m;
}

// Tile descriptors into LDS: alpha rows, the columns' canonical indices and
// the run's orbit column table (16-byte copies).
template <int NT>
__device__ __forceinline__ TileView tile_setup(const StageDev& S, double* smem, bool rev, const TileRec& T,
                                               TileLds& m) {
  TileView t;
  t.R = T.R;
  t.W = T.c1 - T.c0;
  t.E = t.R * t.W;
  t.c0 = T.c0;
  t.invW = 1.0f / (float)t.W;
  t.adesc = T.adesc;
  t.tile = T.tile;
  t.run = S.runs + T.run;
  m = carve_lds(smem, t.E, t.W, S.npairs, rev);
  for (int r = threadIdx.x; r < t.R; r += NT) m.rows[r] = S.arows[T.r0 + r];
  for (int c = threadIdx.x; c < t.W; c += NT) m.ccol[c] = (uint32_t)S.bcol[T.c0 + c];
  const uint4* src = reinterpret_cast<const uint4*>(S.ctab + T.toff);
  uint4* dst = reinterpret_cast<uint4*>(m.tab);
  const int n16 = (int)(tab_entries(S.npairs, t.W) / 8);
  for (int i = threadIdx.x; i < n16; i += NT) dst[i] = src[i];
  __syncthreads();
  return t;
}

// Global <-> LDS copies of a tile for NV vectors (psi | lam), gathered
// through the row and column tables, eight loads per vector in flight.
template <int NT, int NV>
__device__ __forceinline__ void tile_load(const StageDev& S, const TileView& t, const TileLds& m,
                                          const double* __restrict__ g0, const double* __restrict__ g1) {
  constexpr int U = 8;
  for (int e0 = threadIdx.x; e0 < t.E; e0 += NT * U) {
    double v[NV][U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * NT;
      if (e < t.E) {
        const int r = (int)(((float)e + 0.5f) * t.invW);
        const size_t idx = (size_t)m.rows[r] * S.nsb + m.ccol[e - r * t.W];
        v[0][u] = g0[idx];
        if (NV > 1) v[NV - 1][u] = g1[idx];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * NT;
      if (e < t.E) {
        m.X[e] = v[0][u];
        if (NV > 1) m.L[e] = v[NV - 1][u];
      }
    }
  }
}

template <int NT, int NV>
__device__ __forceinline__ void tile_store(const StageDev& S, const TileView& t, const TileLds& m, double* g0,
                                           double* g1) {
  for (int e = threadIdx.x; e < t.E; e += NT) {
    const int r = (int)(((float)e + 0.5f) * t.invW);
    const size_t idx = (size_t)m.rows[r] * S.nsb + m.ccol[e - r * t.W];
    g0[idx] = m.X[e];
    if (NV > 1) g1[idx] = m.L[e];
  }
}

// Work item i over (rows) x (n list entries): row rr and list index k (inv = 1/n).
__device__ __forceinline__ void split_item(int i, int n, float inv, int& rr, int& k) {
  rr = (int)(((float)i + 0.5f) * inv);
  k = i - rr * n;
}

// Row of the lower alpha member of hypercube row rr: a zero bit inserted at abit.
__device__ __forceinline__ int ins0(int rr, int abit) {
  return ((rr >> abit) << (abit + 1)) | (rr & ((1 << abit) - 1));
}

// o = M v (TR: M^T v), M row-major 4x4.
template <bool TR>
__device__ __forceinline__ void mat4(const double (&M)[16], const double (&v)[4], double (&o)[4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double acc = (TR ? M[a] : M[4 * a]) * v[0];
#pragma unroll
    for (int b = 1; b < 4; ++b) acc += (TR ? M[4 * b + a] : M[4 * a + b]) * v[b];
    o[a] = acc;
  }
}

// One pair pass of a forward stage on the tile in LDS.  Orbits come from the
// run's column lists, so every work item is one orbit and a wave runs one
// orbit type except where two lists meet.  Orbit members: 0 = (alpha lower,
// beta lower) at m0, 1 = alpha upper (+ dr), 2 = beta upper (+ dc), 3 = both.
// The pair's matrix values are read into registers before the orbit loops,
// and the 2-member orbits of beta-empty and beta-double columns run as
// separate item ranges, so every matrix value stays a wave-uniform (scalar)
// operand.
template <bool TR>
__device__ __forceinline__ void orbit2(double* X, int m0, int d, const double (&B)[4]) {
  const double v0 = X[m0], v1 = X[m0 + d];
  X[m0] = B[0] * v0 + (TR ? B[2] : B[1]) * v1;
  X[m0 + d] = (TR ? B[1] : B[2]) * v0 + B[3] * v1;
}

template <int NT, bool TR>
__device__ __forceinline__ void fwd_pair(const TileView& t, const TileLds& m, const double* __restrict__ Mp, int j) {
  const uint32_t fa = (t.adesc >> (5 * j)) & 31u;
  const int n4 = t.run->cnt[0][j], n2e = t.run->cnt[1][j], n2d = t.run->cnt[2][j];
  const uint16_t* tab = m.tab + j * t.W;
  double* X = m.X;
  const int W = t.W;
  if ((fa & 3u) >= 2u) {  // alpha pair single: 4-member orbits, then 2-member alpha orbits
    const int abit = (int)(fa >> 2);
    const int Rh = t.R >> 1, dr = W << abit;
    const int nA = Rh * n4, nE = nA + Rh * n2e, nT = nE + Rh * n2d;
    const float iA = t.run->inv[0][j], iE = t.run->inv[1][j], iD = t.run->inv[2][j];
    double M[16], be[4], bd[4];
#pragma unroll
    for (int q = 0; q < 16; ++q) M[q] = Mp[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      be[q] = Mp[16 + q];
      bd[q] = Mp[20 + q];
    }
    for (int i = threadIdx.x; i < nT; i += NT) {
      int rr, k;
      if (i < nA) {
        split_item(i, n4, iA, rr, k);
        const uint32_t e = tab[k];
        const int m0 = ins0(rr, abit) * W + (int)(e & 1023u);
        const int m2 = m0 + (1 << (e >> 10));
        const double v[4] = {X[m0], X[m0 + dr], X[m2], X[m2 + dr]};
        double o[4];
        mat4<TR>(M, v, o);
        X[m0] = o[0];
        X[m0 + dr] = o[1];
        X[m2] = o[2];
        X[m2 + dr] = o[3];
      } else if (i < nE) {  // beta pair empty
        split_item(i - nA, n2e, iE, rr, k);
        orbit2<TR>(X, ins0(rr, abit) * W + (int)(tab[n4 + k] & 1023u), dr, be);
      } else {  // beta pair doubly occupied
        split_item(i - nE, n2d, iD, rr, k);
        orbit2<TR>(X, ins0(rr, abit) * W + (int)(tab[n4 + n2e + k] & 1023u), dr, bd);
      }
    }
  } else {  // alpha pair empty or double: 2-member beta orbits; other amplitudes stay
    const double* Bp = Mp + ((fa & 3u) == 0u ? 24 : 28);
    const double B[4] = {Bp[0], Bp[1], Bp[2], Bp[3]};
    const bool deriv = Mp[kPairVals] == 0.0;
    const int nA = t.R * n4;
    const float iA = t.run->inv[0][j];
    for (int i = threadIdx.x; i < nA; i += NT) {
      int rr, k;
      split_item(i, n4, iA, rr, k);
      const uint32_t e = tab[k];
      orbit2<TR>(X, rr * W + (int)(e & 1023u), 1 << (e >> 10), B);
    }
    if (deriv) {  // derivative pair: T_g annihilates the amplitudes no gate moves
      const int n2 = n2e + n2d, nB = t.R * n2;
      const float iB = t.run->inv[3][j];
      for (int i = threadIdx.x; i < nB; i += NT) {
        int rr, k;
        split_item(i, n2, iB, rr, k);
        X[rr * W + (int)(tab[n4 + k] & 1023u)] = 0.0;
      }
    }
  }
}

// DPP lane moves of a double (bound_ctrl: lanes without a source read 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Row sums (16 lanes) of 8 per-lane values with halving exchanges inside
// each quad (xor 1, xor 2) and row shifts by 4 and 8: lane l with
// l % 16 >= 12 ends with the row sums of values
// q = (l & 1) * 4 + ((l >> 1) & 1) * 2 + {0, 1} in w[0], w[1].
// Every lane of the wave must be active.
__device__ __forceinline__ void row_fold8(const double (&Q)[8], double (&w)[2], int lane) {
  const bool o1 = lane & 1, o2 = lane & 2;
  double w1[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = o1 ? Q[4 + i] : Q[i], send = o1 ? Q[i] : Q[4 + i];
    w1[i] = keep + dpp_move<0xb1, 0xf>(send);  // quad_perm [1,0,3,2]
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = o2 ? w1[2 + i] : w1[i], send = o2 ? w1[i] : w1[2 + i];
    w[i] = keep + dpp_move<0x4e, 0xf>(send);  // quad_perm [2,3,0,1]
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    w[i] += dpp_move<0x114, 0xf>(w[i]);  // row_shr:4
    w[i] += dpp_move<0x118, 0xf>(w[i]);  // row_shr:8
  }
}

// Row sum of one value; lane l with l % 16 == 15 holds it.
__device__ __forceinline__ double row_sum(double v) {
  v += dpp_move<0xb1, 0xf>(v);
  v += dpp_move<0x4e, 0xf>(v);
  v += dpp_move<0x114, 0xf>(v);
  v += dpp_move<0x118, 0xf>(v);
  return v;
}

// Gates of one stage on one vector (canonical order, in place).
// dagger: U^+ (transposed pair matrices).  src may be dst (in place) or the
// sweep's input vector (first stage).
template <int NT>
__global__ __launch_bounds__(NT) void k_stage_fwd(StageDev S, const double* __restrict__ mats, int dagger,
                                                  const double* src, double* dst, int64_t mats_stride,
                                                  int64_t vec_stride) {
  extern __shared__ double smem[];
  mats += blockIdx.y * mats_stride;
  src += blockIdx.y * vec_stride;
  dst += blockIdx.y * vec_stride;
  // block b walks slots b, b + grid, ... (tiles are sorted largest first)
  for (int slot = blockIdx.x; slot < S.nb; slot += gridDim.x) {
    const TileRec T = S.tiles[slot];
    if (T.R == 0) continue;
    TileLds m;
    const TileView t = tile_setup<NT>(S, smem, false, T, m);
    tile_load<NT, 1>(S, t, m, src, nullptr);
    __syncthreads();
    for (int j = 0; j < S.npairs; ++j) {
      if (dagger) fwd_pair<NT, true>(t, m, mats + j * kMatVals, j);
      else fwd_pair<NT, false>(t, m, mats + j * kMatVals, j);
      __syncthreads();
    }
    tile_store<NT, 1>(S, t, m, dst, nullptr);
    __syncthreads();  // the next tile overwrites the LDS tables
  }
}

// One 2-member orbit of the reverse sweep at (m0, m0 + d): returns its
// product lam_0 psi_1 - lam_1 psi_0 and, with wb, un-applies the matrix (B^T)
// on both vectors.
__device__ __forceinline__ double rev_orbit2(double* X, double* L, int m0, int d, const double (&B)[4], bool wb) {
  const double v0 = X[m0], v1 = X[m0 + d], w0 = L[m0], w1 = L[m0 + d];
  if (wb) {
    X[m0] = B[0] * v0 + B[2] * v1;
    X[m0 + d] = B[1] * v0 + B[3] * v1;
    L[m0] = B[0] * w0 + B[2] * w1;
    L[m0 + d] = B[1] * w0 + B[3] * w1;
  }
  return w0 * v1 - w1 * v0;
}

// Reverse stage on (psi, lam), in place: per pair, the antisymmetric orbit
// products Q[a<b] += lam_a psi_b - lam_b psi_a before un-applying the pair
// matrix (M^T) on both; per-tile sums go to partial[poff + (j*10 + v)*ntiles + tile].
// Per pair the wave sums land in red[j][.][wave] (8 values when the alpha
// pair is single, else one: the beta orbit type of the alpha occupation).
template <int NT>
__global__ __launch_bounds__(NT) void k_stage_rev(StageDev S, const double* __restrict__ mats,
                                                  double* __restrict__ psi, double* __restrict__ lam,
                                                  double* __restrict__ partial, int write_back, int64_t mats_stride,
                                                  int64_t vec_stride, int64_t part_stride) {
  mats += blockIdx.y * mats_stride;
  psi += blockIdx.y * vec_stride;
  lam += blockIdx.y * vec_stride;
  partial += blockIdx.y * part_stride;
  constexpr int kWaves = NT / 64;
  extern __shared__ double smem[];
  __shared__ double red[kMaxPairs][8][kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool wb = write_back != 0;
  for (int slot = blockIdx.x; slot < S.nb; slot += gridDim.x) {
    const TileRec T = S.tiles[slot];
    if (T.R == 0) continue;
    TileLds m;
    const TileView t = tile_setup<NT>(S, smem, true, T, m);
    tile_load<NT, 2>(S, t, m, psi, lam);
    __syncthreads();
    const int W = t.W;
    double* X = m.X;
    double* L = m.L;
    for (int j = 0; j < S.npairs; ++j) {
      const double* __restrict__ Mp = mats + j * kMatVals;
      const uint32_t fa = (t.adesc >> (5 * j)) & 31u;
      const int n4 = t.run->cnt[0][j], n2e = t.run->cnt[1][j], n2d = t.run->cnt[2][j];
      const uint16_t* tab = m.tab + j * W;
      double Q[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) Q[q] = 0.0;
      if ((fa & 3u) >= 2u) {
        const int abit = (int)(fa >> 2);
        const int Rh = t.R >> 1, dr = W << abit;
        const int nA = Rh * n4, nE = nA + Rh * n2e, nT = nE + Rh * n2d;
        const float iA = t.run->inv[0][j], iE = t.run->inv[1][j], iD = t.run->inv[2][j];
        double M[16], be[4], bd[4];  // pair matrix in registers (see fwd_pair)
#pragma unroll
        for (int q = 0; q < 16; ++q) M[q] = Mp[q];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          be[q] = Mp[16 + q];
          bd[q] = Mp[20 + q];
        }
        for (int i = threadIdx.x; i < nT; i += NT) {
          if (i < nA) {
            int rr, k;
            split_item(i, n4, iA, rr, k);
            const uint32_t e = tab[k];
            const int m0 = ins0(rr, abit) * W + (int)(e & 1023u);
            const int m2 = m0 + (1 << (e >> 10));
            const double v[4] = {X[m0], X[m0 + dr], X[m2], X[m2 + dr]};
            const double w[4] = {L[m0], L[m0 + dr], L[m2], L[m2 + dr]};
            Q[0] += w[0] * v[1] - w[1] * v[0];
            Q[1] += w[0] * v[2] - w[2] * v[0];
            Q[2] += w[0] * v[3] - w[3] * v[0];
            Q[3] += w[1] * v[2] - w[2] * v[1];
            Q[4] += w[1] * v[3] - w[3] * v[1];
            Q[5] += w[2] * v[3] - w[3] * v[2];
            if (wb) {
              double ov[4], ow[4];
              mat4<true>(M, v, ov);
              mat4<true>(M, w, ow);
              X[m0] = ov[0];
              X[m0 + dr] = ov[1];
              X[m2] = ov[2];
              X[m2 + dr] = ov[3];
              L[m0] = ow[0];
              L[m0 + dr] = ow[1];
              L[m2] = ow[2];
              L[m2 + dr] = ow[3];
            }
          } else if (i < nE) {  // beta pair empty
            int rr, k;
            split_item(i - nA, n2e, iE, rr, k);
            Q[6] += rev_orbit2(X, L, ins0(rr, abit) * W + (int)(tab[n4 + k] & 1023u), dr, be, wb);
          } else {  // beta pair doubly occupied
            int rr, k;
            split_item(i - nE, n2d, iD, rr, k);
            Q[7] += rev_orbit2(X, L, ins0(rr, abit) * W + (int)(tab[n4 + n2e + k] & 1023u), dr, bd, wb);
          }
        }
        double w2[2];
        row_fold8(Q, w2, lane);
#pragma unroll
        for (int i = 0; i < 2; ++i) {  // across the rows of the wave
          w2[i] += __shfl_xor(w2[i], 16, 64);
          w2[i] += __shfl_xor(w2[i], 32, 64);
        }
        if (lane >= 60) {
          const int q = (lane & 1) * 4 + ((lane >> 1) & 1) * 2;
          red[j][q][wave] = w2[0];
          red[j][q + 1][wave] = w2[1];
        }
      } else {
        const double* Bp = Mp + ((fa & 3u) == 0u ? 24 : 28);
        const double B[4] = {Bp[0], Bp[1], Bp[2], Bp[3]};
        const int nA = t.R * n4;
        const float iA = t.run->inv[0][j];
        for (int i = threadIdx.x; i < nA; i += NT) {
          int rr, k;
          split_item(i, n4, iA, rr, k);
          const uint32_t e = tab[k];
          Q[0] += rev_orbit2(X, L, rr * W + (int)(e & 1023u), 1 << (e >> 10), B, wb);
        }
        double s = row_sum(Q[0]);
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (lane == 63) red[j][0][wave] = s;
      }
      __syncthreads();
    }
    for (int x = threadIdx.x; x < S.npairs * kQVals; x += NT) {
      const int j = x / kQVals, q = x - j * kQVals;
      const uint32_t fa = (t.adesc >> (5 * j)) & 31u;
      int src = -1;
      if ((fa & 3u) >= 2u) src = q < 8 ? q : -1;
      else src = q == ((fa & 3u) == 0u ? 8 : 9) ? 0 : -1;
      double v = 0.0;
      if (src >= 0)
        for (int w = 0; w < kWaves; ++w) v += red[j][src][w];
      partial[S.poff + (size_t)x * S.nt + t.tile] = v;
    }
    if (wb)  // not for the first stage of the program: the un-applied states are not needed
      tile_store<NT, 2>(S, t, m, psi, lam);
    __syncthreads();  // the next tile overwrites the LDS tables and red
  }
}

// ---------------------------------------------------------------------------
// Host orchestration.

static double stage_bytes(const Space& s, int vectors) { return 16.0 * (double)s.dim * vectors; }

// Tile grids: one block per dispatch slot up to blocks_per_cu blocks per CU
// (more slots are walked by the same blocks; 256 * blocks_per_cu is a
// multiple of 8, so a block stays on one XCD's share of the slots).
static unsigned tile_grid(int nslots, int blocks_per_cu) {
  return (unsigned)std::max(1, std::min(nslots, 256 * blocks_per_cu));
}

static void launch_mats(const TilePlan& T, const double* theta, int deriv_gate, hipStream_t st) {
  SQ_PROF(PC_OTHER, st, 0.0, 0.0,
          k_stage_mats<<<T.ntiled, 64, 0, st>>>(T.d_info, theta, deriv_gate, T.d_mats, 0, 0));
}

static void launch_fwd1(const StageDev& D, size_t lds, int nt, bool dagger, const double* src, double* dst,
                        hipStream_t st) {
  if (nt == 256) {
    allow_lds(k_stage_fwd<256>, lds);
    k_stage_fwd<256><<<tile_grid(D.nb, 16), 256, lds, st>>>(D, D.mats, dagger ? 1 : 0, src, dst, 0, 0);
    return;
  }
  allow_lds(k_stage_fwd<kNT>, lds);
  k_stage_fwd<kNT><<<tile_grid(D.nb, 8), kNT, lds, st>>>(D, D.mats, dagger ? 1 : 0, src, dst, 0, 0);
}

static void launch_rev1(const StageDev& D, size_t lds, int nt, double* psi, double* lam, double* partial, int wb,
                        hipStream_t st) {
  if (nt == 256) {
    allow_lds(k_stage_rev<256>, lds);
    k_stage_rev<256><<<tile_grid(D.nb, 16), 256, lds, st>>>(D, D.mats, psi, lam, partial, wb, 0, 0, 0);
    return;
  }
  allow_lds(k_stage_rev<kNT>, lds);
  k_stage_rev<kNT><<<tile_grid(D.nb, 8), kNT, lds, st>>>(D, D.mats, psi, lam, partial, wb, 0, 0, 0);
}

// A stage's launches: with a large-tile side launch, it runs on the plan's
// side stream between a fork and a join event, concurrently with the main
// launch on st (the two write disjoint amplitudes).
template <class F>
static void stage_launches(const TilePlan& T, const Stage& S, hipStream_t st, F&& launch) {
  if (S.dev_big.nb == 0) {
    launch(S.dev, false, st);
    return;
  }
  SQ_HIP(hipEventRecord(T.ev_fork, st));
  SQ_HIP(hipStreamWaitEvent(T.side, T.ev_fork, 0));
  launch(S.dev_big, true, T.side);
  launch(S.dev, false, st);
  SQ_HIP(hipEventRecord(T.ev_join, T.side));
  SQ_HIP(hipStreamWaitEvent(st, T.ev_join, 0));
}

static void launch_fwd(const TilePlan& T, const Stage& S, bool dagger, const double* src, double* dst, hipStream_t st) {
  stage_launches(T, S, st, [&](const StageDev& D, bool big, hipStream_t s) {
    launch_fwd1(D, big ? S.lds_fwd_big : S.lds_fwd, T.nt, dagger, src, dst, s);
  });
}

static void launch_rev(const TilePlan& T, const Stage& S, double* psi, double* lam, double* partial, int wb,
                       hipStream_t st) {
  stage_launches(T, S, st, [&](const StageDev& D, bool big, hipStream_t s) {
    launch_rev1(D, big ? S.lds_rev_big : S.lds_rev, T.nt, psi, lam, partial, wb, s);
  });
}

// Forward sweep (dagger: U^+, gates reversed with -theta) over the whole
// program, tiles where possible, psi = U psi_in (canonical order; psi_in may
// be psi).  A first tiled stage reads psi_in itself (its tiles cover every
// amplitude), otherwise psi_in is copied first.
// scratch >= dim doubles (derivative insertion on per-gate stages);
// gwork >= 2*dim doubles for gates that run through the per-gate kernels.
// deriv_gate >= 0 inserts T_g right after gate g (plain gradient; forward only).
void tiles_forward(const Space& s, const Program& p, const TilePlan& T, const double* theta, const double* psi_in,
                   double* psi, double* scratch, double* gwork, int deriv_gate, bool dagger, hipStream_t st) {
  launch_mats(T, theta, dagger ? -1 : deriv_gate, st);
  const int ns = (int)T.stages.size();
  const double* src = psi_in;
  for (int ii = 0; ii < ns; ++ii) {
    const int i = dagger ? ns - 1 - ii : ii;
    const Stage& S = T.stages[i];
    if (!S.tiled) {
      if (src != psi) copy_vec(src, psi, s.dim, st);
      src = psi;
      apply_gate_impl(s, p, S.g0, theta, dagger, psi, gwork, st);
      if (S.g0 == deriv_gate) {
        grad_action_impl(s, p, S.g0, psi, scratch, st);
        copy_vec(scratch, psi, s.dim, st);
      }
      continue;
    }
    SQ_PROF(PC_TILE_FWD, st, stage_bytes(s, 1), 0.0, launch_fwd(T, S, dagger, src, psi, st));
    src = psi;
  }
  if (src != psi) copy_vec(src, psi, s.dim, st);
}

// Reverse sweep on (psi, lam) with gradient dots into grad_dev[g] (device).
// gwork >= 2*dim; partial: reduction scratch of the per-gate kernels.
// mats_ready: the pair matrices of theta (no derivative insertion) are still
// in place from a forward sweep on this stream.
void tiles_reverse(const Space& s, const Program& p, const TilePlan& T, const double* theta, double* psi, double* lam,
                   double* gwork, double* partial, double* grad_dev, bool mats_ready, hipStream_t st) {
  if (!mats_ready) launch_mats(T, theta, -1, st);
  for (int i = (int)T.stages.size() - 1; i >= 0; --i) {
    const Stage& S = T.stages[i];
    if (!S.tiled) {
      reverse_step_impl(s, p, S.g0, theta, psi, lam, grad_dev, gwork, partial, st);
      continue;
    }
    const int wb = i > 0 ? 1 : 0;
    SQ_PROF(PC_TILE_REV, st, stage_bytes(s, wb ? 2 : 1), 0.0, launch_rev(T, S, psi, lam, T.d_partial, wb, st));
  }
  if (T.ntiled > 0)
    SQ_PROF(PC_REDUCE, st, 8.0 * (double)T.partial_len, 0.0,
            k_stage_grad<<<T.ntiled * kMaxPairs, kTB, 0, st>>>(T.d_info, theta, T.d_partial, grad_dev, 0, 0, 0));
}

// ---- groups of instances of one program ------------------------------------

bool tiles_group_ok(const Space& s, const Program& p) {
  const TilePlan* T = tiles_for(s, p);
  if (!T || T->ntiled == 0) return false;
  for (const Stage& S : T->stages)
    if (!S.tiled || S.dev_big.nb != 0) return false;
  return true;
}

int64_t tiles_mats_doubles(const Space& s, const Program& p) {
  const TilePlan* T = tiles_for(s, p);
  return T ? (int64_t)T->ntiled * kMaxPairs * kMatVals : 0;
}

int64_t tiles_partial_doubles(const Space& s, const Program& p) {
  const TilePlan* T = tiles_for(s, p);
  return T ? std::max<int64_t>(T->partial_len, 1) : 0;
}

// The launches of one stage over M.n instances (grid y), each stage's pair
// matrices at mats + stage index * kMaxPairs * kMatVals of every instance.
static void launch_fwd_group(const TilePlan& T, const Stage& S, const double* mats, double* psi, const Multi& M,
                             hipStream_t st) {
  const StageDev& D = S.dev;
  const double* sm = mats + (size_t)S.index * kMaxPairs * kMatVals;
  if (T.nt == 256) {
    allow_lds(k_stage_fwd<256>, S.lds_fwd);
    k_stage_fwd<256><<<dim3(tile_grid(D.nb, 16), M.n), 256, S.lds_fwd, st>>>(D, sm, 0, psi, psi, M.mats, M.vec);
  } else {
    allow_lds(k_stage_fwd<kNT>, S.lds_fwd);
    k_stage_fwd<kNT><<<dim3(tile_grid(D.nb, 8), M.n), kNT, S.lds_fwd, st>>>(D, sm, 0, psi, psi, M.mats, M.vec);
  }
  SQ_HIP(hipGetLastError());
}

static void launch_rev_group(const TilePlan& T, const Stage& S, const double* mats, double* psi, double* lam,
                             double* partial, int wb, const Multi& M, hipStream_t st) {
  const StageDev& D = S.dev;
  const double* sm = mats + (size_t)S.index * kMaxPairs * kMatVals;
  if (T.nt == 256) {
    allow_lds(k_stage_rev<256>, S.lds_rev);
    k_stage_rev<256><<<dim3(tile_grid(D.nb, 16), M.n), 256, S.lds_rev, st>>>(D, sm, psi, lam, partial, wb, M.mats,
                                                                            M.vec, M.part);
  } else {
    allow_lds(k_stage_rev<kNT>, S.lds_rev);
    k_stage_rev<kNT><<<dim3(tile_grid(D.nb, 8), M.n), kNT, S.lds_rev, st>>>(D, sm, psi, lam, partial, wb, M.mats,
                                                                          M.vec, M.part);
  }
  SQ_HIP(hipGetLastError());
}

void tiles_forward_group(const Space& s, const Program& p, const double* theta, double* psi, double* mats,
                         const Multi& M, hipStream_t st) {
  const TilePlan* T = tiles_for(s, p);
  SQ_CHECK(T && tiles_group_ok(s, p) && M.n > 0 && M.n <= 65535, SQ_ERR_ARG, "program has no group tile path");
  SQ_PROF(PC_OTHER, st, 0.0, 0.0,
          k_stage_mats<<<dim3(T->ntiled, M.n), 64, 0, st>>>(T->d_info, theta, -1, mats, M.theta, M.mats));
  for (const Stage& S : T->stages)
    SQ_PROF(PC_TILE_FWD, st, stage_bytes(s, 1) * M.n, 0.0, launch_fwd_group(*T, S, mats, psi, M, st));
}

void tiles_reverse_group(const Space& s, const Program& p, const double* theta, double* psi, double* lam,
                         const double* mats, double* partial, double* grad, const Multi& M, hipStream_t st) {
  const TilePlan* T = tiles_for(s, p);
  SQ_CHECK(T && tiles_group_ok(s, p) && M.n > 0 && M.n <= 65535, SQ_ERR_ARG, "program has no group tile path");
  for (int i = (int)T->stages.size() - 1; i >= 0; --i) {
    const Stage& S = T->stages[i];
    const int wb = i > 0 ? 1 : 0;
    SQ_PROF(PC_TILE_REV, st, stage_bytes(s, wb ? 2 : 1) * M.n, 0.0,
            launch_rev_group(*T, S, mats, psi, lam, partial, wb, M, st));
  }
  SQ_PROF(PC_REDUCE, st, 8.0 * (double)T->partial_len * M.n, 0.0,
          k_stage_grad<<<dim3(T->ntiled * kMaxPairs, M.n), kTB, 0, st>>>(T->d_info, theta, partial, grad, M.theta,
                                                                        M.part, M.grad));
}

}  // namespace sq
