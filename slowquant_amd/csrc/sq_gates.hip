// Gate programs and state propagation (replaces construct_ups_state,
// propagate_unitary and get_grad_action, operator_state_algebra.py:963-2865).
//
// Every gate exp(theta T) of the layout is expanded once, on the host, into
// raw operator products in the engine's blocked sign convention.  Three
// device paths apply them:
//   * GC_ROT / GC_SA_SINGLE: the generator is G - G^+ for ONE string G with
//     distinct indices, so T^3 = -T and the reference's closed form
//     1 + sin T + (1 - cos) T^2 is exactly a pairwise determinant rotation;
//     applied in place by one thread per (determinant, partner) pair.
//   * GC_POLY / GC_TAYLOR: the reference's polynomial in T (multi-frequency
//     spin-adapted doubles, degenerate index sets), built from gather passes
//     out[I] = sum_t coef_t <I|O_t|J> in[J].
//   * fused reverse-sweep kernels used by the adjoint gradient.

#include <cmath>
#include <cstring>

#include "sq_internal.h"

namespace sq {

// ---------------------------------------------------------------------------
// Host-side expansion of gates into blocked-convention terms.

namespace {

struct RawOp {
  int spin_orb;
  bool cre;
};

Term make_term(const std::vector<RawOp>& ops, double coef) {
  Term t{};
  t.coef = coef;
  int swaps = 0;
  for (size_t i = 0; i < ops.size(); ++i)
    for (size_t j = i + 1; j < ops.size(); ++j)
      if ((ops[i].spin_orb & 1) && !(ops[j].spin_orb & 1)) ++swaps;
  if (swaps & 1) t.coef = -t.coef;
  int nca = 0, naa = 0, ncb = 0, nab = 0;
  for (const RawOp& o : ops) {
    SpinOps& so = (o.spin_orb & 1) ? t.b : t.a;
    SQ_CHECK(so.n < 12, SQ_ERR_ARG, "operator string too long");
    so.orb[so.n] = (int8_t)(o.spin_orb >> 1);
    so.cre[so.n] = o.cre ? 1 : 0;
    ++so.n;
    if (o.spin_orb & 1) {
      (o.cre ? ncb : nab)++;
    } else {
      (o.cre ? nca : naa)++;
    }
  }
  SQ_CHECK(nca == naa && ncb == nab, SQ_ERR_SPACE,
           "excitation does not conserve the alpha/beta electron numbers (leaves the CI space)");
  return t;
}

std::vector<RawOp> adjoint(const std::vector<RawOp>& ops) {
  std::vector<RawOp> r;
  for (auto it = ops.rbegin(); it != ops.rend(); ++it) r.push_back({it->spin_orb, !it->cre});
  return r;
}

// G = a+_{c0} a+_{c1} ... a_{a0} a_{a1} ... in the given order.
std::vector<RawOp> product(const std::vector<int>& cre, const std::vector<int>& ann) {
  std::vector<RawOp> r;
  for (int c : cre) r.push_back({c, true});
  for (int a : ann) r.push_back({a, false});
  return r;
}

bool distinct(const std::vector<RawOp>& ops) {
  for (size_t i = 0; i < ops.size(); ++i)
    for (size_t j = i + 1; j < ops.size(); ++j)
      if (ops[i].spin_orb == ops[j].spin_orb) return false;
  return true;
}

// E_pq (spatial) as raw terms: sum_sigma a+_{p sigma} a_{q sigma}.
void push_EE(std::vector<std::pair<std::vector<RawOp>, double>>& out, int p, int q, int r, int s, double c) {
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y)
      out.push_back({product({2 * p + x}, {2 * q + x}), c}), out.back().first.push_back({2 * r + y, true}),
          out.back().first.push_back({2 * s + y, false});
}

}  // namespace

// Host copy of apply_ops: number of strings a spin part applies to.
static int64_t count_applies(const SpinOps& o, const std::vector<uint32_t>& strs) {
  int64_t c = 0;
  for (uint32_t m0 : strs) {
    uint32_t m = m0;
    bool ok = true;
    for (int j = o.n - 1; j >= 0 && ok; --j) {
      uint32_t bit = 1u << o.orb[j];
      if (o.cre[j]) {
        if (m & bit) ok = false;
        m |= bit;
      } else {
        if (!(m & bit)) ok = false;
        m &= ~bit;
      }
    }
    c += ok;
  }
  return c;
}

void spin_part_list(const std::vector<uint32_t>& strs, const SpinOps& o, std::vector<int2>& lists,
                    std::map<std::pair<const void*, std::vector<int>>, std::pair<int, int>>& cache, int& off,
                    int& len) {
  if (o.n == 0) {
    off = -1;
    len = (int)strs.size();
    return;
  }
  std::vector<int> key{o.n};
  for (int j = 0; j < o.n; ++j) key.push_back(o.orb[j] * 2 + (o.cre[j] ? 1 : 0));
  const auto ck = std::make_pair((const void*)&strs, key);
  const auto it = cache.find(ck);
  if (it != cache.end()) {
    off = it->second.first;
    len = it->second.second;
    return;
  }
  std::map<uint32_t, int> rank;
  for (int i = 0; i < (int)strs.size(); ++i) rank[strs[i]] = i;
  off = (int)lists.size();
  for (int i = 0; i < (int)strs.size(); ++i) {
    uint32_t m = strs[i];
    int par = 0;
    bool ok = true;
    for (int j = o.n - 1; j >= 0 && ok; --j) {
      const uint32_t bit = 1u << o.orb[j];
      par += __builtin_popcount(m & (bit - 1u));
      if (o.cre[j]) {
        if (m & bit) ok = false;
        m |= bit;
      } else {
        if (!(m & bit)) ok = false;
        m &= ~bit;
      }
    }
    if (!ok) continue;
    lists.push_back(int2{i, rank.at(m) | (int)((uint32_t)(par & 1) << 31)});
  }
  len = (int)lists.size() - off;
  cache[ck] = {off, len};
}

static void add_gate(Program& P, const sq_gate_desc& g, int n_orb) {
  auto check_idx = [&](int count, int limit) {
    SQ_CHECK(g.n_idx == count, SQ_ERR_ARG, "wrong number of indices for excitation type");
    for (int i = 0; i < count; ++i)
      SQ_CHECK(g.idx[i] >= 0 && g.idx[i] < limit, SQ_ERR_SPACE, "orbital index outside the active space");
  };
  GateDev gd{};
  gd.scale = 1.0;
  gd.term_begin = (int)P.terms.size();
  int pp = -1, pq = -1;
  auto add_rot_or_taylor = [&](const std::vector<RawOp>& G, double coef) {
    if (distinct(G)) {
      gd.kind = GC_ROT;
      P.terms.push_back(make_term(G, coef));
      P.terms.push_back(make_term(adjoint(G), -coef));
      gd.scale = std::fabs(coef);
    } else {
      gd.kind = GC_TAYLOR;
      P.terms.push_back(make_term(G, coef));
      P.terms.push_back(make_term(adjoint(G), -coef));
    }
  };
  switch (g.kind) {
    case SQ_GATE_SINGLE: {
      check_idx(2, 2 * n_orb);
      add_rot_or_taylor(product({g.idx[1]}, {g.idx[0]}), 1.0);
      break;
    }
    case SQ_GATE_DOUBLE: {
      check_idx(4, 2 * n_orb);
      const int* x = g.idx;
      add_rot_or_taylor(product({x[2], x[3]}, {x[1], x[0]}), 1.0);
      int a = x[0] >> 1, b = x[2] >> 1;
      if (x[0] == 2 * a && x[1] == 2 * a + 1 && x[2] == 2 * b && x[3] == 2 * b + 1 && b == a + 1) {
        pp = a;
        pq = b;
      }
      break;
    }
    case SQ_GATE_TRIPLE:
    case SQ_GATE_QUADRUPLE:
    case SQ_GATE_QUINTUPLE:
    case SQ_GATE_SEXTUPLE: {
      int r = g.kind - SQ_GATE_TRIPLE + 3;
      check_idx(2 * r, 2 * n_orb);
      std::vector<int> ann, cre;
      for (int i = r - 1; i >= 0; --i) ann.push_back(g.idx[i]);
      for (int i = 0; i < r; ++i) cre.push_back(g.idx[r + i]);
      add_rot_or_taylor(product(cre, ann), 1.0);
      break;
    }
    case SQ_GATE_SA_SINGLE: {
      check_idx(2, n_orb);
      int i = g.idx[0], a = g.idx[1];
      gd.kind = GC_SA_SINGLE;
      if (i == a) {
        // G1(2i,2i) - h.c. = 0: the gate is the identity.
        gd.kind = GC_TAYLOR;
      } else {
        for (int s = 0; s < 2; ++s) {
          auto G = product({2 * a + s}, {2 * i + s});
          P.terms.push_back(make_term(G, 1.0));
          P.terms.push_back(make_term(adjoint(G), -1.0));
        }
        if (a == i + 1) {
          pp = i;
          pq = a;
        }
      }
      break;
    }
    case SQ_GATE_SA_SINGLE_UCC: {
      check_idx(2, n_orb);
      int i = g.idx[0], a = g.idx[1];
      gd.kind = GC_TAYLOR;
      const double c = std::pow(2.0, -0.5);
      for (int s = 0; s < 2; ++s) {
        auto G = product({2 * a + s}, {2 * i + s});
        P.terms.push_back(make_term(G, c));
        P.terms.push_back(make_term(adjoint(G), -c));
      }
      break;
    }
    case SQ_GATE_SA_DOUBLE_1:
    case SQ_GATE_SA_DOUBLE_2:
    case SQ_GATE_SA_DOUBLE_3:
    case SQ_GATE_SA_DOUBLE_4:
    case SQ_GATE_SA_DOUBLE_5: {
      check_idx(4, n_orb);
      int i = g.idx[0], j = g.idx[1], a = g.idx[2], b = g.idx[3];
      int cas = g.kind - SQ_GATE_SA_DOUBLE_1 + 1;
      if (cas == 1 && i == j && a == b && i != a) {
        // 1/2 E_ai E_ai = a+_{a,al} a_{i,al} a+_{a,be} a_{i,be}: one string, T^3 = -T.
        std::vector<RawOp> G = {{2 * a, true}, {2 * i, false}, {2 * a + 1, true}, {2 * i + 1, false}};
        gd.kind = GC_ROT;
        P.terms.push_back(make_term(G, 1.0));
        P.terms.push_back(make_term(adjoint(G), -1.0));
        break;
      }
      double fac, sgn;
      if (cas <= 4) {
        double f = (a == b ? 2.0 : 1.0) * (i == j ? 2.0 : 1.0);
        fac = 0.5 / std::sqrt(f);
        sgn = 1.0;
      } else {
        fac = 1.0 / (2.0 * std::sqrt(3.0));
        sgn = -1.0;
      }
      std::vector<std::pair<std::vector<RawOp>, double>> raw;
      push_EE(raw, a, i, b, j, fac);        // E_ai E_bj
      push_EE(raw, a, j, b, i, sgn * fac);  // +- E_aj E_bi
      size_t nfwd = raw.size();
      for (size_t t = 0; t < nfwd; ++t) raw.push_back({adjoint(raw[t].first), -raw[t].second});
      for (auto& r : raw) P.terms.push_back(make_term(r.first, r.second));
      gd.kind = (cas == 1) ? GC_TAYLOR : GC_POLY;
      gd.poly = cas;
      break;
    }
    default:
      throw Error{SQ_ERR_TYPE, "Got unknown excitation type, kind=" + std::to_string(g.kind)};
  }
  gd.term_end = (int)P.terms.size();
  if (gd.kind == GC_ROT || gd.kind == GC_SA_SINGLE) {
    const Space& sp = *P.space;
    for (int h = 0; h < (gd.kind == GC_SA_SINGLE ? 2 : 1); ++h) {
      const Term& T = P.terms[gd.term_begin + 2 * h];
      gd.pairs[h] = (double)count_applies(T.a, sp.h_astr) * (double)count_applies(T.b, sp.h_bstr);
    }
  }
  P.gates.push_back(gd);
  P.pair_p.push_back(pp);
  P.pair_q.push_back(pq);
}

// ---------------------------------------------------------------------------
// Device helpers.

// Frequencies and T^k coefficients of the multi-frequency SA doubles
// (operator_state_algebra.py:1086-1279).  Row k-1 multiplies T^k; odd k use
// sin(S_f th), even k use cos(S_f th) - 1.
struct PolyTable {
  int nf, nk;
  double S[5];
  double K[10][5];
};

__constant__ PolyTable c_poly[4];  // index: case 2,3 -> 0; 4 -> 1; 5 -> 2; Taylor -> 3

static PolyTable host_poly(int which) {
  PolyTable t{};
  const double r2 = std::sqrt(2.0), r3 = std::sqrt(3.0);
  if (which == 0) {
    t.nf = 2;
    t.nk = 4;
    double S[] = {1, r2 / 2};
    double K[4][2] = {{-1, 2 * r2}, {1, -4}, {-2, 2 * r2}, {2, -4}};
    std::memcpy(t.S, S, sizeof(S));
    for (int k = 0; k < 4; ++k)
      for (int f = 0; f < 2; ++f) t.K[k][f] = K[k][f];
  } else if (which == 1) {
    t.nf = 4;
    t.nk = 8;
    double S[] = {1, r2, r2 / 2, 0.5};
    double K[8][4] = {{2. / 3, -r2 / 42, -8 * r2 / 3, 128. / 21}, {-2. / 3, 1. / 42, 16. / 3, -256. / 21},
                      {13. / 3, -r2 / 6, -44 * r2 / 3, 64. / 3},  {-13. / 3, 1. / 6, 88. / 3, -128. / 3},
                      {22. / 3, -r2 / 3, -52 * r2 / 3, 64. / 3},  {-22. / 3, 1. / 3, 104. / 3, -128. / 3},
                      {8. / 3, -4 * r2 / 21, -16 * r2 / 3, 128. / 21}, {-8. / 3, 4. / 21, 32. / 3, -256. / 21}};
    std::memcpy(t.S, S, sizeof(S));
    for (int k = 0; k < 8; ++k)
      for (int f = 0; f < 4; ++f) t.K[k][f] = K[k][f];
  } else if (which == 2) {
    t.nf = 5;
    t.nk = 10;
    double S[] = {r2, r2 / 2, r3 / 3, r3 / 2, r3 / 6};
    double K[10][5] = {
        {r2 / 1150, 8 * r2 / 5, -54 * r3 / 25, -16 * r3 / 75, 432 * r3 / 115},
        {-1. / 1150, -16. / 5, 162. / 25, 32. / 75, -2592. / 115},
        {11 * r2 / 690, 404 * r2 / 15, -171 * r3 / 5, -56 * r3 / 15, 2952 * r3 / 115},
        {-11. / 690, -808. / 15, 513. / 5, 112. / 15, -17712. / 115},
        {133 * r2 / 1725, 308 * r2 / 3, -2718 * r3 / 25, -1192 * r3 / 75, 1368 * r3 / 23},
        {-133. / 1725, -616. / 3, 8154. / 25, 2384. / 75, -8208. / 23},
        {16 * r2 / 115, 608 * r2 / 5, -576 * r3 / 5, -112 * r3 / 5, 6192 * r3 / 115},
        {-16. / 115, -1216. / 5, 1728. / 5, 224. / 5, -37152. / 115},
        {48 * r2 / 575, 192 * r2 / 5, -864 * r3 / 25, -192 * r3 / 25, 1728 * r3 / 115},
        {-48. / 575, -384. / 5, 2592. / 25, 384. / 25, -10368. / 115}};
    std::memcpy(t.S, S, sizeof(S));
    for (int k = 0; k < 10; ++k)
      for (int f = 0; f < 5; ++f) t.K[k][f] = K[k][f];
  } else {
    // 1 + sin(th) T + (1 - cos(th)) T^2 == rows (1), (-1) with S = 1.
    t.nf = 1;
    t.nk = 2;
    t.S[0] = 1.0;
    t.K[0][0] = 1.0;
    t.K[1][0] = -1.0;
  }
  return t;
}

static int poly_slot(const GateDev& g) {
  if (g.kind == GC_TAYLOR) return 3;
  if (g.poly == 2 || g.poly == 3) return 0;
  if (g.poly == 4) return 1;
  return 2;
}

__device__ __forceinline__ double poly_coef(const PolyTable& t, int k, double th) {
  // k is 1-based power of T
  double c = 0.0;
  for (int f = 0; f < t.nf; ++f) {
    double a = t.S[f] * th;
    c += t.K[k - 1][f] * ((k & 1) ? sin(a) : (cos(a) - 1.0));
  }
  return c;
}

// Apply a spin-resolved op product to a ket mask (ops[n-1] acts first).
__device__ __forceinline__ bool apply_ops(const SpinOps& o, uint32_t& m, int& par) {
  for (int j = o.n - 1; j >= 0; --j) {
    int orb = o.orb[j];
    uint32_t bit = 1u << orb;
    par += __popc(m & (bit - 1u));
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

// Apply the adjoint of a spin-resolved product to a bra mask: gives the ket
// J with <I|O|J> != 0 and its sign (ops[0]^+ acts first).
__device__ __forceinline__ bool apply_ops_adjoint(const SpinOps& o, uint32_t& m, int& par) {
  for (int j = 0; j < o.n; ++j) {
    int orb = o.orb[j];
    uint32_t bit = 1u << orb;
    par += __popc(m & (bit - 1u));
    if (!o.cre[j]) {
      if (m & bit) return false;
      m |= bit;
    } else {
      if (!(m & bit)) return false;
      m &= ~bit;
    }
  }
  return true;
}

struct SpaceView {
  int nsa, nsb;
  const uint32_t* astr;
  const uint32_t* bstr;
  const int32_t* arank;
  const int32_t* brank;
};

static SpaceView view(const Space& s) { return SpaceView{s.nsa, s.nsb, s.d_astr, s.d_bstr, s.d_arank, s.d_brank}; }

constexpr int kBlk = 256;

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += red[w];
  }
  return r;
}

// One rotation half over its spin-part lists: thread (row r, column c) owns
// the pair of determinants (A_r, B_c) -> (image of A_r, image of B_c), with
// A, B list entries (string, image | parity << 31) or, for an identity spin
// part, every string.  Only touched pairs get threads; along an identity beta
// part consecutive lanes touch consecutive amplitudes.  The rotation's cosine
// and sine are computed once per block.
__device__ __forceinline__ int2 rot_entry(const int2* __restrict__ lists, int off, int i) {
  return off < 0 ? int2{i, i} : lists[off + i];
}

__global__ void k_rotate_list(int nsb, const int2* __restrict__ lists, int aoff, int na, int boff, int nb,
                              double csign, const double* __restrict__ theta, int gidx, double scale, double sgn,
                              double* __restrict__ psi) {
  __shared__ double cs[2];
  const double t = theta[gidx];
  if (fabs(t) < 1e-28) return;  // uniform over the grid
  if (threadIdx.x == 0) sincos(sgn * scale * t, &cs[1], &cs[0]);
  __syncthreads();
  const double c = cs[0], s = cs[1];
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= nb) return;
  const int2 B = rot_entry(lists, boff, col);
  const int jb = B.y & 0x7fffffff, pb = (int)((uint32_t)B.y >> 31);
  for (int row = blockIdx.y; row < na; row += gridDim.y) {
    const int2 A = rot_entry(lists, aoff, row);
    const int ja = A.y & 0x7fffffff, pa = (int)((uint32_t)A.y >> 31);
    const double sigma = ((pa ^ pb) ? -1.0 : 1.0) * csign;
    const size_t D = (size_t)A.x * nsb + B.x, J = (size_t)ja * nsb + jb;
    const double pd = psi[D], pj = psi[J];
    psi[D] = c * pd - sigma * s * pj;
    psi[J] = sigma * s * pd + c * pj;
  }
}

// Reverse-sweep step of one rotation half over its lists: partial dots
// scale * <lam|T psi> (before un-applying) per block, then both vectors rotate
// by -theta.
__global__ void k_rotate_list_adj(int nsb, const int2* __restrict__ lists, int aoff, int na, int boff, int nb,
                                  double csign, const double* __restrict__ theta, int gidx, double scale,
                                  double* __restrict__ psi, double* __restrict__ lam, double* __restrict__ partial) {
  __shared__ double red[kBlk / 64];
  __shared__ double cs[2];
  const double t = theta[gidx];
  const bool skip = fabs(t) < 1e-28;
  if (threadIdx.x == 0) sincos(-scale * t, &cs[1], &cs[0]);
  __syncthreads();
  const double c = cs[0], s = cs[1];
  double contrib = 0.0;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col < nb) {
    const int2 B = rot_entry(lists, boff, col);
    const int jb = B.y & 0x7fffffff, pb = (int)((uint32_t)B.y >> 31);
    for (int row = blockIdx.y; row < na; row += gridDim.y) {
      const int2 A = rot_entry(lists, aoff, row);
      const int ja = A.y & 0x7fffffff, pa = (int)((uint32_t)A.y >> 31);
      const double sigma = ((pa ^ pb) ? -1.0 : 1.0) * csign;
      const size_t D = (size_t)A.x * nsb + B.x, J = (size_t)ja * nsb + jb;
      const double pd = psi[D], pj = psi[J], ld = lam[D], lj = lam[J];
      contrib += scale * sigma * (lj * pd - ld * pj);
      if (!skip) {
        psi[D] = c * pd - sigma * s * pj;
        psi[J] = sigma * s * pd + c * pj;
        lam[D] = c * ld - sigma * s * lj;
        lam[J] = sigma * s * ld + c * lj;
      }
    }
  }
  const double tot = block_sum(contrib, red);
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

// Generic gather: nxt[I] = sum_t coef_t <I|O_t|J> cur[J];
// if acc: acc[I] += ca * cur[I] + cb * nxt[I] with ca, cb from the gate's
// polynomial table (powers ka, kb; 0 means no term).
__global__ void k_tapply(SpaceView v, const Term* __restrict__ terms, int t0, int t1,
                         const double* __restrict__ cur, double* __restrict__ nxt, double* __restrict__ acc,
                         const double* __restrict__ theta, int gidx, int slot, int ka, int kb, double tsign) {
  int ia = blockIdx.y;
  int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= v.nsb) return;
  uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
  double val = 0.0;
  for (int t = t0; t < t1; ++t) {
    const Term& T = terms[t];
    uint32_t a = a0, b = b0;
    int par = 0;
    if (!apply_ops_adjoint(T.a, a, par)) continue;
    if (!apply_ops_adjoint(T.b, b, par)) continue;
    int ra = v.arank[a], rb = v.brank[b];
    if (ra < 0 || rb < 0) continue;
    double x = cur[(size_t)ra * v.nsb + rb];
    val += ((par & 1) ? -T.coef : T.coef) * x;
  }
  size_t I = (size_t)ia * v.nsb + ib;
  if (nxt) nxt[I] = val;
  if (acc) {
    double th = theta ? tsign * theta[gidx] : 0.0;
    const PolyTable& P = c_poly[slot];
    double r = 0.0;
    if (ka > 0) r += poly_coef(P, ka, th) * cur[I];
    if (kb > 0) r += poly_coef(P, kb, th) * val;
    acc[I] += r;
  }
}

// Dense matrix of an operator: row I of M gets coef * sign at column J for
// every term whose adjoint maps I to J, i.e. M[I][J] = <I|O|J> (the rows are
// the gather of k_tapply).  With ref != 0 both indices are in the reference's
// interleaved sign convention: M_ref[I][J] = s(I) s(J) M[I][J].  M must be
// zero on entry.
__device__ __forceinline__ int interleave_parity(uint32_t a, uint32_t b) {
  int par = 0;
  while (a) {
    const int p = __ffs(a) - 1;
    a &= a - 1;
    par += __popc(b & ((1u << p) - 1u));
  }
  return par;
}

__global__ void k_operator_matrix(SpaceView v, const Term* __restrict__ terms, int t0, int t1, int ref,
                                  double* __restrict__ M) {
  const int ia = blockIdx.y;
  const int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= v.nsb) return;
  const uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
  const size_t dim = (size_t)v.nsa * v.nsb;
  double* row = M + ((size_t)ia * v.nsb + ib) * dim;
  const int pI = ref ? interleave_parity(a0, b0) : 0;
  for (int t = t0; t < t1; ++t) {
    const Term& T = terms[t];
    uint32_t a = a0, b = b0;
    int par = pI;
    if (!apply_ops_adjoint(T.a, a, par)) continue;
    if (!apply_ops_adjoint(T.b, b, par)) continue;
    const int ra = v.arank[a], rb = v.brank[b];
    if (ra < 0 || rb < 0) continue;
    if (ref) par += interleave_parity(a, b);
    row[(size_t)ra * v.nsb + rb] += (par & 1) ? -T.coef : T.coef;
  }
}

__global__ void k_fill_zero(double* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// Partial dots sum_I lam[I] * (T psi)[I] per block.
__global__ void k_tapply_dot(SpaceView v, const Term* __restrict__ terms, int t0, int t1,
                             const double* __restrict__ psi, const double* __restrict__ lam,
                             double* __restrict__ partial) {
  __shared__ double red[kBlk / 64];
  int ia = blockIdx.y;
  int ib = blockIdx.x * blockDim.x + threadIdx.x;
  double val = 0.0;
  if (ib < v.nsb) {
    uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
    for (int t = t0; t < t1; ++t) {
      const Term& T = terms[t];
      uint32_t a = a0, b = b0;
      int par = 0;
      if (!apply_ops_adjoint(T.a, a, par)) continue;
      if (!apply_ops_adjoint(T.b, b, par)) continue;
      int ra = v.arank[a], rb = v.brank[b];
      if (ra < 0 || rb < 0) continue;
      val += ((par & 1) ? -T.coef : T.coef) * psi[(size_t)ra * v.nsb + rb];
    }
    val *= lam[(size_t)ia * v.nsb + ib];
  }
  double tot = block_sum(val, red);
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

// Sum partials in a fixed order into out[slot] (+= when accumulate).
__global__ void k_reduce(const double* __restrict__ partial, int64_t n, double* __restrict__ out, int slot,
                         int accumulate, double factor) {
  __shared__ double red[kBlk / 64];
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v += partial[i];
  double tot = block_sum(v, red);
  if (threadIdx.x == 0) out[slot] = (accumulate ? out[slot] : 0.0) + factor * tot;
}

__global__ void k_copy(const double* __restrict__ in, double* __restrict__ out, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

__global__ void k_scale_theta(const double* __restrict__ in, double* __restrict__ out, int n, double f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f * in[i];
}

// ---------------------------------------------------------------------------
// Host orchestration.

static dim3 grid2d(const Space& s) { return dim3((s.nsb + kBlk - 1) / kBlk, s.nsa); }

void copy_vec(const double* in, double* out, int64_t n, hipStream_t st) {
  if (in == out || n == 0) return;
  SQ_PROF(PC_COPY, st, 16.0 * (double)n, 0.0, k_copy<<<(unsigned)((n + kBlk - 1) / kBlk), kBlk, 0, st>>>(in, out, n));
}

int64_t nblocks2d(const Space& s) { return (int64_t)((s.nsb + kBlk - 1) / kBlk) * s.nsa; }

// Rotation half h of gate g over its spin-part lists: columns over the beta
// list, rows over the alpha list (grid.y capped, rows strided).
static dim3 rot_grid(const int4& L) {
  return dim3((unsigned)std::max(1, (L.w + kBlk - 1) / kBlk), (unsigned)std::max(1, std::min(L.y, 65535)));
}

static void rotate_half(const Space& s, const Program& p, int g, int h, const double* theta_dev, double scale,
                        double sgn, double* psi, hipStream_t st) {
  const int4 L = p.rot_lists[2 * g + h];
  if (L.y == 0 || L.w == 0) return;
  const Term& T = p.terms[p.gates[g].term_begin + 2 * h];
  SQ_PROF(PC_ROTATE, st, 32.0 * (double)L.y * (double)L.w, 0.0,
          k_rotate_list<<<rot_grid(L), kBlk, 0, st>>>(s.nsb, p.d_lists, L.x, L.y, L.z, L.w, T.coef < 0 ? -1.0 : 1.0,
                                                      theta_dev, g, scale, sgn, psi));
}

// Reverse step of one rotation half; returns the number of partial dots.
static int64_t rotate_half_adj(const Space& s, const Program& p, int g, int h, const double* theta_dev, double scale,
                               double* psi, double* lam, double* partial, hipStream_t st) {
  const int4 L = p.rot_lists[2 * g + h];
  if (L.y == 0 || L.w == 0) return 0;
  const Term& T = p.terms[p.gates[g].term_begin + 2 * h];
  const dim3 grid = rot_grid(L);
  SQ_PROF(PC_ROTATE_ADJ, st, 64.0 * (double)L.y * (double)L.w, 0.0,
          k_rotate_list_adj<<<grid, kBlk, 0, st>>>(s.nsb, p.d_lists, L.x, L.y, L.z, L.w, T.coef < 0 ? -1.0 : 1.0,
                                                   theta_dev, g, scale, psi, lam, partial));
  return (int64_t)grid.x * grid.y;
}

// theta_dev points at a device theta vector; negate applies -theta (dagger).
// work: >= 2*dim doubles (polynomial temporaries).
void apply_gate_impl(const Space& s, const Program& p, int g, const double* theta_dev, bool negate, double* psi,
                     double* work, hipStream_t st) {
  const GateDev& gd = p.gates[g];
  SpaceView v = view(s);
  dim3 grid = grid2d(s);
  double sgn = negate ? -1.0 : 1.0;
  if (gd.kind == GC_ROT) {
    rotate_half(s, p, g, 0, theta_dev, gd.scale, sgn, psi, st);
  } else if (gd.kind == GC_SA_SINGLE) {
    // the reference applies the alpha rotation, then the beta one
    rotate_half(s, p, g, 0, theta_dev, 1.0, sgn, psi, st);
    rotate_half(s, p, g, 1, theta_dev, 1.0, sgn, psi, st);
  } else {
    if (gd.term_end == gd.term_begin) return;  // identity generator
    int slot = poly_slot(gd);
    PolyTable t = host_poly(slot);
    double* bufs[2] = {work, work + s.dim};
    const double* cur = psi;
    for (int k = 1; k <= t.nk; ++k) {
      double* nxt = bufs[(k - 1) & 1];
      const double d8 = 8.0 * (double)s.dim;
      if (k == 1) {
        SQ_PROF(PC_TAPPLY, st, 2 * d8, 0.0,
                k_tapply<<<grid, kBlk, 0, st>>>(v, p.d_terms, gd.term_begin, gd.term_end, cur, nxt, nullptr,
                                                theta_dev, g, slot, 0, 0, sgn));
      } else {
        SQ_PROF(PC_TAPPLY, st, 4 * d8, 0.0,
                k_tapply<<<grid, kBlk, 0, st>>>(v, p.d_terms, gd.term_begin, gd.term_end, cur, nxt, psi, theta_dev,
                                                g, slot, k == 2 ? 1 : 0, k, sgn));
      }
      cur = nxt;
    }
  }
}

void grad_action_impl(const Space& s, const Program& p, int g, const double* in, double* out, hipStream_t st) {
  const GateDev& gd = p.gates[g];
  SpaceView v = view(s);
  SQ_PROF(PC_TAPPLY, st, 16.0 * (double)s.dim, 0.0,
          k_tapply<<<grid2d(s), kBlk, 0, st>>>(v, p.d_terms, gd.term_begin, gd.term_end, in, out, nullptr, nullptr, g,
                                              0, 0, 0, 1.0));
}

// Device-side partial dot <lam|T_g psi> into partial buffer; returns #blocks.
int64_t tdot_impl(const Space& s, const Program& p, int g, const double* psi, const double* lam, double* partial,
                  hipStream_t st) {
  const GateDev& gd = p.gates[g];
  SQ_PROF(PC_TAPPLY, st, 16.0 * (double)s.dim, 0.0,
          k_tapply_dot<<<grid2d(s), kBlk, 0, st>>>(view(s), p.d_terms, gd.term_begin, gd.term_end, psi, lam,
                                                  partial));
  return nblocks2d(s);
}

void reduce_impl(const double* partial, int64_t n, double* out_dev, int slot, bool accumulate, double factor,
                 hipStream_t st) {
  SQ_PROF(PC_REDUCE, st, 8.0 * (double)n, 0.0,
          k_reduce<<<1, kBlk, 0, st>>>(partial, n, out_dev, slot, accumulate ? 1 : 0, factor));
}

// Reverse-sweep step for gate g on (psi, lam): grad_dev[g] = 2 <lam|T_g psi>,
// then both vectors un-apply gate g.  work: >= 3*dim + partial space.
void reverse_step_impl(const Space& s, const Program& p, int g, const double* theta_dev, double* psi, double* lam,
                       double* grad_dev, double* work, double* partial, hipStream_t st) {
  const GateDev& gd = p.gates[g];
  int64_t nb = nblocks2d(s);
  if (gd.kind == GC_ROT) {
    const int64_t np = rotate_half_adj(s, p, g, 0, theta_dev, gd.scale, psi, lam, partial, st);
    reduce_impl(partial, np, grad_dev, g, false, 2.0, st);
  } else if (gd.kind == GC_SA_SINGLE) {
    // un-apply beta then alpha; each half contributes its own <lam|T psi>
    int64_t np = rotate_half_adj(s, p, g, 1, theta_dev, 1.0, psi, lam, partial, st);
    reduce_impl(partial, np, grad_dev, g, false, 2.0, st);
    np = rotate_half_adj(s, p, g, 0, theta_dev, 1.0, psi, lam, partial, st);
    reduce_impl(partial, np, grad_dev, g, true, 2.0, st);
  } else {
    if (gd.term_end == gd.term_begin) {
      reduce_impl(partial, 0, grad_dev, g, false, 0.0, st);
      return;
    }
    tdot_impl(s, p, g, psi, lam, partial, st);
    reduce_impl(partial, nb, grad_dev, g, false, 2.0, st);
    apply_gate_impl(s, p, g, theta_dev, true, psi, work, st);
    apply_gate_impl(s, p, g, theta_dev, true, lam, work, st);
  }
}

void upload_poly_tables() {
  static bool done = false;
  if (done) return;
  PolyTable t[4];
  for (int i = 0; i < 4; ++i) t[i] = host_poly(i);
  SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_poly), t, sizeof(t)));
  done = true;
}

}  // namespace sq

using namespace sq;

extern "C" {

int sq_program_create(const sq_space* sp, const sq_gate_desc* gates, int n_gates, sq_program** out) {
  return guarded([&] {
    SQ_CHECK(sp && out && (gates || n_gates == 0) && n_gates >= 0, SQ_ERR_ARG, "invalid arguments");
    upload_poly_tables();
    auto* pr = new sq_program();
    Program& P = pr->p;
    P.space = &sp->s;
    try {
      for (int i = 0; i < n_gates; ++i) add_gate(P, gates[i], sp->s.n);
      {
        std::vector<int2> lists;
        std::map<std::pair<const void*, std::vector<int>>, std::pair<int, int>> cache;
        P.rot_lists.assign(2 * P.gates.size(), int4{-1, 0, -1, 0});
        for (size_t g = 0; g < P.gates.size(); ++g) {
          const GateDev& gd = P.gates[g];
          if (gd.kind != GC_ROT && gd.kind != GC_SA_SINGLE) continue;
          for (int h = 0; h < (gd.kind == GC_SA_SINGLE ? 2 : 1); ++h) {
            const Term& T = P.terms[gd.term_begin + 2 * h];
            int4& L = P.rot_lists[2 * g + h];
            spin_part_list(sp->s.h_astr, T.a, lists, cache, L.x, L.y);
            spin_part_list(sp->s.h_bstr, T.b, lists, cache, L.z, L.w);
          }
        }
        SQ_HIP(hipMalloc(&P.d_lists, sizeof(int2) * std::max<size_t>(lists.size(), 1)));
        if (!lists.empty())
          SQ_HIP(hipMemcpy(P.d_lists, lists.data(), sizeof(int2) * lists.size(), hipMemcpyHostToDevice));
      }
      SQ_HIP(hipMalloc(&P.d_gates, sizeof(GateDev) * std::max<size_t>(P.gates.size(), 1)));
      SQ_HIP(hipMalloc(&P.d_terms, sizeof(Term) * std::max<size_t>(P.terms.size(), 1)));
      if (!P.gates.empty())
        SQ_HIP(hipMemcpy(P.d_gates, P.gates.data(), sizeof(GateDev) * P.gates.size(), hipMemcpyHostToDevice));
      if (!P.terms.empty())
        SQ_HIP(hipMemcpy(P.d_terms, P.terms.data(), sizeof(Term) * P.terms.size(), hipMemcpyHostToDevice));
    } catch (...) {
      sq_program_destroy(pr);
      throw;
    }
    *out = pr;
  });
}

int sq_program_destroy(sq_program* pr) {
  return guarded([&] {
    if (!pr) return;
    if (pr->p.d_gates) (void)hipFree(pr->p.d_gates);
    if (pr->p.d_terms) (void)hipFree(pr->p.d_terms);
    if (pr->p.d_lists) (void)hipFree(pr->p.d_lists);
    if (pr->p.d_term_fast) (void)hipFree(pr->p.d_term_fast);
    if (pr->p.ucc_csr) ucc_csr_free(pr->p.ucc_csr);
    if (pr->p.tiles) destroy_tile_plan(pr->p.tiles);
    delete pr;
  });
}

int sq_program_ngates(const sq_program* pr) { return pr ? (int)pr->p.gates.size() : -1; }

int sq_program_set_tile_shape(sq_program* pr, int target, int threads) {
  return guarded([&] {
    SQ_CHECK(pr && target >= 0 && (threads == 0 || threads == 256 || threads == 512), SQ_ERR_ARG,
             "tile target must be >= 0 and threads 0, 256 or 512");
    if (pr->p.tile_target == target && pr->p.tile_threads == threads) return;
    SQ_HIP(hipDeviceSynchronize());  // the old plan's buffers may be in use by queued kernels
    if (pr->p.tiles) destroy_tile_plan(pr->p.tiles);
    pr->p.tiles = nullptr;
    pr->p.tiles_built = false;
    ++pr->p.plan_gen;
    pr->p.tile_target = target;
    pr->p.tile_threads = threads;
  });
}

int sq_operator_create(const sq_space* sp, const sq_op_string* strings, int n_strings, int drop_outside,
                       sq_operator** out) {
  return guarded([&] {
    SQ_CHECK(sp && out && (strings || n_strings == 0) && n_strings >= 0, SQ_ERR_ARG, "invalid arguments");
    auto* op = new sq_operator();
    op->space = &sp->s;
    try {
      for (int i = 0; i < n_strings; ++i) {
        const sq_op_string& st = strings[i];
        SQ_CHECK(st.n_ops >= 0 && st.n_ops <= 24, SQ_ERR_ARG, "operator string longer than 24");
        std::vector<RawOp> ops;
        for (int j = 0; j < st.n_ops; ++j) {
          SQ_CHECK(st.spin_orb[j] >= 0 && st.spin_orb[j] < 2 * sp->s.n, SQ_ERR_SPACE,
                   "spin orbital outside the active space");
          ops.push_back({st.spin_orb[j], st.is_creator[j] != 0});
        }
        try {
          op->terms.push_back(make_term(ops, st.coef));
        } catch (const Error& e) {
          if (e.code == SQ_ERR_SPACE && drop_outside) continue;
          throw;
        }
      }
      SQ_HIP(hipMalloc(&op->d_terms, sizeof(Term) * std::max<size_t>(op->terms.size(), 1)));
      if (!op->terms.empty())
        SQ_HIP(hipMemcpy(op->d_terms, op->terms.data(), sizeof(Term) * op->terms.size(), hipMemcpyHostToDevice));
    } catch (...) {
      sq_operator_destroy(op);
      throw;
    }
    *out = op;
  });
}

int sq_operator_destroy(sq_operator* op) {
  return guarded([&] {
    if (!op) return;
    if (op->d_terms) (void)hipFree(op->d_terms);
    operator_pattern_free(*op);
    delete op;
  });
}

int sq_operator_apply(const sq_space* sp, const sq_operator* op, const double* in, double* out, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && op && in && out && in != out, SQ_ERR_ARG, "invalid arguments (in and out must differ)");
    SQ_CHECK(op->space == &sp->s, SQ_ERR_ARG, "operator built for another space");
    const Space& s = sp->s;
    if (s.dim == 0) return;
    // repeated applications (linear response applies the same operators many
    // times) go through the operator's sparsity pattern
    sq_operator& o = const_cast<sq_operator&>(*op);
    if (++o.applies >= 2 && operator_apply_pattern(s, o, in, out, as_stream(stream))) return;
    k_tapply<<<grid2d(s), kBlk, 0, as_stream(stream)>>>(view(s), op->d_terms, 0, (int)op->terms.size(), in, out,
                                                        nullptr, nullptr, 0, 0, 0, 0, 1.0);
    SQ_HIP(hipGetLastError());
  });
}

int sq_operator_matrix(const sq_space* sp, const sq_operator* op, int reference_convention, double* M,
                       void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && op && M, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(op->space == &sp->s, SQ_ERR_ARG, "operator built for another space");
    const Space& s = sp->s;
    if (s.dim == 0) return;
    hipStream_t st = as_stream(stream);
    const size_t n = (size_t)s.dim * (size_t)s.dim;
    k_fill_zero<<<(unsigned)std::min<size_t>((n + kBlk - 1) / kBlk, 65536), kBlk, 0, st>>>(M, n);
    SQ_HIP(hipGetLastError());
    k_operator_matrix<<<grid2d(s), kBlk, 0, st>>>(view(s), op->d_terms, 0, (int)op->terms.size(),
                                                  reference_convention ? 1 : 0, M);
    SQ_HIP(hipGetLastError());
  });
}

int sq_state_prep(const sq_space* sp, const sq_program* pr, const double* theta, const double* psi_in,
                  double* psi_out, int dagger, double* work, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && pr && psi_in && psi_out, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(pr->p.space == &sp->s, SQ_ERR_ARG, "program built for another space");
    const Space& s = sp->s;
    const Program& p = pr->p;
    SQ_CHECK(theta || p.gates.empty(), SQ_ERR_ARG, "theta is NULL");
    hipStream_t st = as_stream(stream);
    const TilePlan* T = tiles_for(s, p);
    if (T) {
      tiles_forward(s, p, *T, theta, psi_in, psi_out, work + 2 * s.dim, work, -1, dagger != 0, st);
      return;
    }
    copy_vec(psi_in, psi_out, s.dim, st);
    int n = (int)p.gates.size();
    for (int k = 0; k < n; ++k) {
      int g = dagger ? n - 1 - k : k;
      apply_gate_impl(s, p, g, theta, dagger != 0, psi_out, work, st);
    }
  });
}

int sq_apply_gate(const sq_space* sp, const sq_program* pr, int gate, const double* theta, double* psi, double* work,
                  void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && pr && psi && theta, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(pr->p.space == &sp->s, SQ_ERR_ARG, "program built for another space");
    SQ_CHECK(gate >= 0 && gate < (int)pr->p.gates.size(), SQ_ERR_ARG, "gate index out of range");
    apply_gate_impl(sp->s, pr->p, gate, theta, false, psi, work, as_stream(stream));
  });
}

int sq_grad_action(const sq_space* sp, const sq_program* pr, int gate, const double* in, double* out, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && pr && in && out && in != out, SQ_ERR_ARG, "invalid arguments (in and out must differ)");
    SQ_CHECK(pr->p.space == &sp->s, SQ_ERR_ARG, "program built for another space");
    SQ_CHECK(gate >= 0 && gate < (int)pr->p.gates.size(), SQ_ERR_ARG, "gate index out of range");
    grad_action_impl(sp->s, pr->p, gate, in, out, as_stream(stream));
  });
}

}  // extern "C"
