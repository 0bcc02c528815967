// Internal definitions shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sqhip.h"

namespace sq {

void set_error(const std::string& msg);

struct Error {
  int code;
  std::string msg;
};

#define SQ_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      throw ::sq::Error{SQ_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};      \
    }                                                                                       \
  } while (0)

#define SQ_CHECK(cond, code, msg)                    \
  do {                                               \
    if (!(cond)) throw ::sq::Error{(code), (msg)};   \
  } while (0)

template <class F>
int guarded(F&& f) {
  try {
    f();
    return SQ_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return SQ_ERR_ARG;
  }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Maximum active orbitals: rank tables are direct-indexed by the 2^n bit mask.
constexpr int kMaxOrb = 16;

// One spin-resolved raw operator product: ops applied right to left
// (ops[n-1] first).  orb = spatial orbital, cre = creation flag.
struct SpinOps {
  int32_t n;
  int8_t orb[12];
  int8_t cre[12];
};

// A raw product term of a generator in the blocked convention:
// coef * (alpha ops) (beta ops).
struct Term {
  double coef;
  SpinOps a;
  SpinOps b;
};

// A term's adjoint on a determinant in mask form, for terms whose spin parts
// touch each orbital at most once (flags bit 0): it applies iff the string
// has every creator orbital (s) and no annihilator orbital (c); the image is
// string ^ s ^ c and the sign parity is popc(string & x) + flags bit 1,
// summed over both spin parts.  Other terms (flags bit 0 clear) are applied
// operator by operator.
struct TermFast {
  uint32_t sa, ca, xa, sb, cb, xb;
  uint32_t flags, pad;
  double coef;
};

// Device-side gate record.
struct GateDev {
  int32_t kind;          // internal class, see GateClass
  int32_t term_begin;    // generator terms [term_begin, term_end)
  int32_t term_end;
  int32_t poly;          // for POLY gates: index of the frequency table (2..5)
  double scale;          // angle multiplier for rotation gates
  double pairs[2];       // rotation pairs touched by each rotation half (for byte counts)
};

enum GateClass : int32_t {
  GC_ROT = 1,        // single string G with distinct indices: exp(th (G - G^+)) = pair rotation
  GC_SA_SINGLE = 2,  // two commuting rotations (alpha, beta) with the same angle
  GC_POLY = 3,       // multi-frequency polynomial in T (sa_double_2..5)
  GC_TAYLOR = 4      // generic: closed form 1 + sin T + (1 - cos) T^2 through T applications
};

struct Space {
  int n, na, nb;
  int nsa, nsb;
  int64_t dim;
  int npair;              // n(n+1)/2 packed (p>=q) pairs
  int device;
  uint32_t* d_astr = nullptr;  // [nsa]
  uint32_t* d_bstr = nullptr;  // [nsb]
  int32_t* d_arank = nullptr;  // [1<<n]
  int32_t* d_brank = nullptr;
  int32_t* d_aexc = nullptr;   // [npair][nsa]: sign*(target+1) for E+_k, p>q; occupation for p==q
  int32_t* d_bexc = nullptr;   // [npair][nsb]
  int32_t* d_aexc_full = nullptr;  // [n*n][nsa]: sign*(target+1) for E_pq|s> (x = p*n+q)
  int32_t* d_bexc_full = nullptr;
  int2* d_gta = nullptr;  // sigma GEMM operand tables [K4 + 16][nsa], [K4 + 16][nsb] (sq_sigma.hip)
  int2* d_gtb = nullptr;
  // row sigma tables (sq_rowsigma.hip): operand rows [K8 + 16][nsa], [K8 + 32][nsb],
  // beta neighbour lists by partner column with per-tile offsets
  int2* d_rta = nullptr;
  int2* d_rtb = nullptr;
  uint32_t* d_nbr = nullptr;
  int32_t* d_nbr_off = nullptr;
  uint8_t* d_nbr_tile = nullptr;
  // row sigma shape (sq_sigma_configure): 0 / -1 = automatic
  int sig_R = 0, sig_nsplit = 0, sig_det = -1;
  std::vector<uint32_t> h_astr, h_bstr;
};

void build_gemm_tables(Space& s);
void build_row_tables(Space& s);
void free_row_tables(Space& s);

struct TilePlan;
TilePlan* build_tile_plan(const Space& sp, const struct Program& P);
void destroy_tile_plan(TilePlan* T);
bool tile_plan_has_tiles(const TilePlan* T);

struct Program {
  const Space* space;
  mutable TilePlan* tiles = nullptr;  // built on first use (sq_tiles.hip)
  mutable bool tiles_built = false;
  int tile_target = 0;   // elements per tile (0: automatic; sq_program_set_tile_shape)
  int tile_threads = 0;  // threads per tile block, 256 or 512 (0: 512)
  int plan_gen = 0;      // bumped whenever the tile plan is dropped (batched groups check it)
  std::vector<GateDev> gates;
  std::vector<Term> terms;
  GateDev* d_gates = nullptr;
  Term* d_terms = nullptr;
  std::vector<int> pair_p, pair_q;  // orbital pair (p, p+1) of "brick" gates, -1 otherwise
  // Rotation gates: per gate and half (2g + h) the spin-part lists
  // {alpha offset, alpha length, beta offset, beta length} into d_lists
  // (offset -1: the spin part is the identity, every string).
  std::vector<int4> rot_lists;
  int2* d_lists = nullptr;
  // Per term, its adjoint in mask form (built on first use by the UCC
  // kernels, sq_ucc.hip).
  mutable struct TermFast* d_term_fast = nullptr;
  // T's sparsity pattern as rows of (source determinant, term | sign) entries
  // (built on first use by the UCC kernels, sq_ucc.hip; freed by
  // ucc_csr_free).
  mutable struct UccCsr* ucc_csr = nullptr;
  mutable bool ucc_csr_tried = false;
};

void ucc_csr_free(struct UccCsr* c);

// Strings a spin part applies to, as (string index, image | parity << 31)
// entries appended to lists (shared per (strings, operator) through cache);
// off = -1 and len = all strings for an identity part.
void spin_part_list(const std::vector<uint32_t>& strs, const SpinOps& o, std::vector<int2>& lists,
                    std::map<std::pair<const void*, std::vector<int>>, std::pair<int, int>>& cache, int& off,
                    int& len);

struct Hamiltonian {
  const Space* space;
  int npair;
  int kpad;            // padded K (= M) of the packed GEMM, multiple of 16
  double e_core;
  double* d_A = nullptr;  // [kpad][kpad] transposed coefficient matrix (see sq_sigma.hip)
  double* d_Ma = nullptr;  // [nsa][nsa] alpha-string Hamiltonian + e_core (sq_rowsigma.hip)
};

// Row sigma (sq_rowsigma.hip)
bool row_sigma_applies(const Space& s);
int64_t row_sigma_work_doubles(const Space& s);
int64_t row_sigma_used_doubles(const Space& s);
void build_alpha_ham(const Space& s, Hamiltonian& H);
// allow_library = false keeps every step in this library's own kernels (no
// library call that may allocate).
int64_t row_sigma(const Space& s, const Hamiltonian& H, const double* psi, double* sigma, double* work,
                  double** partial_out, hipStream_t st, bool allow_library = true);

// Kernel timing classes (sq_profile.hip).
enum ProfClass {
  PC_ROTATE = 0,
  PC_ROTATE_ADJ,
  PC_TAPPLY,
  PC_DBUILD,
  PC_GEMM,
  PC_SIGMA_GATHER,
  PC_REDUCE,
  PC_COPY,
  PC_TILE_FWD,
  PC_TILE_REV,
  PC_PERM,
  PC_RDM_GRAM,
  PC_BATCH,
  PC_OTHER,
  PC_SIGMA_ROWS,
  PC_ALPHA_GEMM,
  PC_ORBITAL,
  kProfClasses
};
bool prof_on();
bool prof_pause();
void prof_resume(bool was);
void prof_begin(int cls, hipStream_t st);
void prof_end(int cls, hipStream_t st, double bytes, double flops);

#define SQ_PROF(cls, st, bytes, flops, ...)          \
  do {                                               \
    ::sq::prof_begin((cls), (st));                   \
    __VA_ARGS__;                                     \
    SQ_HIP(hipGetLastError());                       \
    ::sq::prof_end((cls), (st), (bytes), (flops));   \
  } while (0)

// Workspace layout helpers (bytes).
int64_t sigma_work_bytes(const Space& s);

// Implemented in sq_sigma.hip
void sigma_impl(const Space& s, const Hamiltonian& H, const double* psi, double* sigma, double* energy_host,
                double* work, hipStream_t st);
void sigma_energy_dev(const Space& s, const Hamiltonian& H, const double* psi, double* sigma, double* e_dev,
                      double* work, hipStream_t st);
void dot_dev(const double* x, const double* y, int64_t n, double* out_dev, int slot, double factor, double* partial,
             hipStream_t st);
int64_t rdm_work_bytes(const Space& s, int* rt_out, int* nchunks_out, int64_t* chunk_out);
void rdm_impl(const Space& s, const double* psi, double* rdm1, double* rdm2, double* work, hipStream_t st);

// Implemented in sq_gates.hip
void apply_gate_impl(const Space& s, const Program& p, int g, const double* theta_dev, bool negate, double* psi,
                     double* work, hipStream_t st);
void grad_action_impl(const Space& s, const Program& p, int g, const double* in, double* out, hipStream_t st);
void copy_vec(const double* in, double* out, int64_t n, hipStream_t st);
int64_t nblocks2d(const Space& s);
int64_t tdot_impl(const Space& s, const Program& p, int g, const double* psi, const double* lam, double* partial,
                  hipStream_t st);
void reduce_impl(const double* partial, int64_t n, double* out_dev, int slot, bool accumulate, double factor,
                 hipStream_t st);
void reverse_step_impl(const Space& s, const Program& p, int g, const double* theta_dev, double* psi, double* lam,
                       double* grad_dev, double* work, double* partial, hipStream_t st);

// Implemented in sq_tiles.hip
const TilePlan* tiles_for(const Space& s, const Program& p);  // nullptr when no tiled stage applies

// Several instances of one (space, program) evaluated together (batched
// groups, sq_batch.hip): the tile, gradient and row-sigma kernels run with
// blockIdx.y = instance and add y * stride to each per-instance buffer;
// n = 1 with zero strides is one instance.
struct Multi {
  int n = 1;
  int64_t theta = 0, mats = 0, vec = 0, part = 0, grad = 0;
};
// Group path of a program: every stage tiled, no side launches.
bool tiles_group_ok(const Space& s, const Program& p);
int64_t tiles_mats_doubles(const Space& s, const Program& p);     // pair matrices per instance
int64_t tiles_partial_doubles(const Space& s, const Program& p);  // gradient partials per instance
// psi (in place, start states already there) = U(theta) psi for every instance
void tiles_forward_group(const Space& s, const Program& p, const double* theta, double* psi, double* mats,
                         const Multi& M, hipStream_t st);
// reverse sweep on (psi, lam) with dE/dtheta into grad (n x gates)
void tiles_reverse_group(const Space& s, const Program& p, const double* theta, double* psi, double* lam,
                         const double* mats, double* partial, double* grad, const Multi& M, hipStream_t st);
// sigma = H psi for n instances of one space (row form with its alpha GEMM
// tiles in the launch); A and Ma per instance at the given strides; energy
// partials per instance (row_sigma_group_partials of them) at partial.
bool row_sigma_group_ok(const Space& s);
int64_t row_sigma_group_partials(const Space& s);
void row_sigma_group(const Space& s, const double* A, int64_t A_stride, const double* Ma, int64_t Ma_stride,
                     const double* psi, double* sigma, double* partial, int n, hipStream_t st);
void tiles_forward(const Space& s, const Program& p, const TilePlan& T, const double* theta, const double* psi_in,
                   double* psi, double* scratch, double* gwork, int deriv_gate, bool dagger, hipStream_t st);
void tiles_reverse(const Space& s, const Program& p, const TilePlan& T, const double* theta, double* psi, double* lam,
                   double* gwork, double* partial, double* grad_dev, bool mats_ready, hipStream_t st);

// Dynamic LDS above 64 KB must be enabled per kernel; the request covers
// exactly what a launch needs (raised only, kept per kernel).
template <class K>
inline void allow_lds(K kernel, size_t bytes) {
  static std::mutex mu;
  static std::map<const void*, size_t> allowed;
  const void* key = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lock(mu);
  size_t& a = allowed[key];
  if (a == 0) a = 64 * 1024;
  if (bytes > a) {
    SQ_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    a = bytes;
  }
}

}  // namespace sq

struct sq_space {
  sq::Space s;
};
struct sq_program {
  sq::Program p;
};
struct sq_hamiltonian {
  sq::Hamiltonian h;
};
struct sq_operator {
  const sq::Space* space = nullptr;
  std::vector<sq::Term> terms;
  sq::Term* d_terms = nullptr;
  // sparsity pattern of the operator on its space, built at the second
  // application (an operator applied once keeps the term scan)
  int applies = 0;
  bool csr_tried = false;
  struct sq::UccCsr* csr = nullptr;
  sq::TermFast* d_fast = nullptr;
};

namespace sq {
// out = O in through the operator's sparsity pattern; false (nothing
// launched) when the pattern is not built (first application, too large, or
// SQ_UCC_CSR=0).
bool operator_apply_pattern(const Space& s, sq_operator& op, const double* in, double* out, hipStream_t st);
void operator_pattern_free(sq_operator& op);
}  // namespace sq
