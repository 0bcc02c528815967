/*
 * sqhip.h -- C ABI of the MI355X state-vector engine for SlowQuant's
 * unitary product-state (UPS / tUPS / fUCC) and UCC wave functions.
 *
 * The reference exposes this path as a Python API, not an FFI; each entry
 * point below names the reference function(s) it replaces.  All functions:
 *   - are extern "C", take plain pointers and sizes (no torch types);
 *   - return 0 on success and a negative SQ_ERR_* code otherwise, with a
 *     message available from sq_last_error() (thread-local);
 *   - never throw;
 *   - take device pointers for state vectors (FP64, length sq_space_dim())
 *     owned by the caller, and a hipStream_t (passed as void*) on which all
 *     work is enqueued.  Library-owned tables live in the sq_space / sq_program
 *     objects.
 *
 * State vectors handed to the engine are in the engine's internal sign
 * convention (determinants ordered alpha block then beta block).  Use
 * sq_state_from_reference / sq_state_to_reference to convert from/to the
 * reference's interleaved convention (ci_spaces.py:101-104); the index order
 * (alpha string outer, beta string inner, lexicographic) is the reference's.
 */
#ifndef SQHIP_H
#define SQHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SQ_OK = 0,
  SQ_ERR_ARG = -1,      /* invalid argument (maps to ValueError) */
  SQ_ERR_HIP = -2,      /* HIP runtime error (maps to RuntimeError) */
  SQ_ERR_SPACE = -3,    /* operator leaves the CI space (maps to KeyError) */
  SQ_ERR_TYPE = -4,     /* unknown excitation type (maps to ValueError) */
  SQ_ERR_NOMEM = -5
};

/* Gate kinds: the excitation types of UpsStructure / UccStructure (util.py:547-1073). */
enum {
  SQ_GATE_SA_SINGLE = 1,  /* (i,a) spatial  */
  SQ_GATE_SINGLE = 2,     /* (i,a) spin orbitals */
  SQ_GATE_DOUBLE = 3,     /* (i,j,a,b) spin orbitals */
  SQ_GATE_TRIPLE = 4,
  SQ_GATE_QUADRUPLE = 5,
  SQ_GATE_QUINTUPLE = 6,
  SQ_GATE_SEXTUPLE = 7,
  SQ_GATE_SA_DOUBLE_1 = 11, /* (i,j,a,b) spatial, cases 1..5 */
  SQ_GATE_SA_DOUBLE_2 = 12,
  SQ_GATE_SA_DOUBLE_3 = 13,
  SQ_GATE_SA_DOUBLE_4 = 14,
  SQ_GATE_SA_DOUBLE_5 = 15,
  SQ_GATE_SA_SINGLE_UCC = 21 /* G1_sa generator (2^-1/2 E_ai - h.c.), UCC only */
};

typedef struct sq_space sq_space;
typedef struct sq_program sq_program;

typedef struct {
  int32_t kind;      /* SQ_GATE_* */
  int32_t n_idx;     /* number of valid entries in idx */
  int32_t idx[12];   /* indices exactly as in the layout tuple */
} sq_gate_desc;

/* Message of the last failing call on this thread. */
const char* sq_last_error(void);
/* Library build identification string. */
const char* sq_version(void);

/* ---- CI space: ci_spaces.get_indexing (ci_spaces.py:76-116) --------------
 * Builds, on the current device, the lexicographic alpha/beta string tables,
 * closed-form rank tables (no hash map) and the E_pq single-excitation tables
 * with their signs. */
int sq_space_create(int n_orb, int n_alpha, int n_beta, sq_space** out);
int sq_space_destroy(sq_space* s);
int64_t sq_space_dim(const sq_space* s);
int sq_space_nstrings(const sq_space* s, int32_t* n_alpha_str, int32_t* n_beta_str);
/* Copy string bit masks (bit p = spatial orbital p) to host arrays. */
int sq_space_copy_strings(const sq_space* s, uint32_t* h_alpha, uint32_t* h_beta);
/* Copy the reference-convention determinant integers idx2det (ci_spaces.py:99-107)
 * generated on the device from the string tables. */
int sq_space_idx2det(const sq_space* s, int64_t* h_idx2det, void* stream);
/* Sign conversion between the reference's interleaved convention and the
 * engine's blocked convention (in-place allowed). */
int sq_state_from_reference(const sq_space* s, const double* ref, double* internal, void* stream);
int sq_state_to_reference(const sq_space* s, const double* internal, double* ref, void* stream);

/* ---- Gate programs: the layout of UpsStructure / UccStructure ------------ */
int sq_program_create(const sq_space* s, const sq_gate_desc* gates, int n_gates, sq_program** out);
int sq_program_destroy(sq_program* p);
int sq_program_ngates(const sq_program* p);
/* Tile shape of the fused rotation stages: elements per tile (0: automatic,
 * 256..2048 by the space's dimension; at most 4096) and threads per tile block
 * (256 or 512; 0: 512).  No counterpart in the reference (its gates have no
 * tiles).  Synchronises the device and drops the program's tile plan, which
 * the next call rebuilds. */
int sq_program_set_tile_shape(sq_program* p, int target, int threads);

/* construct_ups_state (operator_state_algebra.py:963-1412): psi_out = U(theta) psi_in,
 * or U(theta)^dagger psi_in when dagger != 0.  theta is a DEVICE array of
 * n_gates doubles.  psi_in may equal psi_out.  work: sq_work_bytes(s, 0). */
int sq_state_prep(const sq_space* s, const sq_program* p, const double* theta, const double* psi_in,
                  double* psi_out, int dagger, double* work, void* stream);
/* propagate_unitary (operator_state_algebra.py:1867-2309): apply gate `gate` only. */
int sq_apply_gate(const sq_space* s, const sq_program* p, int gate, const double* theta, double* psi,
                  double* work, void* stream);
/* get_grad_action (operator_state_algebra.py:2757-2865): out = T_gate psi_in. */
int sq_grad_action(const sq_space* s, const sq_program* p, int gate, const double* psi_in, double* out,
                   void* stream);

/* ---- Hamiltonian ----------------------------------------------------------
 * H = e_core + sum_pq h_pq E_pq + 1/2 sum_pqrs g_pqrs (E_pq E_rs - delta_qr E_ps)
 * on the active space (hamiltonian_0i_0a + get_folded_operator,
 * operators.py:476-529, fermionic_operator.py:379-471).  h: host n*n, g: host
 * n^4 chemist-notation, both must have the 8-fold real symmetry. */
typedef struct sq_hamiltonian sq_hamiltonian;
int sq_hamiltonian_create(const sq_space* s, const double* h, const double* g, double e_core,
                          sq_hamiltonian** out);
int sq_hamiltonian_destroy(sq_hamiltonian* H);
/* Bytes of device workspace needed by sq_sigma / gradients / RDMs. */
int64_t sq_work_bytes(const sq_space* s, int what);

/* sigma = H psi and, if energy != NULL, energy = <psi|H|psi> (host double)
 * (propagate_state([H]) + expectation_value, operator_state_algebra.py:472-824).
 * Spaces of up to 4096 beta strings: Ma C (the alpha-string Hamiltonian times
 * psi as a string matrix, FP64 MFMA) and one row kernel per group of alpha
 * rows that contracts g with E_pq psi on FP64 MFMA and applies the beta
 * excitations with the intermediate kept in LDS.  Larger spaces: D build
 * (E_pq psi gathers) -> G = g.D (FP64 MFMA) -> sigma gather. */
int sq_sigma(const sq_space* s, const sq_hamiltonian* H, const double* psi, double* sigma, double* energy,
             double* work, void* stream);

/* Kernel shape of sq_sigma's row path on space s; no counterpart in the
 * reference (its sigma has no kernel shapes).  rows_per_block: alpha rows per
 * block (0: automatic); nsplit: partial products of the alpha-string GEMM
 * (1..4, 0: automatic); deterministic: 1 selects the staged form whose sums
 * run in a fixed order (bit-identical results run to run), 0 the scatter form
 * (LDS atomics: equal to rounding, faster), -1 the SQ_SIGMA_DETERMINISTIC
 * environment default (scatter).  Takes effect at the next call; workspace
 * sizes do not change. */
int sq_sigma_configure(sq_space* s, int rows_per_block, int nsplit, int deterministic);

/* sigma = H psi with no allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  Same kernels as sq_sigma's row path, with its scratch (alpha GEMM
 * partials and block partials) in a 64 MiB device array that each GPU the
 * library runs on holds until the process ends; spaces needing more, or
 * above 4096 beta strings, fail with SQ_ERR_ARG.  energy: DEVICE double or
 * NULL; nothing is synchronised.  Calls that overlap on one GPU share the
 * array and race. */
int sq_sigma_noalloc(const sq_space* s, const sq_hamiltonian* H, const double* psi, double* sigma, double* energy,
                     void* stream);

/* Energy and full theta gradient of a UPS state at theta (device array):
 * ups_wavefunction._calc_energy_optimization + _calc_gradient_optimization
 * (ups_wavefunction.py:1019-1142) for theta parameters.  grad: host array of
 * n_gates doubles, energy: host double.  psi0: reference state (device,
 * internal convention).  Adjoint = one forward sweep, one sigma, one reverse
 * sweep that un-applies the gates on (psi, lambda) in place and fuses
 * g_i = 2 <lambda|T_i psi>. */
int sq_energy_grad_adjoint(const sq_space* s, const sq_program* p, const sq_hamiltonian* H,
                           const double* theta, const double* psi0, double* energy, double* grad,
                           double* work, void* stream);
/* Same quantities by per-parameter derivative propagation (O(N^2) gate
 * applications): d_i psi = U_N..U_{i+1} U_i T_i U_{i-1}..U_1 psi0. */
int sq_energy_grad_plain(const sq_space* s, const sq_program* p, const sq_hamiltonian* H,
                         const double* theta, const double* psi0, double* energy, double* grad,
                         double* work, void* stream);
/* The reference's own product-rule sweep (ups_wavefunction.py:1090-1141). */
int sq_energy_grad_reference_sweep(const sq_space* s, const sq_program* p, const sq_hamiltonian* H,
                                   const double* theta, const double* psi0, double* energy, double* grad,
                                   double* work, void* stream);

/* Reverse half of the adjoint gradient, for a torch.autograd.Function whose
 * forward ran sq_state_prep (psi = U(theta) psi0) and sq_sigma (lam = H psi):
 * one reverse sweep un-applies every gate on psi and lam in place (both are
 * consumed) and writes dE/dtheta_i = 2 <lam_i|T_i psi_i> to grad (host).
 * Replaces the theta block of ups_wavefunction.py:1060-1142.
 * work: sq_work_bytes(s, 2). */
int sq_adjoint_reverse(const sq_space* s, const sq_program* p, const double* theta, double* psi, double* lam,
                       double* grad, double* work, void* stream);

/* ---- RDMs (ups_wavefunction.py:409-476, density_matrix.py:5-136) ---------
 * rdm1: host n*n, rdm2: host n^4 (p,q,r,s) = <E_pq E_rs> - delta_qr rdm1_ps. */
int sq_rdm12(const sq_space* s, const double* psi, double* rdm1, double* rdm2, double* work, void* stream);

/* ---- Unfactorised UCC (construct_ucc_state, operator_state_algebra.py:870-960;
 * ucc_wavefunction.py:966-1098).  The program holds the UCC excitation list;
 * its generators T_k (G1_sa for "sa_single_ucc", G2_sa, G1..G6 anti-Hermitian)
 * form T(theta) = sum_k theta_k T_k.  work: sq_ucc_work_bytes(). */
int64_t sq_ucc_work_bytes(const sq_space* s, const sq_program* p);
/* out = exp(t T(theta)) in, matrix-free scaled Taylor series (in != out). */
int sq_ucc_expm(const sq_space* s, const sq_program* p, const double* theta, double t, const double* in,
                double* out, void* work, void* stream);
/* energy = <psi|H|psi>, psi = exp(T(theta)) psi0, and (if grad != NULL) the
 * analytic dE/dtheta_i = 2 int_0^1 <e^{-sT} lam|T_i|e^{-sT} psi> ds (lam = H psi),
 * by composite 8-point Gauss-Legendre quadrature over s. */
int sq_ucc_energy_grad(const sq_space* s, const sq_program* p, const sq_hamiltonian* H, const double* theta,
                       const double* psi0, double* energy, double* grad, void* work, void* stream);

/* ---- Batched instances (the reference batches equal-dimension states
 * column-wise in its _SA family: apply_operator_SA_threaded,
 * operator_state_algebra.py:352-410; construct_ups_state_SA, :1415-1866;
 * propagate_unitary_SA, :2312-2756; expectation_value_SA, :827-868).
 * Here each instance has its own space, program and Hamiltonian; one
 * persistent launch computes every instance's energy and adjoint gradient
 * from the HF determinant.  Programs may hold rotation gates only (single,
 * double, sa_single, pair doubles: tUPS / QNP / fUCC-SD layouts); spaces up
 * to 12 orbitals.  n_workgroups <= 0 picks 256.  Instances of 16384 or more
 * determinants whose programs are all brick stages run instead in groups
 * of equal (space, gate layout) through the tile-stage, row-sigma and
 * gradient kernels with one grid row per instance (SQ_BATCH_GROUPS=0 keeps
 * them in the persistent launch).  Such a batch uses the programs' tile
 * plans at every launch: the programs must outlive it, and changing a
 * grouped program's tile shape (sq_program_set_tile_shape) or a grouped
 * space's row-sigma shape (sq_sigma_configure) afterwards makes its launches
 * fail with SQ_ERR_ARG (create the batch again). */
typedef struct sq_batch sq_batch;
int sq_batch_create(int n_inst, const sq_space* const* spaces, const sq_program* const* programs,
                    const sq_hamiltonian* const* hams, int n_workgroups, sq_batch** out);
int sq_batch_destroy(sq_batch* b);
/* Total number of parameters (sum of the programs' gate counts). */
int64_t sq_batch_num_params(const sq_batch* b);
/* theta: host array, instances' parameters concatenated in instance order. */
int sq_batch_set_theta(sq_batch* b, const double* theta, void* stream);
/* Start state of every instance (construct_ups_state_SA from arbitrary
 * reference states, operator_state_algebra.py:1415-1864, as the SA-UPS
 * RotoSolve batch uses it, sa_ups_wavefunction.py:1005-1082): psi0 is a host
 * array of n_inst device pointers, each to its instance's state in the engine
 * convention (space dim doubles), or NULL for the HF determinant (the
 * default).  Read at every launch and graph replay: the states must stay
 * allocated while the batch runs. */
int sq_batch_set_start_states(sq_batch* b, const double* const* psi0, void* stream);
/* Kernel shape picked by the autotuner: rotation pairs of determinants per
 * thread in flight (1, 2 or 4) and sigma GEMM 16-column tiles per wave (1 or
 * 2); no counterpart in the reference (its batching has no kernel shapes).
 * Drops a captured graph. */
int sq_batch_configure(sq_batch* b, int rot_pairs, int gemm_tiles);
/* One launch on stream. */
int sq_batch_run(sq_batch* b, void* stream);
/* Same launch captured once into a hipGraph, then replayed. */
int sq_batch_run_graph(sq_batch* b, void* stream);
/* Host copies: energies[n_inst] (electronic + core), grads[num_params]. */
int sq_batch_results(const sq_batch* b, double* energies, double* grads, void* stream);

/* ---- Generic operators: propagate_state for FermionicOperator lists ------
 * (operator_state_algebra.py:472-630, scan kernels :139-219).  A string is a
 * raw product of creation/annihilation operators, ops[0] leftmost, with
 * interleaved spin-orbital indices of the active space (2p alpha, 2p+1 beta)
 * and a coefficient.  A string that changes N_alpha or N_beta leaves the CI
 * space: creation fails with SQ_ERR_SPACE unless drop_outside != 0, in which
 * case it is skipped (the reference's do_unsafe). */
typedef struct {
  double coef;
  int32_t n_ops;
  int32_t spin_orb[24];
  int32_t is_creator[24];
} sq_op_string;
typedef struct sq_operator sq_operator;
int sq_operator_create(const sq_space* s, const sq_op_string* strings, int n_strings, int drop_outside,
                       sq_operator** out);
int sq_operator_destroy(sq_operator* op);
/* out = O in (device vectors, internal convention, in != out).  The first
 * application scans the terms per determinant; the second builds the
 * operator's sparsity pattern (kept with the operator, up to 2^25 entries),
 * which serves that and every later application. */
int sq_operator_apply(const sq_space* s, const sq_operator* op, const double* in, double* out, void* stream);
/* build_operator_matrix (operator_state_algebra.py:413-469): the dense
 * dim x dim matrix M[I][J] = <I|O|J> (device, row-major) in one launch, in the
 * reference's sign convention when reference_convention != 0. */
int sq_operator_matrix(const sq_space* s, const sq_operator* op, int reference_convention, double* M,
                       void* stream);

/* ---- Orbital rotations and integrals ---------------------------------------
 * Matrices are row-major FP64 device arrays; every contraction is an FP64-MFMA
 * GEMM.  work: sq_orbital_work_bytes(n_ao, n_mo, n_inactive, n_active, n_kappa)
 * bytes (the largest of the three calls below for those sizes). */
int64_t sq_orbital_work_bytes(int n_ao, int n_mo, int n_inactive, int n_active, int n_kappa);
/* c_mo = C_old expm(-kappa) (ups_wavefunction.py:367-385): c_old and c_new
 * are n_ao x n device arrays, kappa the n x n antisymmetric HOST matrix of the
 * rotations moved since the last expansion point.  expm by scaling and
 * squaring of a degree-14 Taylor polynomial. */
int sq_orbital_rotation(int n, int n_ao, const double* c_old, const double* kappa, double* c_new, double* work,
                        void* stream);
/* one_electron_integral_transform / two_electron_integral_transform
 * (integralfunctions.py:186-211): h_mo = C^T h_ao C (n_mo x n_mo) and
 * g_mo[ijkl] = sum C_ai C_bj C_ck C_dl g_ao[abcd] (four quarter transforms);
 * either pair may be NULL.  n_mo <= 255. */
int sq_ao2mo(int n_ao, int n_mo, const double* c, const double* h_ao, const double* g_ao, double* h_mo,
             double* g_mo, void* work, void* stream);
/* get_orbital_gradient (density_matrix.py:181-230): <0|[kappa_mn, H]|0> for
 * each (m, n) of kappa_idx (host, 2 * n_kappa ints) from the active-space
 * rdm1 (n_active^2) and rdm2 (n_active^4) and the MO integrals (device);
 * grad: host, n_kappa doubles. */
int sq_orbital_gradient(int n_orb, int n_inactive, int n_active, const double* h_mo, const double* g_mo,
                        const double* rdm1, const double* rdm2, int n_kappa, const int32_t* kappa_idx, double* grad,
                        void* work, void* stream);

/* ---- Kernel timing -------------------------------------------------------
 * No counterpart in the reference (its timing is wall-clock prints).  While
 * enabled, every engine kernel launch is bracketed by HIP events on its
 * stream; collect returns, per kernel class, the summed duration (ms), the
 * launch count and the algorithmic bytes and flops of those launches. */
int sq_profile_enable(int on);
/* Kernel classes timed while enabled: bit c of class_mask for class c
 * (default all).  Timing one class adds one event pair per launch of it. */
int sq_profile_select(uint32_t class_mask);
int sq_profile_nclasses(void);
const char* sq_profile_class_name(int cls);
int sq_profile_collect(int n_classes, double* total_ms, int64_t* count, double* bytes, double* flops);

/* Test hook for the sigma contraction kernel: G[m][c] = sum_k At[k][m] D[k][c]
 * for m < npair, k < K4 (device arrays; At is K4 x kpad, D is K4 x dim,
 * G is npair x dim).  No counterpart in the reference: it exposes the
 * FP64-MFMA kernel of sq_sigma for numerical checks. */
int sq_debug_gemm(const double* At, const double* D, double* G, int npair, int K4, int kpad, int64_t dim,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SQHIP_H */
