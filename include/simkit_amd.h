/*
 * simkit_amd — C ABI of the MI355X (gfx950) implicit-FEM backward-Euler path.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (hipMalloc'd or torch tensors'
 *     data_ptr()) unless the name ends in _host. Row-major, C-contiguous.
 *   - Indices are int32, reals FP64 unless stated.
 *   - x / U vectors are flattened vertex positions (x[3*v + i]).
 *   - F is row-major per element (F[9*e + 3*i + c], i spatial, c material);
 *     9x9 Hessians index row-major vec(F).
 *   - Every call is ordered on the given stream (NULL = the null stream).
 *     Calls that return host scalars synchronise that stream.
 *   - Return value: SK_OK (0) or a negative sk_status; sk_last_error() gives
 *     the calling thread's last message.
 *
 * The reference (simkit v0.1.4) is pure Python/numpy/scipy; each entry point
 * names the reference function it replaces as file:line.
 */
#ifndef SIMKIT_AMD_H
#define SIMKIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SK_VERSION 1

typedef void* sk_stream_t; /* hipStream_t */

enum sk_status {
  SK_OK = 0,
  SK_ERR_INVALID = -1,
  SK_ERR_HIP = -2,
  SK_ERR_LS_FAIL = -3,       /* Armijo search found no step (backtracking_line_search.py:66-67 returns t=0) */
  SK_ERR_PCG_BREAKDOWN = -4, /* p^T A p <= 0: system not positive definite */
  SK_ERR_PCG_MAXIT = -5,     /* PCG hit its iteration limit before rtol */
  SK_ERR_NONFINITE = -6,     /* NaN/Inf in a reduction */
  SK_ERR_INVERTED = -7,      /* sk_sim_step: more elements with det F <= 0 at x_next than sk_sim_params.max_inverted */
  SK_ERR_NEWTON_MAXIT = -8,  /* Newton hit its iteration limit before the gradient tolerance */
  SK_ERR_UNSUPPORTED = -9,
  SK_ERR_EIG_MAXIT = -10
};

/* Element densities (energies/ in the reference): stable_neo_hookean.py,
 * arap.py, linear_elasticity.py, macklin_mueller_neo_hookean.py, fcr.py.
 * Every density but ARAP reads lam. */
enum sk_material { SK_SNH = 0, SK_ARAP = 1, SK_LINEAR = 2, SK_MMNH = 3, SK_FCR = 4 };
enum sk_precision {
  SK_PREC_FP64 = 0,  /* every stage in FP64 */
  SK_PREC_MIXED = 1  /* FP32 element Hessian, PSD projection, assembled values and inner PCG, wrapped in FP64
                        iterative refinement; FP64 gradient, energies and near-inverted ARAP polar SVD */
};
enum sk_psd_method { SK_PSD_NONE = 0, SK_PSD_PROJ = 1, SK_PSD_ABS = 2 };

const char* sk_last_error(void);
const char* sk_status_string(int status);
int sk_version(void);

/* ------------------------------------------------------------------------ */
/* Kinematics and rest measures                                              */
/* ------------------------------------------------------------------------ */

/* Per-tet gradient operator D_e = (H Dm^-1)^T, stored D[12*e + 4*c + v], and
 * signed rest volume det(Dm)/6 (vol may be NULL).
 * Replaces deformation_jacobian.py:58-61 (and deformation_gradient.py:43-62),
 * tetrahedron_volumes.py:26-27 / volume.py:14-41. */
int sk_tet_operator(const double* X, const int32_t* T, int64_t nv, int64_t nt, double* D, double* vol,
                    sk_stream_t stream);

/* F[e] = (D_e U[T_e])^T. Replaces deformation_gradient.py:65-67 (and J @ x,
 * stable_neo_hookean.py:471,499,531). */
int sk_deformation_gradient(const double* D, const int32_t* T, const double* U, int64_t nt, double* F,
                            sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Element tier                                                              */
/* ------------------------------------------------------------------------ */

/* psi (nt), P (nt x 9), H (nt x 81) of material on F; any output may be NULL;
 * lam is ignored (may be NULL) for ARAP.
 * Replaces stable_neo_hookean_{energy,gradient,hessian}_element_F
 * (energies/stable_neo_hookean.py:65-443) and arap_{energy,gradient,hessian}_element_F
 * (energies/arap.py:71-145). */
int sk_element_F(int material, const double* F, const double* mu, const double* lam, int64_t nt, double* psi,
                 double* P, double* H, sk_stream_t stream);

/* Batched eigenvalue projection of n symmetric d x d matrices (lower triangle
 * read), method SK_PSD_PROJ (floor 1e-6) or SK_PSD_ABS; any d >= 1 (d <= 12 in
 * registers, larger d in device scratch).
 * Replaces psd_project.py:12-47. */
int sk_psd_project(const double* H, int64_t n, int d, int method, double* out, sk_stream_t stream);

/* Signed SVD F = U diag(S) V^T (U, V rotations, S[2] carries sign det F) and
 * R = U V^T; any output may be NULL.
 * Replaces svd_rv.py:8-53 and polar_svd.py:8-89. */
int sk_polar_svd(const double* F, int64_t n, double* U, double* S, double* V, double* R, sk_stream_t stream);

/* dR/dF (n x 81). Replaces rotation_gradient.py:12-75 (3D branch). */
int sk_rotation_gradient(const double* F, int64_t n, double* dRdF, sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Mesh handle: sorted CSR pattern of the vertex-pair x 3x3 block structure  */
/* and the element -> nnz map used for atomic-free assembly.                 */
/* ------------------------------------------------------------------------ */
typedef struct sk_mesh sk_mesh;

/* T_host: nt x 4 connectivity (host); D: device operator from sk_tet_operator
 * (or extracted from a user's deformation Jacobian). Replaces the pattern
 * work of sp.sparse.block_diag + J.T @ H @ J (stable_neo_hookean.py:536-537). */
int sk_mesh_create(const int32_t* T_host, const double* D, int64_t nv, int64_t nt, sk_mesh** out);
int sk_mesh_destroy(sk_mesh* mesh);
int sk_mesh_info(const sk_mesh* mesh, int64_t* nv, int64_t* nt, int64_t* nnz, int64_t* nblocks);

/* Scalar CSR pattern of the 3nv x 3nv Hessian: indptr (3nv+1), indices (nnz),
 * columns sorted within each row. */
int sk_mesh_csr_pattern(const sk_mesh* mesh, int32_t* indptr, int32_t* indices, sk_stream_t stream);

/* elem_nnz[144*e + 12*(3*i+p) + 3*j+q] = CSR position of entry
 * (3*T[e][i]+p, 3*T[e][j]+q): where K_e's entries land. */
int sk_mesh_element_nnz(const sk_mesh* mesh, int32_t* elem_nnz, sk_stream_t stream);

/* Lumped per-vertex mass (sum_e vol_e rho_e)/4; rho per element, may be NULL
 * (rho = 1). Replaces massmatrix.py:16-50 (diagonal of M) and, scaled by the
 * acceleration, gravity_force.py:8-42. */
int sk_mesh_lumped_mass(sk_mesh* mesh, const double* vol, const double* rho, double* mass, sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Global x tier on a mesh (mu, lam, vol are per-element device arrays).     */
/* F0 (nt x 9, may be NULL) is added to every F = J x: with x = u and        */
/* F0 = J x_bar these are the u tier (stable_neo_hookean.py:544-648,         */
/* arap.py:331-424).                                                         */
/* ------------------------------------------------------------------------ */

/* E = sum_e vol_e psi(F_e(x)) -> *energy_host; elem_energy (nt) optional.
 * Replaces stable_neo_hookean_energy_x (:449-474) / arap_energy_x (arap.py:242-265). */
int sk_mesh_energy(sk_mesh* mesh, int material, const double* x, const double* F0, const double* mu, const double* lam,
                   const double* vol, double* energy_host, double* elem_energy, sk_stream_t stream);

/* g = J^T vec(vol P) (3nv). Replaces stable_neo_hookean_gradient_x (:477-503) /
 * arap_gradient_x (arap.py:268-292). */
int sk_mesh_gradient(sk_mesh* mesh, int material, const double* x, const double* F0, const double* mu, const double* lam,
                     const double* vol, double* g, sk_stream_t stream);

/* CSR values (nnz, pattern of sk_mesh_csr_pattern) of sum_e B_e^T psd(vol_e H_e) B_e
 * (stable neo-Hookean: psd from the Hessian's closed-form eigensystem, other densities and
 * near-singular F by the 9x9 Jacobi).
 * psd: SK_PSD_NONE / SK_PSD_PROJ / SK_PSD_ABS. Replaces stable_neo_hookean_hessian_x
 * (:506-538) / arap_hessian_x (arap.py:295-325) including psd_project and the
 * block_diag / J.T @ H @ J assembly. */
int sk_mesh_hessian(sk_mesh* mesh, int material, const double* x, const double* F0, const double* mu, const double* lam,
                    const double* vol, int psd, double* values, sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Kinetic (inertial) terms                                                  */
/* ------------------------------------------------------------------------ */

/* y = A x for a general CSR matrix (M @ d in kinetic.py:107-119). */
int sk_csr_spmv(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values, const double* x,
                double* y, sk_stream_t stream);

/* With x_curr/x_prev: x~ = x_c + h (x_c - x_p)/h, v = (x_c - x_p)/h, d = x - x~;
 * without: d = x. Then energy = 0.5 c d.(M d) / h^2 (host) and
 * grad = c (M d) / h^2. Every output may be NULL.
 * Replaces kinetic.py:41-60 (velocity_be), :87-90 (be_target), :107-119
 * (kinetic_energy / _gradient) and :125-175 (the _be variants). */
int sk_kinetic(int64_t n, const double* x, const double* x_curr, const double* x_prev, double h, double c,
               const int32_t* M_indptr, const int32_t* M_indices, const double* M_values, double* energy_host,
               double* grad, double* x_tilde, double* velocity, sk_stream_t stream);

/* BDF2 kinetic terms (kinetic.py bdf2_target, velocity_bdf2,
 * kinetic_energy_bdf2 / _gradient_bdf2, c = 9/4):
 * x~ = (4/3) x_c - (1/3) x_p + (8h/9) v_c - (2h/9) v_p with
 * v_c = (3 x_c - 4 x_p + x_p2)/(2h), v_p = (3 x_p - 4 x_p2 + x_p3)/(2h);
 * d = x - x~, energy = 0.5 (9/4) d.(M d) / h^2 (host), grad = (9/4) (M d) / h^2;
 * velocity receives v_c. Every output may be NULL; the four history levels
 * are required. */
int sk_kinetic_bdf2(int64_t n, const double* x, const double* x_curr, const double* x_prev, const double* x_prev2,
                    const double* x_prev3, double h, const int32_t* M_indptr, const int32_t* M_indices,
                    const double* M_values, double* energy_host, double* grad, double* x_tilde, double* velocity,
                    sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Linear solve                                                              */
/* ------------------------------------------------------------------------ */

/* Jacobi-preconditioned CG on a general SPD CSR matrix, x0 = 0, stop at
 * ||r|| <= rtol ||b||. Replaces spsolve in newton.py:52 for the generic
 * (user-closure) Newton path. *iters and *relres are host outputs. */
int sk_pcg_csr(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values, const double* b,
               double* x, double rtol, int64_t maxit, int64_t* iters, double* relres, sk_stream_t stream);

/* sk_pcg_csr without any allocation: the same kernels on a static device
 * workspace (one per GPU, 25.2 MB, held until the process ends) instead of
 * buffers allocated per call; n <= 524288 (larger n returns SK_ERR_INVALID).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc). All calls on one GPU share its workspace, so calls that overlap
 * on one GPU race; results equal sk_pcg_csr's bit for bit. */
int sk_pcg_csr_noalloc(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values,
                       const double* b, double* x, double rtol, int64_t maxit, int64_t* iters, double* relres,
                       sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Device-resident backward-Euler simulator                                  */
/* ------------------------------------------------------------------------ */
typedef struct sk_sim sk_sim;

typedef struct sk_sim_params {
  int material;          /* sk_material */
  int psd;               /* SK_PSD_PROJ (reference default psd=True) or SK_PSD_NONE */
  double h;              /* time step */
  double newton_rtol;    /* stop when ||g_free|| <= newton_rtol * ||g_free(x0)|| */
  int newton_max_iter;
  double pcg_rtol;       /* block-Jacobi PCG relative residual */
  int pcg_max_iter;
  double ls_alpha;       /* Armijo constants, backtracking_line_search.py:8-17 */
  double ls_beta;
  int ls_max_iter;
  double ls_threshold;
  int timing;            /* 1: record per-phase times in sk_step_info */
  int precision;         /* sk_precision */
  double inner_rtol;     /* SK_PREC_MIXED: each time the FP32 PCG residual has dropped by this factor, the
                            correction is accumulated in FP64 and the FP64 residual b - A x replaces the
                            FP32 one (CG continues with its search direction) */
  int max_refine;        /* SK_PREC_MIXED: residual replacements per Newton solve at most */
  int spmv_blocks;       /* SpMV launch width cap (workgroups walk chunk ranges); 0 = one chunk per wave */
  int update_blocks;     /* launch width cap of the vertex-wise PCG kernels; 0 = one lane per vertex */
  int pcg_chunk;         /* PCG iterations enqueued between convergence checks (the check of one chunk
                            waits until the next is enqueued); 0 = default (8 block-Jacobi; two-level: 4
                            FP64, 2 mixed) */
  int max_inverted;      /* < 0: no limit; otherwise a step whose x_next has more elements with det F <= 0
                            returns SK_ERR_INVERTED (x_next is still written). The reference's densities
                            accept inverted elements, so the default is no limit. */
  int pcg_variant;       /* block-Jacobi PCG: 0 classic (two launches per iteration: A p, then the vector
                            update), 1 pipelined (one launch per iteration, Ghysels-Vanroose recurrences on
                            A B^-1 formed once per Newton system), SK_PREC_FP64 only: the two-level
                            preconditioner and the mixed path (whose residual replacement breaks the
                            pipelined recurrences in FP32) run the classic form */
  int psd_jacobi;        /* 1: PSD projection by the in-register 9x9 Jacobi for every density; 0: from the
                            closed-form Hessian eigensystem where the density has one (stable neo-Hookean;
                            elements with F near singular still take the Jacobi) */
} sk_sim_params;

typedef struct sk_step_info {
  int status;            /* sk_status of the step */
  int newton_iters;
  int64_t pcg_iters;
  int ls_trials;
  double g0_norm;
  double g_norm;
  double ms_gradient;    /* per-phase device times (timing=1) */
  double ms_hessian;
  double ms_pcg;
  double ms_line_search;
  int refine_steps;      /* SK_PREC_MIXED: FP64 residual replacements summed over the Newton iterations */
  int guess_used;        /* sk_sim_step_guess(es): k if Newton started from the k-th supplied guess (1-based),
                            0 if from x~ */
  double guess_dE;       /* sk_sim_step_guess(es): E(first guess) - E(x~), FP64 */
  double guess_g_norm;   /* sk_sim_step_guess(es): ||g_free|| at the guess taken, else 0 */
  int stalled;           /* 1: stopped because the update fell below position resolution (see sk_sim_step) */
  int inverted_elems;    /* elements with det F <= 0 at x_next */
  int64_t polar_fp64_elems; /* SK_PREC_MIXED, ARAP/FCR: element Hessians evaluated in FP64 because
                               det F < 1e-3, summed over the step's Newton iterations */
  double start_dE;       /* sk_sim_step_guess(es): E(start) - E(x~) of the start taken (0 for x~) */
} sk_step_info;

/* mu, lam, vol: per element; mass: per DOF (3nv, M = kron(M_lumped, I3));
 * f_ext: per DOF external force (may be NULL); fixed: per DOF 0/1 (may be
 * NULL). Arrays are copied. Replaces the closure set-up of backward_euler.py:76-83
 * (with kinetic.py:125-194). */
int sk_sim_create(sk_mesh* mesh, const sk_sim_params* params, const double* mu, const double* lam,
                  const double* vol, const double* mass, const double* f_ext, const uint8_t* fixed, sk_sim** out);
int sk_sim_destroy(sk_sim* sim);

/* One backward-Euler step: Newton on E(x) = V(x) + 1/(2h^2)|x - x~|_M^2 - f_ext.x
 * from x0 = x~ = x_curr + h (x_curr - x_prev)/h, fixed DOFs held at x_bc
 * (or at x_curr when x_bc is NULL). Writes x_next; info is a host struct.
 * Newton stops when ||g_free|| <= newton_rtol ||g_free(x0)||, or when the
 * accepted update no longer moves the positions (max |t dx| <= 4 eps max |x|
 * over the free DOFs; info->stalled = 1): then x0 was already an equilibrium
 * to within the resolution of the positions (a damped mesh at rest) and the
 * relative test cannot be met in floating point.
 * Replaces backward_euler.py:27-91 -> newton.py:7-75 -> backtracking_line_search.py:8-67. */
int sk_sim_step(sk_sim* sim, const double* x_curr, const double* x_prev, const double* x_bc, double* x_next,
                sk_step_info* info, sk_stream_t stream);

/* sk_sim_step with a candidate start point x_guess (3nv, device). Newton
 * starts from x_guess (fixed DOFs replaced by their prescribed values) when
 * its incremental potential is lower than at x~, else from x~ as
 * sk_sim_step does; the stopping test stays ||g_free|| <= newton_rtol *
 * ||g_free(x~)||, so both start points meet the same absolute tolerance.
 * No counterpart in the reference (backward_euler.py:85 always starts from
 * x~); it is the hook for a learned warm start. */
int sk_sim_step_guess(sk_sim* sim, const double* x_curr, const double* x_prev, const double* x_bc,
                      const double* x_guess, double* x_next, sk_step_info* info, sk_stream_t stream);

/* sk_sim_step_guess with up to 8 candidate start points (guesses: host array
 * of n_guess device pointers, caller's numbering): Newton starts from the
 * candidate with the lowest incremental potential when that is below E(x~)
 * (ties go to the earlier candidate), else from x~; info.guess_used is the
 * 1-based index of the start taken. For a learned warm start with a
 * non-learned fallback: a prediction that is worse than the fallback is
 * rejected in favour of it rather than of x~. No counterpart in the
 * reference. */
int sk_sim_step_guesses(sk_sim* sim, const double* x_curr, const double* x_prev, const double* x_bc, int n_guess,
                        const double* const* guesses, double* x_next, sk_step_info* info, sk_stream_t stream);

/* One BDF2 step (bdf2.py:31-101 with kinetic.py bdf2_target / velocity_bdf2,
 * c = 9/4): Newton on V(x) + c/(2h^2) |x - x~|_M^2 - f_ext.x from
 * x0 = x~ = (4/3) x_c - (1/3) x_p + (8h/9) v_c - (2h/9) v_p, with
 * v_c = (3 x_c - 4 x_p + x_p2)/(2h) and v_p = (3 x_p - 4 x_p2 + x_p3)/(2h).
 * Same simulator, parameters and status codes as sk_sim_step. */
int sk_sim_step_bdf2(sk_sim* sim, const double* x_curr, const double* x_prev, const double* x_prev2,
                     const double* x_prev3, const double* x_bc, double* x_next, sk_step_info* info,
                     sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Batched implicit steps over many independent meshes                       */
/* No counterpart in the reference, whose backward_euler.py:27-91 advances one */
/* mesh per call; each problem here gets exactly that step (the same Newton,  */
/* line-search and PCG rules on its own scalars) with the mixed-precision     */
/* linear solve.                                                             */
/* ------------------------------------------------------------------------ */
typedef struct sk_batch sk_batch;

typedef struct sk_batch_params {
  int material;          /* sk_material */
  int psd;               /* sk_psd_method */
  double h;
  double newton_rtol;    /* per problem: stop when ||g_free|| <= newton_rtol ||g_free(x~)|| */
  int newton_max_iter;
  double pcg_rtol;       /* FP64 residual target of the per-problem solve */
  int pcg_max_iter;
  double inner_rtol;     /* FP32 residual drop between FP64 residual replacements */
  double ls_alpha;
  double ls_beta;
  int ls_max_iter;
  double ls_threshold;
  int block;             /* threads per workgroup of the per-problem kernels (128, 256, 512, 1024); 0 = 256 */
  int lanes;             /* lanes per matrix row in the per-problem SpMV (1 or 4); 0 = 1 */
  int lds;               /* 1: the per-problem PCG keeps the problem's vectors and row data in LDS (when the largest problem fits) */
  int use_graph;         /* 1: one Newton iteration of the batch is captured in a hipGraph and replayed */
} sk_batch_params;

typedef struct sk_batch_info {
  int status;
  int newton_iters;
  int64_t pcg_iters;
  int ls_trials;
  int refine_steps;
  double g0_norm;
  double g_norm;
  int stalled;           /* as in sk_step_info */
} sk_batch_info;

/* mesh: the union of the problems' meshes, problem p owning vertices
 * [vstart[p], vstart[p+1]) and tets [tstart[p], tstart[p+1]) (host arrays of
 * nprob+1; T_host (4 nt) is the union's tet array, checked against the
 * ranges). mu, lam, vol per union element; mass, f_ext, fixed per union DOF
 * (device; f_ext and fixed may be NULL). */
int sk_batch_create(sk_mesh* mesh, int nprob, const int64_t* vstart, const int64_t* tstart, const int32_t* T_host,
                    const sk_batch_params* params, const double* mu, const double* lam, const double* vol,
                    const double* mass, const double* f_ext, const uint8_t* fixed, sk_batch** out);
int sk_batch_destroy(sk_batch* batch);
/* One backward-Euler step of every problem (union DOF vectors on the device;
 * x_bc may be NULL). info (host, nprob entries) and rounds (batch Newton
 * iterations launched) may be NULL. */
int sk_batch_step(sk_batch* batch, const double* x_curr, const double* x_prev, const double* x_bc, double* x_next,
                  sk_batch_info* info, int* rounds, sk_stream_t stream);
/* Sets the per-problem kernel shape (a recorded autotuning choice). */
int sk_batch_set_shape(sk_batch* batch, int block, int lanes, int lds);

/* Two-level preconditioner for the per-problem PCG, no counterpart in the
 * reference: block-Jacobi plus an additive coarse correction W E^-1 W^T per
 * problem, where W holds the rigid motions (about the centroid, restricted to
 * the free DOFs, orthonormalised) of each run of agg_rows consecutive
 * vertices of the problem (a multiple of 64, at most 512). E = W^T A W is
 * formed from the system assembled at the start of the next sk_batch_step
 * (period = 0: only then; k >= 1: again at the start of every k-th step after
 * it, 1 keeping it matched to large rotations) and inverted per problem on the
 * device. X: host
 * rest positions of the union mesh (nv x 3). agg_rows <= 0 returns to
 * block-Jacobi. Consecutive vertices should be spatially close (as the
 * cantilever generator numbers them). */
int sk_batch_set_coarse(sk_batch* batch, const double* X, int agg_rows, int period);
/* Times the per-problem PCG kernel on the first Newton system of a step from
 * (x_curr, x_prev, x_bc) for every (lds, block, lanes) in {0, 1} x {128, 256,
 * 512, 1024} x {1, 4} (reps launches each, after one warm-up; lds = 1 only
 * when the largest problem's vectors fit) and keeps the fastest. ms_out
 * (host, 16 entries, may be NULL) receives the mean times in that order (-1:
 * not run). The batch state is reset by the next sk_batch_step. */
int sk_batch_autotune(sk_batch* batch, const double* x_curr, const double* x_prev, const double* x_bc, int reps,
                      int* block_out, int* lanes_out, int* lds_out, double* ms_out, sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Linear modal analysis                                                     */
/* ------------------------------------------------------------------------ */
typedef struct sk_modal_info {
  int sweeps;                /* filter + Rayleigh-Ritz sweeps */
  int64_t spmvs;             /* block sparse products */
  int block;                 /* subspace block size (64 or 128) */
  double max_rel_residual;   /* max_j ||A y_j - theta_j y_j|| / max(|theta_j|, 1e-6 spectrum_bound), j < k */
  double spectrum_bound;     /* Gershgorin bound of D (K + shift M) D */
  double ms;                 /* wall time of the call */
} sk_modal_info;

/* Lowest k (1..96) eigenpairs of (K + shift M) phi = lambda M phi over the
 * free DOFs: K the assembled element Hessian of `material` at positions x
 * (3nv, device; no PSD projection), M = diag(mass) (3nv per-DOF lumped mass,
 * device), fixed (3nv 0/1, device, may be NULL) DOFs eliminated and returned
 * as zero rows. basis (device, 3nv x k, row-major: k coefficients per DOF)
 * receives M-orthonormal eigenvectors, eigenvalues (host, k) ascending.
 * Chebyshev-filtered subspace iteration (degree `degree` per sweep, at most
 * max_sweeps) until every returned pair has relative residual <= tol
 * (against max(|lambda|, 1e-6 spectrum_bound), so zero modes converge), from a
 * block seeded by `seed` (deterministic). Returns SK_ERR_EIG_MAXIT (outputs
 * still written) when max_sweeps is reached first. Replaces
 * linear_modal_analysis.py:18-73 (material = SK_LINEAR, mu = lam = 1,
 * shift 1e-6 with bI given) and its eigs.py:66-110 shift-invert solve. */
int sk_modal_basis(sk_mesh* mesh, int material, const double* x, const double* mu, const double* lam,
                   const double* vol, const double* mass, const uint8_t* fixed, int k, double shift, double tol,
                   int degree, int max_sweeps, uint64_t seed, double* basis, double* eigenvalues,
                   sk_modal_info* info, sk_stream_t stream);

/* Tall-skinny helpers for modal coordinates (no counterpart in the
 * reference): q_j = sum_r basis[r][j] w[r] (x[r] - base[r]) for j < k
 * (k <= 256; w, base may be NULL; q on the device), out = base + basis c
 * (base may be NULL), and out = a x + b y. basis is rows x k row-major. */
int sk_dense_project(int64_t rows, int k, const double* basis, const double* weight, const double* x,
                     const double* base, double* q, sk_stream_t stream);
int sk_dense_expand(int64_t rows, int k, const double* basis, const double* coef, const double* base, double* out,
                    sk_stream_t stream);
int sk_axpby(int64_t n, double a, const double* x, double b, const double* y, double* out, sk_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Learned warm start: BF16 MFMA multilayer perceptron                       */
/* No counterpart in the reference.                                          */
/* ------------------------------------------------------------------------ */

/* C (M x N, FP32, row stride ldc) = A B with BF16 operands given by element
 * strides: A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; M, N, K
 * positive multiples of 16; FP32 accumulation on the matrix cores. */
int sk_gemm_bf16(int M, int N, int K, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                 int64_t sbn, float* C, int64_t ldc, sk_stream_t stream);

/* MLP dims[0] -> dims[1] -> ... -> dims[n_layers] with exact-erf GELU after
 * every layer but the last; BF16 weights and activations, FP32 accumulation,
 * FP32 master weights and Adam moments. Parameters (host, FP32) are laid out
 * layer by layer as W_l (dims[l+1] x dims[l], row-major) then b_l. */
typedef struct sk_mlp sk_mlp;
int sk_mlp_create(int n_layers, const int* dims, int max_batch, sk_mlp** out);
int sk_mlp_destroy(sk_mlp* net);
int64_t sk_mlp_num_params(const sk_mlp* net);
int sk_mlp_set_params(sk_mlp* net, const float* params, sk_stream_t stream);   /* also resets Adam */
int sk_mlp_get_params(sk_mlp* net, float* params, sk_stream_t stream);
/* Y (device, batch x dims[L], FP32) = net(X) (device, batch x dims[0], FP32). */
int sk_mlp_forward(sk_mlp* net, const float* X, int batch, float* Y, sk_stream_t stream);
/* A normalized predictor in one call, FP64 in and out (device arrays): Y = net((X - in_mu) / in_sd)
 * * out_sd + out_mu, with per-column in_mu, in_sd (dims[0]) and out_mu, out_sd (dims[L]); the
 * normalized input is rounded to FP32, then to BF16 as in sk_mlp_forward. No counterpart in the
 * reference (the learned warm start's inference). */
int sk_mlp_predict(sk_mlp* net, const double* X, int batch, const double* in_mu, const double* in_sd,
                   const double* out_mu, const double* out_sd, double* Y, sk_stream_t stream);
/* Forward + backward of the mean squared error against T (device, batch x
 * dims[L]); keeps the gradient for sk_mlp_adam / sk_mlp_get_grad; *loss_host
 * (may be NULL: no synchronization) receives the loss. */
int sk_mlp_backward(sk_mlp* net, const float* X, const float* T, int batch, double* loss_host, sk_stream_t stream);
int sk_mlp_get_grad(sk_mlp* net, float* grad, sk_stream_t stream);
/* One Adam step (bias-corrected) with the last gradient. */
int sk_mlp_adam(sk_mlp* net, float lr, float beta1, float beta2, float eps, sk_stream_t stream);

/* Two-level PCG preconditioner, no counterpart in the reference (which
 * factorises each Newton system with SuperLU, newton.py:52): block-Jacobi
 * plus an additive coarse correction W E^-1 W^T over piecewise-rigid
 * aggregates. X (host, nv x 3) gives the rest positions; the bounding box is
 * cut into cubic bins of about per_agg vertices, each non-empty bin is an
 * aggregate with its 6 rigid motions restricted to the free DOFs
 * (orthonormalised, degenerate ones dropped). E = W^T A W is formed from the
 * first system the simulator then solves and inverted once on the device;
 * later solves reuse it (any symmetric positive definite E gives a valid
 * preconditioner). per_agg <= 0 returns to block-Jacobi. Fails with
 * SK_ERR_INVALID when the coarse space would exceed 4096 unknowns. */
int sk_sim_set_coarse(sk_sim* sim, const double* X, int per_agg);
/* Host-only, no counterpart in the reference: the vertices of the tet mesh
 * T (host, nt x 4) in the row order of the simulator's sliced-ELL matrix
 * (stable, by descending number of neighbour vertices including itself;
 * order[nv]). A mesh renumbered in this order has the identity row order,
 * so the SpMV's per-row stores are contiguous. */
int sk_sell_order(const int32_t* T, int64_t nv, int64_t nt, int32_t* order);
/* Host-only, no counterpart in the reference: the vertices of X (host,
 * nv x 3) grouped by the aggregates sk_sim_set_coarse would form with
 * per_agg, ascending within each aggregate (order[nv]), the number of
 * aggregates (n_agg, may be NULL) and where each aggregate's run starts
 * (agg_ptr[n_agg + 1], at most nv + 1 entries; may be NULL). A mesh
 * renumbered in this order gives every aggregate a contiguous vertex range,
 * which the two-level kernels then read with unit stride. */
int sk_coarse_order(const double* X, int64_t nv, int per_agg, int32_t* order, int* n_agg, int32_t* agg_ptr);
/* Host-side vertex numbering of the states, no counterpart in the reference:
 * order_host (nv entries, a permutation) says that the simulator's mesh
 * vertex t is the caller's vertex order_host[t]. Every state the step entry
 * points read (x_curr, x_prev, x_prev2, x_prev3, x_bc, x_guess) and x_next are
 * then in the caller's numbering; the step's first and last kernels permute
 * them. The mesh, mass, f_ext and fixed given to sk_sim_create are in the
 * simulator's numbering. NULL restores the identity. Lets a caller keep its
 * own vertex order while the simulator numbers the mesh in the order its
 * kernels read it (sk_sell_order, sk_coarse_order). */
int sk_sim_set_order(sk_sim* sim, const int32_t* order_host);
/* Aggregates, coarse unknowns, whether E^-1 has been formed, and the time it took (ms). */
int sk_sim_coarse_info(const sk_sim* sim, int* n_agg, int* nc, int* built, double* build_ms);

/* Measurement hook, no counterpart in the reference: re-initialises PCG
 * scratch vectors on the system assembled by the last sk_sim_step (RHS = the
 * last gradient) and enqueues `reps` block-CSR SpMV launches back to back on
 * `stream`, so a caller can bracket them with events. The simulator's state
 * (positions, gradient, assembled values) is not modified. */
int sk_sim_spmv_reps(sk_sim* sim, int reps, sk_stream_t stream);

/* Measurement hook, no counterpart in the reference: enqueues `reps` launches
 * of the simulator's element pass (F, 9x9 Hessian, vol weight, PSD
 * projection, staged blocks) at the current positions on `stream`. Only the
 * element staging buffer is overwritten; the next step re-assembles. */
int sk_sim_hess_reps(sk_sim* sim, int reps, sk_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SIMKIT_AMD_H */
