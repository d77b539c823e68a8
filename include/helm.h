/*
 * helm.h -- C ABI of libhelm.so, the MI355X frequency-domain Helmholtz solver.
 *
 * 2-D variable-density viscoacoustic Helmholtz operator in flux form with a
 * complex-coordinate-stretched quadratic PML, multiplied by sx*sz so that the
 * matrix is complex symmetric (A^T = A), exp(-i omega t) convention:
 *
 *   (A u)[i,j] = k0^2 sz_i sx_j kappa_c[i,j] u[i,j]
 *              + sum_{4 edges e} F_e (u_neighbour - u[i,j]),     u_ghost = 0
 *   F_e = stretch_e * mean(eta_rel over the edge ends, ghost end = 1) / h^2
 *
 * eta_rel = 1 + m_eta = rho0/rho,  kappa_c = (1 + m_kappa) (1 + i alpha c/omega)^2.
 * This is the operator of nsrno/theranostic/wave.py:127-188 (Solver.matrix),
 * extended by the complex kappa_c for attenuation (op_kind HELM_OP_FLUX5).
 *
 * op_kind HELM_OP_STAG8 holds the spectral operator of nsrno/wave_operator.py
 * (:105-109) on a periodic box with an absorbing sponge instead of a PML:
 *
 *   L u = [D-x(ax D+x u) + D-z(az D+z u)] / k0^2 + kappa_S u,
 *   (D+ u)_j = sum_m a_m (u_{j+m+1} - u_{j-m}) / h,
 *   (D- g)_j = sum_m a_m (g_{j+m} - g_{j-m-1}) / h,
 *   a = (1225/1024, -245/3072, 49/5120, -5/7168),  indices mod the box size,
 *
 * whose Fourier symbol is the reference's (wave_operator.py:22-28).  That
 * operator uses the opposite sign convention (outgoing exp(-i k r), damping
 * as -i terms), so the plan stores it in this library's convention,
 * A = k0^2 conj(L) (real ax, az; kappa_c = conj(kappa_S)); the adjoint
 * A^H = k0^2 L is the reference's operator, and the reference's solve
 * L u = f is helm_solve(adjoint = 1) with right-hand side k0^2 f.
 *
 * Conventions: arrays are [rhs][z][x] row-major, complex numbers are
 * interleaved (re, im); c64 = 2 x float32, c128 = 2 x float64.  All data
 * pointers are DEVICE pointers unless stated.  Every call is asynchronous on
 * the given HIP stream (hipStream_t passed as void*, NULL = default stream),
 * except helm_solve, which synchronises once per restart cycle.  Every entry
 * point returns 0 on success or a HELM_ERR_* code; helm_error_string() turns
 * it into text and helm_last_error() gives the detail of the last failure on
 * the calling thread.  A plan is not thread safe: one plan per stream.
 *
 * The reference has no C ABI on this path; the native precedent it follows is
 * nsrno/native_time.cu:117-155 (extern "C" int entry points, device pointers
 * from torch, a stream argument, an error-string function).
 */
#ifndef HELM_H
#define HELM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HELM_OK 0
#define HELM_ERR_ARG 1     /* invalid argument (shape, size, value)        */
#define HELM_ERR_HIP 2     /* a HIP runtime call failed                    */
#define HELM_ERR_ALLOC 3   /* device allocation failed                     */
#define HELM_ERR_STATE 4   /* call order (e.g. solve before set_medium)    */

/* termination reasons, nsrno/wave_solver.py:17,36 and nsrno/dense_wave.py:131-145 */
#define HELM_CONVERGED 0
#define HELM_ITERATION_BUDGET 1
#define HELM_BREAKDOWN 2
#define HELM_STAGNATION 3     /* no progress over the stagnation window     */
#define HELM_DIVERGED 4       /* BiCGSTAB: the recomputed residual exploded (dense_wave.py:140-142) */

#define HELM_GMRES 0
#define HELM_BICGSTAB 1

#define HELM_PC_NONE 0     /* plain (unpreconditioned) Krylov               */
#define HELM_PC_MG 1       /* complex-shifted-Laplacian multigrid V-cycle   */
#define HELM_PC_LEARNED 2  /* learned conv-net correction + MG V-cycle      */
#define HELM_PC_SPECTRAL 3 /* Fourier multiplier gain * ifft2(M fft2(v))    */
#define HELM_PC_SPECTRAL_PLAIN 4 /* the same without the gain: ifft2(M fft2(v)),
                                  the preconditioner of solve_frequency's GMRES
                                  (nsrno/wave_solver.py:56-59)             */

#define HELM_OP_FLUX5 0    /* 5-point flux form, stretched PML (wave.py)   */
#define HELM_OP_STAG8 1    /* 8th-order staggered, periodic + sponge       */
#define HELM_OP_COND8 2    /* 8th-order staggered, split PML, masked box   */

typedef struct helm_plan helm_plan;

/* Grid, frequency and PML: the arguments of Grid/Solver/coefficients,
 * nsrno/theranostic/wave.py:29-156.  Coordinates in mm, frequency in MHz. */
typedef struct {
    int32_t nz, nx;            /* nodes including the PML                    */
    int32_t nrhs;              /* capacity: right-hand sides per call        */
    int32_t restart;           /* capacity: Krylov basis length m (0: none)  */
    double x0_mm, z0_mm;       /* coordinate of node [0,0]                   */
    double h_mm;               /* equal spacing in x and z                   */
    double pml_mm;             /* PML thickness on every side (0: none)      */
    double frequency_mhz;
    double c0_m_s;             /* k0 = omega/c0 and sigma_max reference      */
    double pml_strength;       /* sigma_max = pml_strength*c0/pml_width      */
    int32_t op_kind;           /* HELM_OP_FLUX5, HELM_OP_STAG8, HELM_OP_COND8 */
} helm_desc;

typedef struct {
    int32_t method;            /* HELM_GMRES or HELM_BICGSTAB                */
    int32_t precond;           /* HELM_PC_*                                  */
    int32_t restart;           /* m, 1 <= m <= helm_desc.restart             */
    int32_t max_iterations;    /* budget of Krylov steps per right-hand side */
    double tolerance;          /* on the true FP64 ||b - A x|| / ||b||       */
    int32_t adjoint;           /* nonzero: solve A^H x = b                   */
    int32_t use_x0;            /* nonzero: x holds the initial guess         */
    /* called once per restart with the true relative residuals (host array
     * [nrhs]) and the operator applications so far, like the callback of
     * nsrno/wave_solver.py:9-17; may be NULL */
    void (*callback)(void* user, int32_t applications, const double* relative, int32_t nrhs);
    void* user;
    /* nonzero: GMRES replays every full restart cycle after the first as one
     * captured HIP graph (the cycle's kernels, preconditioner included) */
    int32_t use_graph;
    /* Automatic fallback (nsrno/dense_wave.py:212-216, 233-235):
     * - GMRES with HELM_PC_LEARNED: a right-hand side whose true residual has
     *   not fallen below stagnation_ratio times its value stagnation_steps
     *   Krylov steps earlier is stagnating; with fallback != 0 the solve goes
     *   on with the multigrid preconditioner (info.fallback = 1), otherwise
     *   that right-hand side stops with HELM_STAGNATION;
     * - BiCGSTAB: with fallback != 0 the right-hand sides that end without
     *   converging (breakdown, divergence, budget) are solved on by FGMRES with
     *   the multigrid preconditioner (plain GMRES when none is set up) from
     *   the iterate BiCGSTAB left (from the initial guess if that is not
     *   finite), within the remaining iteration budget.
     * stagnation_steps <= 0: 60; stagnation_ratio <= 0: 0.5. */
    int32_t fallback;
    int32_t stagnation_steps;
    double stagnation_ratio;
} helm_solve_opts;

typedef struct {
    int32_t reason;            /* HELM_CONVERGED ...                         */
    int32_t iterations;        /* Krylov steps (preconditioned applies)      */
    int32_t applications;      /* all operator applications incl. residuals  */
    int32_t cycles;            /* restarts / refinement cycles               */
    double relative_residual;  /* true FP64 ||b - A x|| / ||b|| at return    */
    int32_t fallback;          /* 1: finished by the fallback solver          */
    int32_t first_reason;      /* reason of the first solver (before any fallback) */
    int32_t graph_fallbacks;   /* restart cycles run directly because their graph
                                  capture failed (use_graph; 0 otherwise)     */
} helm_solve_info;

/* ---- library ---------------------------------------------------------- */
const char* helm_version(void);
const char* helm_error_string(int32_t code);
const char* helm_last_error(void);

/* ---- plan: replaces Solver.__init__/coefficients/matrix (wave.py:99-188) */
int32_t helm_plan_create(const helm_desc* desc, helm_plan** plan);
int32_t helm_plan_destroy(helm_plan* plan);
/* eta_rel [N] float64 and kappa_c [N] c128 (device).  Replaces the
 * m_kappa/m_eta arguments of Solver.matrix (wave.py:174-179). */
int32_t helm_set_medium(helm_plan* plan, const double* eta_rel, const double* kappa_c, void* stream);
/* Physical maps (device float64 [N]): c [m/s], rho [kg/m^3], alpha [Np/m];
 * eta_rel = rho0/rho, kappa_c = rho0 c0^2/(rho c^2) (1 + i alpha c/omega)^2.
 * Replaces setup.physical_maps (setup.py:299-307) run backwards and the
 * alpha power law of nsrno/wave_operator.py:56,67. alpha may be NULL. */
int32_t helm_set_medium_physical(helm_plan* plan, const double* c, const double* rho,
                                 const double* alpha, double rho0, void* stream);

/* ---- operator: replaces Solver.matrix(...) @ u (wave.py:174-188) ------- */
/* out = A u (adjoint != 0: A^H u) for nrhs fields; precision 32: c64 in/out
 * computed in FP32; 64: c128 in/out computed in FP64. */
int32_t helm_apply(helm_plan* plan, const void* u, void* out, int32_t nrhs,
                   int32_t precision, int32_t adjoint, void* stream);

/* ---- ports: replaces Solver.solve's RHS = -W/h^2 and Solver.sample ------
 * (wave.py:260, 275-278).  The sparse port weights W (N x n_ports, real) are
 * passed twice: node-major CSR (node_ptr [n_nodes+1], node_idx [n_nodes],
 * port_of_tap [nnz], w_node [nnz]) for injection, port-major CSC (port_ptr
 * [n_ports+1], node_of_tap [nnz], w_port [nnz]) for sampling; int32/float64. */
/* rhs[r] = -(1/h^2) sum_p amp[r][p] W[:,p]  (rhs c128 [nrhs][N], zeroed
 * first; amp c128 [nrhs][n_ports]) */
int32_t helm_inject(helm_plan* plan, int32_t n_ports, int32_t n_nodes, const int32_t* node_ptr,
                    const int32_t* node_idx, const int32_t* port_of_tap, const double* w_node,
                    const void* amp, void* rhs, int32_t nrhs, void* stream);
/* traces[r][p] = sum_n W[n,p] field[r][n]; field c128 (precision 64) or c64
 * (precision 32), traces c128 [nrhs][n_ports] */
int32_t helm_sample(helm_plan* plan, int32_t n_ports, const int32_t* port_ptr,
                    const int32_t* node_of_tap, const double* w_port, const void* field,
                    int32_t precision, void* traces, int32_t nrhs, void* stream);

/* ---- Krylov solve: replaces Solver.solve's splu (wave.py:253-261) and
 * nsrno/wave_solver.gmres (wave_solver.py:9-36) / dense_wave.solve_system
 * (dense_wave.py:167-225).  rhs and x are c128 [nrhs][N]; info is a HOST
 * array [nrhs].  Independent Krylov spaces per right-hand side, advanced in
 * lock step; FP32 stencil and basis vectors, FP64 inner products,
 * Hessenberg/Givens least squares and true-residual refinement at every
 * restart. */
int32_t helm_solve(helm_plan* plan, const void* rhs, void* x, int32_t nrhs,
                   const helm_solve_opts* opts, helm_solve_info* info, void* stream);

/* ---- batched solve over plans of different grid sizes: replaces the
 * per-frequency loop of nsrno/reconstruct_dense.py:91 (frequencies handed to
 * processes, each frequency's angles batched by dense_wave.solve_system,
 * dense_wave.py:167-225).  plans[k], rhs[k], x[k] (c128 [nrhs][N_k]) for
 * k < nplans (<= 256), the same nrhs (<= 64) for all; info is a HOST array
 * [nplans][nrhs].  Multigrid FGMRES only (opts.method HELM_GMRES, precond
 * HELM_PC_MG; every plan set up with pre = post = 2 and a finest level above
 * the LDS tail size); opts.callback, fallback and stagnation are not used.
 * Each kernel of a restart cycle is one segmented launch over every plan
 * still running; with use_graph the cycle is captured once and replayed.
 * Per plan the numerics are those of helm_solve on the same plan with its
 * default launch shapes (the autotuned stencil shape is not used). */
int32_t helm_batch_solve(int32_t nplans, helm_plan* const* plans, const void* const* rhs,
                         void* const* x, int32_t nrhs, const helm_solve_opts* opts,
                         helm_solve_info* info, void* stream);

/* ---- reverse-communication FGMRES: replaces nsrno/wave_solver.py:9-36
 * (gmres(operator, precondition, rhs, x, callback, restart, max_cycles,
 * tolerance)) for operators and preconditioners the CALLER applies (any
 * tensor -> tensor code), and the batched cycle of nsrno/dense_wave.py:84-116.
 * The library keeps the Krylov state of nrhs independent spaces of n complex
 * entries: FP64 Hessenberg columns and Givens rotations, DCGS2 partial sums;
 * the caller owns the vectors, all device memory:
 *   V  c64 [restart + 1][nrhs][n]  basis        Z  c64 [restart][nrhs][n]  directions
 *   w  c64 [nrhs][n]               A z_j        r, b, x  c128 [nrhs][n]
 * One cycle: helm_krylov_begin(r = b - A x) -> rel; for j = 0..m-1 the
 * caller writes Z_j = M^{-1}(V_j) and w = A Z_j, then helm_krylov_step(j);
 * helm_krylov_end adds Z y to x.  A right-hand side whose Givens estimate
 * reaches tolerance, or whose new basis vector vanishes (happy breakdown),
 * stops its cycle early (later steps skip it); convergence is decided on the
 * true residual r at the next helm_krylov_begin. */
typedef struct helm_krylov helm_krylov;
int32_t helm_krylov_create(int64_t n, int32_t nrhs, int32_t restart, helm_krylov** k);
int32_t helm_krylov_destroy(helm_krylov* k);
/* V_0 = r / ||r|| per RHS; first != 0 also takes ||b|| (b may be NULL
 * otherwise).  rel[nrhs] (host) = ||r||/||b||, active[nrhs] (host) =
 * rel >= tolerance.  Synchronises the stream. */
int32_t helm_krylov_begin(helm_krylov* k, const void* b, const void* r, void* V, double tolerance, int32_t first,
                          double* rel, int32_t* active, void* stream);
/* orthogonalise w = A Z_j against V_0..V_j (DCGS2, FP64 sums), Hessenberg
 * column and rotation, V_{j+1} = w / ||w||; last != 0 ends the cycle at j */
int32_t helm_krylov_step(helm_krylov* k, int32_t j, int32_t last, void* V, void* w, double tolerance, void* stream);
/* x += Z y (FP64); steps (host, may be NULL) = steps used per RHS this cycle */
int32_t helm_krylov_end(helm_krylov* k, const void* Z, void* x, int32_t* steps, void* stream);

/* ---- adjoint gradient: replaces the material-gradient step of
 * dense_wave.ImplicitChannels.backward (nsrno/dense_wave.py:240-256).
 * u, lam: c128 [nrhs][N] forward and adjoint (A^H lam = P^T g) fields;
 * s_eta, s_kap: c128 [N] outputs,
 *   s_kap[q] = sum_r conj(lam_rq) k0^2 sz sx u_rq,
 *   s_eta[q] = sum_r lam_r^H (dA/d eta_rel[q]) u_r,
 * so dL/dm = -Re(sum_q s_eta[q] d eta_rel[q]/dm + s_kap[q] d kappa_c[q]/dm). */
int32_t helm_material_sensitivity(helm_plan* plan, const void* u, const void* lam, int32_t nrhs,
                                  void* s_eta, void* s_kap, void* stream);

/* ---- preconditioners --------------------------------------------------- */
/* Complex-shifted-Laplacian multigrid: kappa_c -> kappa_c (1 + i shift).
 * levels >= 1 (1: smoother only), pre/post Jacobi sweeps, coarse sweeps on
 * the coarsest level.  omega = 0 picks a complex Jacobi weight per node
 * from the local stencil's Fourier symbol (convergent on coarse Helmholtz
 * grids and in the PML); omega in (0, 1.5] forces that real weight. */
int32_t helm_mg_setup(helm_plan* plan, int32_t levels, double shift, int32_t pre, int32_t post,
                      int32_t coarse_sweeps, double omega, void* stream);
/* z = M^{-1} v for nrhs c64 fields (the preconditioner the solve applies). */
int32_t helm_precond_apply(helm_plan* plan, int32_t precond, const void* v, void* z, int32_t nrhs,
                           void* stream);
/* the same, adjoint != 0: z = M^{-H} v (what adjoint solves apply; for the
 * spectral plan, k0^2 M^{-H} preconditions the reference's own operator L) */
int32_t helm_precond_apply_ex(helm_plan* plan, int32_t precond, const void* v, void* z, int32_t nrhs,
                              int32_t adjoint, void* stream);

/* ---- Fourier-multiplier preconditioner: replaces the spectral
 * preconditioner of nsrno/wave_operator.py:93-102,129-135 (direction
 * gain * ifft2(multiplier * fft2(r)), incl. a trained model's context gain
 * and spectral modulation) and conditioned_wave.py:90-92 (gain = the mask).
 * multiplier, gain: device c128 [nz][nx] on the plan's periodic box (gain
 * NULL: 1).  HELM_PC_SPECTRAL then applies it in helm_warm_start, helm_solve
 * and helm_precond_apply_ex: the map as given preconditions the plan's own
 * operator (forward solves), except on spectral plans (A = k0^2 conj(L)),
 * where adjoint solves (the reference's L) take it as given; the other
 * direction applies the conjugate map conj(P(conj v)).  FFTs are mixed-radix Stockham passes in FP32; the
 * plan allocates two transform buffers of nrhs-capacity fields. */
int32_t helm_spectral_set_preconditioner(helm_plan* plan, const void* multiplier, const void* gain, void* stream);
/* 2-D DFT of batch boxes [batch][H][W], c64 (precision 32) or c128 (64)
 * (inverse != 0: e^{+i}, unscaled), H, W >= 2; work: a buffer of the same
 * size; in, out, work distinct.  Replaces torch.fft.fft2/ifft2 on the path. */
int32_t helm_fft2(int32_t H, int32_t W, int32_t batch, int32_t inverse, int32_t precision, const void* in, void* out,
                  void* work, void* stream);

/* ---- learned preconditioner network ----------------------------------
 * A 4-level, 32-channel 3x3-conv U-Net (BF16 MFMA implicit-GEMM kernels, FP32
 * accumulation).  Replaces the reference's LearnedPreconditioner and its
 * training harness (nsrno/conditioned_wave.py:94-117,
 * nsrno/train_preconditioner.py:45-109): with HELM_PC_LEARNED the flexible
 * GMRES direction is z = MG(v) + N(v).  Tensors are NHWC bf16
 * [batch][H][W][32] on an H x W canvas (multiples of 8, >= the grid; the
 * grid sits at the top-left).  Parameters: fp32 [12][9*32*32 + 32]
 * (weights [tap][out][in], then biases), device memory owned by the caller. */
typedef struct helm_net helm_net;
int32_t helm_net_create(int32_t H, int32_t W, int32_t batch, int32_t training, helm_net** net);
int32_t helm_net_destroy(helm_net* net);
int32_t helm_net_param_count(void);
/* linear != 0: identity activations and only the residual channels as input,
 * so N(v) is a linear operator of v (keep the biases at zero) */
int32_t helm_net_set_linear(helm_net* net, int32_t linear);
int32_t helm_net_set_params(helm_net* net, const float* params, void* stream);
int32_t helm_net_forward(helm_net* net, const void* in, void* out, int32_t batch, void* stream);
int32_t helm_net_backward(helm_net* net, const void* dout, float* grad, int32_t batch, void* stream);
int32_t helm_adam_step(int64_t count, float* params, const float* grad, float* m, float* v, int32_t step,
                       double lr, double beta1, double beta2, double eps, double grad_scale, void* stream);
/* attach (not owned) the net used by HELM_PC_LEARNED; NULL detaches */
int32_t helm_plan_set_net(helm_plan* plan, helm_net* net);
/* z = N(v): the network correction alone, c64 [nrhs][N] */
int32_t helm_net_correction(helm_plan* plan, const void* v, void* z, int32_t nrhs, void* stream);
/* the net's input features for residuals v (bf16 [nrhs][H][W][32]) */
int32_t helm_net_features(helm_plan* plan, const void* v, void* features, int32_t nrhs, void* stream);

/* ---- spectral (S) operator: replaces WaveResidualOperator.setup /
 * incident / apply_operator / _iteration (nsrno/wave_operator.py:46-135) and
 * the warm start of solve_frequency (nsrno/wave_solver.py:50-55).  Plans of
 * op_kind HELM_OP_STAG8 only (pml_mm = 0; nz, nx = the padded box H, W;
 * h_mm = 1e3 dx; c0_m_s = the reference speed 1540). */
typedef struct {
    int32_t h, w;              /* medium size inside the box                 */
    int32_t pz, px;            /* offset of the medium in the box            */
    double x0_m, z0_m;         /* coordinates of medium node [0,0] [m]; box
                                  node [i,j] sits at ((j-px) h, (i-pz) h)
                                  + (x0_m, z0_m)                             */
    double rho0;               /* reference density (1000)                   */
    double alpha_scale;        /* clamp(exp(alpha_log), .01, 3): alpha [Np/m]
                                  = alpha_scale*100*ln(10)/20*f_MHz^1.1     */
    int32_t relaxation;        /* 0: power law; 1: discrete relaxation
                                  modulus (wave_operator.py:59-65)          */
    int32_t sponge_cells;      /* sponge width (48)                          */
    double sponge_amplitude;   /* sponge peak (2)                            */
    double eta_per_unit_alpha[2];  /* relaxation only                        */
    double taus_s[2];              /* relaxation only                        */
    double relax_dt_s;             /* relaxation time step (2.5e-9)          */
} helm_spectral_medium;

/* c, rho: device float32 [h][w] (the float medium of wave_operator.py:48);
 * replicate-padded into the box, a = rho0/rho, b = a (c0/c)^2, harmonic face
 * means ax, az (periodic), tissue mask, kappa_S with attenuation and sponge
 * (wave_operator.py:46-70), stored as A = k0^2 conj(L).  k0 = omega/c0, or
 * the relaxation form of wave_operator.py:55.  Rebuilds the multigrid
 * preconditioner (a 5-point shifted operator of the same coefficients, zero
 * Dirichlet ghosts) when one is set up. */
int32_t helm_set_medium_spectral(helm_plan* plan, const float* c, const float* rho,
                                 const helm_spectral_medium* medium, void* stream);
/* Plane-wave incident field (wave_operator.py:111-127), c128 [H][W]: a Hann
 * aperture of half width aperture_m times exp(-i k0 x sin(theta)) on row
 * z = 0, propagated to every row with the discrete dispersion k_z of the
 * staggered stencil (6 Newton steps, evanescent branch imaginary).
 * Needs helm_set_medium_spectral first (geometry and k0). */
int32_t helm_spectral_incident(helm_plan* plan, double sin_theta, double aperture_m, void* field,
                               void* stream);
/* No counterpart in the reference: helm_spectral_incident for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and output, bit
 * for bit; calls no hipMalloc/hipFree.  Its work buffer is a file-scope
 * static device array of 4 x 4096 complex doubles (256 KiB per GPU, held
 * until the process ends), so boxes wider than 4096 nodes are rejected with
 * HELM_ERR_ARG.  Every call on one GPU shares that GPU's buffer: calls that
 * overlap on one GPU (two streams, two threads) race and must be ordered by
 * the caller.  The first call on a device resolves the buffer's address
 * (hipGetSymbolAddress), which allocates nothing. */
int32_t helm_spectral_incident_noalloc(helm_plan* plan, double sin_theta, double aperture_m, void* field,
                                       void* stream);
/* Scattered-field right-hand side of the reference's solve, scaled for
 * helm_solve(adjoint = 1): rhs = -k0^2 (L - L0) u_inc, where L0 has
 * ax = az = 1 and kappa = 1 - i sponge (wave_operator.py:142-143,
 * wave_solver.py:53).  inc, rhs: c128 [nrhs][H][W]. */
int32_t helm_spectral_rhs(helm_plan* plan, const void* inc, void* rhs, int32_t nrhs, void* stream);

/* ---- conditioned split-PML operator: replaces FrequencyGeometry,
 * ConditionedFrequencySystem (nsrno/conditioned_wave.py:12-92) and
 * dense_wave.projected_operator / adjoint_operator / source_rhs
 * (nsrno/dense_wave.py:59-82).  Plans of op_kind HELM_OP_COND8 (pml_mm = 0;
 * nz, nx = the box H, W; h_mm = 1e3 dx; c0_m_s = 1540).  The plan holds the
 * reference's operator itself, in its convention (D = exp(+i omega dt)):
 *
 *   A p = M [ scalar sum_axes ( left D-(F) + right D-(g F) ) + mass M p ] + (1 - M) p,
 *   F = a_face sv D+(M p),  scalar = (D-1)/(D k0^2),  k0 = 2 sin(pi f dt)/(dt c0),
 *
 * with D+/D- the 8-tap staggered differences of the spectral operator above,
 * M the native-medium mask and the split-PML factors sv, g, left, right of
 * conditioned_wave.py:33-41,67-69.  helm_apply(adjoint = 0) is
 * projected_operator, adjoint = 1 adjoint_operator; helm_solve(adjoint = 0)
 * is solve_system's forward solve, adjoint = 1 its adjoint solve. */
typedef struct {
    int32_t h, w;              /* native medium size                         */
    int32_t pz, px;            /* offset of the medium in the box (>= 8,
                                  and >= 8 nodes of box beyond it)            */
    int32_t pml;               /* split-PML width [cells]; also the rim that
                                  is cropped and replicated (:46-49)          */
    double dt_s;               /* native time step                           */
    double rho0;               /* reference density (1000)                   */
    double eta_per_unit_alpha[2];
    double taus_s[2];
} helm_conditioned_medium;

/* c, rho, alpha: device float64 [h][w] native maps (alpha may be NULL: 0).
 * Crops and replicates the pml rim, pads into the box (:45-49), forms the
 * masked harmonic face means, the relaxation modulus and the mass (:56-66)
 * and the per-column/row split-PML factors.  Rebuilds the multigrid
 * preconditioner when one is set up: a 5-point shifted operator of the same
 * coefficients whose stretch is taken from the split-PML factors, applied to
 * the input scaled by k0^2 sx sz (identity outside the medium). */
int32_t helm_set_medium_conditioned(helm_plan* plan, const double* c, const double* rho, const double* alpha,
                                    const helm_conditioned_medium* medium, void* stream);
/* Source term of dense_wave.source_rhs (:59-62): rhs[r] (c128 [nrhs][H][W],
 * written everywhere) = line[r] (c128 [nrhs][w]) on box row `face`, times
 * a c0^2 2 dt / (c_src dx (D - 1)), c_src = c at native (source_z, source_x)
 * after cropping.  c, rho: the maps given to helm_set_medium_conditioned. */
int32_t helm_conditioned_source(helm_plan* plan, const double* c, const double* rho, const void* line,
                                int32_t face, int32_t source_z, int32_t source_x, void* rhs, int32_t nrhs,
                                void* stream);
/* Material gradient of the implicit adjoint (replaces the autograd step of
 * dense_wave.ImplicitChannels.backward, nsrno/dense_wave.py:250-255):
 *   G(c, rho) = -Re sum_r <lam_r, M (A(c, rho) v_r - s_r(c, rho))>,
 * v = M (u + pref) (the library applies M), lam the adjoint fields, s the
 * source term of helm_conditioned_source for `line` (NULL: no source term);
 * v, lam c128 [nrhs][H][W], line c128 [nrhs][w].  grad_c, grad_rho: device
 * float64 [h][w], overwritten with dG/dc and dG/drho of the native maps
 * (through the face means, the relaxation modulus, the source term and the
 * crop/replicate of the rim); alpha is held fixed.  c, rho, alpha: the maps
 * given to helm_set_medium_conditioned. */
int32_t helm_conditioned_gradient(helm_plan* plan, const double* c, const double* rho, const double* alpha,
                                  const void* v, const void* lam, const void* line, int32_t nrhs, int32_t face,
                                  int32_t source_z, int32_t source_x, double* grad_c, double* grad_rho,
                                  void* stream);

/* ---- warm start: `steps` minimal-residual steps per right-hand side,
 * d = M^{-1} r, alpha = <A d, r>/|A d|^2, x += alpha d (the residual never
 * grows), as WaveResidualOperator._iteration (wave_operator.py:129-135) with
 * this library's preconditioner in place of the spectral one.  b, x: c128
 * [nrhs][N]; x holds the initial guess.  adjoint as in helm_solve. */
int32_t helm_warm_start(helm_plan* plan, int32_t precond, const void* b, void* x, int32_t nrhs,
                        int32_t steps, int32_t adjoint, void* stream);

/* ---- Born Jacobian at zero contrast: replaces the reciprocal field products
 * of nsrno/theranostic/wave.py:331-389 (Solver.born_jacobian).  fields c128
 * [n_nodes][n_src] (row = node), pairs (tx[p], rx[p]) int32 [n_pairs], the
 * material basis as CSC over nodes (node_ptr [n_basis+1], node_idx, node_w
 * c128 = h^2 k0^2 sz sx B) and over edges (edge_ptr [n_basis+1], edge_i,
 * edge_j (-1: ghost end), edge_w c128 = -stretch_e (avg B)_e); J c128
 * [n_pairs][2 n_basis] = [m_kappa columns | m_eta columns].  The caller
 * guarantees every index is in range (nodes < n_nodes, sources < n_src). */
int32_t helm_born_jacobian(int64_t n_nodes, int32_t n_src, const void* fields, int32_t n_pairs, const int32_t* tx,
                           const int32_t* rx, int32_t n_basis, const int32_t* node_ptr, const int32_t* node_idx,
                           const void* node_w, const int32_t* edge_ptr, const int32_t* edge_i,
                           const int32_t* edge_j, const void* edge_w, void* J, void* stream);

/* ---- time -> frequency capture: replaces accumulate_pressure of
 * nsrno/native_capture.cu:4-11 (the running DFT of the native time stepper's
 * pressure into complex128 frequency fields).  pressure float32 [snapshots][n]
 * (1 to 32 consecutive captures), phase c128 [frequencies][captures]
 * (exp(-2 pi i f t_c) / 2401 in the reference, native_capture.py:37-40),
 * first = capture index of pressure[0]; fields c128 [frequencies][n] +=
 * sum_k pressure[k] phase[f][first + k]. */
int32_t helm_capture_accumulate(const float* pressure, int32_t snapshots, const void* phase, int32_t first,
                                int32_t captures, int32_t frequencies, int64_t n, void* fields, void* stream);

/* ---- band RF synthesis: replaces the irfft(..., n=2401, norm='forward') of
 * WaveResidualOperator.predict_band_rf (nsrno/wave_operator.py:169-180).
 * spectra c64 [traces][n_bins] hold the spectrum at bins[b] (int32, distinct,
 * 0 <= bins[b] <= samples/2); every other bin is zero.  rf float32
 * [traces][samples] (samples <= 8192), summed in FP64. */
int32_t helm_band_rf(const void* spectra, const int32_t* bins, int32_t n_bins, int32_t traces, int32_t samples,
                     float* rf, void* stream);

/* ---- autotuning of launch shapes: no counterpart in the reference.  Times,
 * with HIP events on the plan's own grid and nrhs right-hand sides, the
 * FP32 stencil's streaming shapes (rows per run x right-hand sides per group,
 * streaming or block form)
 * and the multigrid grid-level workgroup target, and keeps the fastest in the
 * plan (later launches use them).  Call after the medium and the multigrid
 * are set; allocates two scratch fields for the duration of the call. */
typedef struct {
    int32_t stencil_rows;      /* rows per wavefront run of the FP32 stencil */
    int32_t stencil_rhs;       /* right-hand sides per group                 */
    int64_t mg_workgroups;     /* workgroups a multigrid launch aims for     */
    double stencil_us;         /* the chosen stencil's time per apply        */
    double vcycle_us;          /* the chosen V-cycle's time                  */
    int32_t stencil_block;     /* 1: the block form (whole run loaded first),
                                  2: the same in XCD-contiguous order,
                                  3: the paired block form (two columns per
                                  lane), 4: the same in XCD-contiguous order,
                                  5 / 6: the block / paired form in
                                  XCD-contiguous order with the RHS groups
                                  of a run next to each other */
} helm_tuning;
int32_t helm_autotune(helm_plan* plan, int32_t nrhs, helm_tuning* result, void* stream);

/* ---- fixed launch shape of the FP32 flux stencil: no counterpart in the
 * reference.  form 0: streaming (rows 2..1024 per run), 1: block (rows 2, 4,
 * 8 or 16 loaded up front), 2: block in XCD-contiguous workgroup order, 3:
 * paired block (two columns per lane; rows 2, 4 or 8; the block form where
 * the grid has an odd row length or the fields are not 16-byte aligned), 4:
 * paired block in XCD-contiguous order, 5 / 6: block / paired block in
 * XCD-contiguous order with the RHS groups of a run next to each other;
 * rhs 1, 2 or 4 per group (also 8 in forms 1-6).  rows = 0 returns to the
 * built-in default.  For tests and for callers that tuned once and reuse
 * the result. */
int32_t helm_set_stencil_shape(helm_plan* plan, int32_t rows, int32_t rhs, int32_t form);

/* ---- in-solve kernel timing: no counterpart in the reference (its path
 * has no device kernels); for benchmarks that need one kernel's duration
 * inside a running solve.  With timing on, helm_solve brackets every launch
 * of the operator apply, the preconditioner, the two Arnoldi passes and the
 * FP64 true-residual apply with HIP events on the solve's stream, and
 * accumulates per kernel the launches, the summed event time and the
 * algorithmic bytes (each input and output vector counted once). */
typedef struct {
    char name[32];
    int64_t launches;
    double total_ms;
    double bytes;
} helm_kernel_stat;
/* mask != 0 turns timing on for the kernels whose bit is set (bit k =
 * entry k of helm_profile_read: 0 operator apply, 1 preconditioner, 2 DCGS2
 * dots, 3 DCGS2 update, 4 true residual, 5 multigrid smoothing + residual +
 * restriction, 6 multigrid prolongation + smoothing, both on every level,
 * 7 and 8 the same two on the finest level only) and clears the totals; 0
 * turns it off.  Each timed launch adds an event pair to the stream, so time only the
 * kernels needed. */
int32_t helm_profile(helm_plan* plan, int32_t mask);
/* copies up to capacity entries (host array); *count = entries available */
int32_t helm_profile_read(helm_plan* plan, helm_kernel_stat* stats, int32_t capacity, int32_t* count);

#ifdef __cplusplus
}
#endif
#endif
