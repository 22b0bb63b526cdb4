/*
 * edus.h -- C-ABI of libedus_hip.so, the MI355X (gfx950) engine that advances the
 * semiconductor Bloch equations of EDUS on one GPU with the whole k-grid resident
 * in HBM.
 *
 * The reference has no FFI: its hot path is a set of C++ free functions called by
 * every OpenMP thread inside one parallel region.  Each entry point below names the
 * reference function(s) it replaces (paths relative to the reference checkout).
 *
 * Conventions (all entry points):
 *   - plain pointers and sizes; complex numbers are interleaved (re, im) doubles;
 *   - host arrays use the reference layouts: density matrix P[k][i][j] = rho_ji
 *     ("transposed", headers/InitialPopulation.h:18-35), Hamiltonian/Unitary
 *     [k][i][j], Dipole and GradientEnergy [k][i][j][xyz];
 *   - one host thread per context; every call is stream-ordered on the context's
 *     own HIP stream and returns when its results are visible to the host;
 *   - return 0 on success, a negative EDUS_ERR_* code otherwise; the message of
 *     the last failure on the calling thread is returned by edus_last_error().
 *   - no exceptions cross the ABI.
 */
#ifndef EDUS_H
#define EDUS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDUS_OK               0
#define EDUS_ERR_INVALID     -1  /* bad argument, shape or table */
#define EDUS_ERR_HIP         -2  /* HIP runtime failure (message has the hipError_t) */
#define EDUS_ERR_NOMEM       -3  /* device allocation failed */
#define EDUS_ERR_UNSUPPORTED -4  /* configuration this library does not implement */
#define EDUS_ERR_NODEVICE    -5  /* no gfx950 device visible */

#define EDUS_STEPPER_RK4     0
#define EDUS_STEPPER_AB      1
#define EDUS_STEPPER_TAYLOR  2

/* Arithmetic of the stage kernel.  FP64 is the reference's (std::complex<double>
 * throughout).  MIXED keeps the state, the stencil drift and every increment in
 * FP64 but stores H, D and U as FP32 and forms the commutator and dephasing
 * products in FP32 (relative RHS error ~1e-6); band counts 5..32 only.  No
 * reference counterpart: an option for long runs where FP32 operators suffice. */
#define EDUS_PRECISION_FP64  0
#define EDUS_PRECISION_MIXED 1
/* MIXED32: MIXED plus FP32 storage of what only feeds a right-hand side or an
 * FP64 sum: the RK4 stage inputs (y_n + c dt k, read by stages 2-4 and their
 * stencils) and the Adams-Bashforth / Taylor derivative histories.  The state at
 * t_n, the RK4 accumulator and every update stay FP64 (the master copy).  Band
 * counts 5..32; not for batches.  No reference counterpart. */
#define EDUS_PRECISION_MIXED32 2
/* BF16X3: MIXED with the commutator and dephasing products taken as BF16x3 split
 * products on the bf16 matrix cores: each FP32 operand x is split into
 * hi = bf16(x) and lo = bf16(x - hi), every real product is lo*hi + hi*lo + hi*hi
 * with FP32 accumulation (a product carries at most about 3 2^-16 of relative
 * error, against 2^-24 per FP32 multiply).  State, stencil drift and increments
 * stay FP64.  Band counts 9..32 (the matrix-core tiles); not for batches.  No
 * reference counterpart. */
#define EDUS_PRECISION_BF16X3 3

#define EDUS_MAX_NB          32  /* largest band count with a kernel in this library */
#define EDUS_MAX_ORDER        8  /* largest Adams-Bashforth / Taylor order */

typedef struct edus_ctx edus_ctx;

/* Hartree-Fock mean field as a truncated separable Fourier series of V(k-k')
 * (Coulomb.h:273-541, coefficients as read by Fourier_Coulomb.cpp:505-622), in
 * the Wannier basis or, with diag_basis, in the eigenbasis.  Sample in the yz
 * plane (grid 1 x n1 x n2). */
typedef struct edus_coulomb {
    int32_t ncut;            /* Fourier modes per direction (Coulomb_set.Ncut)       */
    int32_t diag_basis;      /* Coulomb_diag_basis: sums over P in the eigenbasis of U,
                                equilibrium the filled bands (RungeKutta.h:41-55, 97-124) */
    const double *A, *B, *C, *D;  /* [ncut*ncut][nb*nb] complex                      */
    const double *screen;    /* [nb*nb] complex Screen_const[0][*]                   */
    const double *v_hartree; /* [nb][nb] complex V_Hartree                           */
} edus_coulomb;

/*
 * Everything the time-stepping path reads, in reference layout.  Built on the host
 * by the setup code (headers/InitParameters.h, CalculateWeigths.h, TightBinding.h,
 * Wannier.h, Source_Main/save_HUD.cpp, runge_kutta_init.cpp) and uploaded once.
 * Pointers are only read during edus_create().
 */
typedef struct edus_problem {
    int32_t nk;          /* k points (the whole grid, nk = Nk0*Nk1*Nk2)               */
    int32_t nb;          /* bands Ncv = Nb[0]+Nb[1]+Nb[2]                               */
    int32_t nch;         /* core bands Nb[0] (probe field / RWA selection)              */
    int32_t ncc;         /* filled bands Nb[0]+Nb[1] (hole source, Losses holes)        */
    int32_t n_nbr;       /* gradient neighbours per k, all shells (12 for 2D hexagonal) */
    int32_t stepper;     /* EDUS_STEPPER_*                                               */
    int32_t order;       /* Adams-Bashforth / Taylor order (ignored for RK4)             */
    int32_t precision;   /* EDUS_PRECISION_FP64 (the reference's arithmetic) or _MIXED  */
    const int32_t *nbr;  /* [nk][n_nbr]  GradientIndex[ik][shell][b], shells in order    */
    const double  *nbr_wb; /* [n_nbr][3] Weigths[shell] * Bvector[shell][b][xyz]        */
    const double  *H;    /* [nk][nb][nb] complex  Hamiltonian                            */
    const double  *D;    /* [nk][nb][nb][3] complex Dipole, symmetric or Hermitian in ij */
    const double  *U;    /* [nk][nb][nb] complex  Unitary (eigenbasis rows)              */
    const double  *T;    /* [nb][nb] real decoherence rates (Initialize_T)               */
    const double  *rho0; /* [nk][nb][nb] complex  initial P0                             */
    const double  *gradH;/* [nk][nb][nb][3] complex GradientEnergy, or NULL (no Current) */
    double wx2;          /* probe carrier (pulse2.wl), RWA term RungeKutta.h:209          */
    double dt;           /* time step (a.u.)                                              */
    double j_norm;       /* dk0*dk1*dk2/det_a        (Observables_MPI.h:698,757)          */
    double tabs_norm;    /* |det M|*dk0*dk1*dk2       (Observables_MPI.h:568)             */
    double nktot;        /* Nk0*Nk1*Nk2               (Losses and TAbs normalisation)     */
    const double *taylor_tau_inv; /* [(order-1)^2] row-major tau^{-1} for EDUS_STEPPER_TAYLOR */
    const edus_coulomb *coulomb;  /* NULL: no mean-field term (nb <= 4, grid 1 x n1 x n2)    */
    int32_t nk_dims[3];           /* grid sizes (Nk0, Nk1, Nk2), needed with coulomb         */
    int32_t reserved1;
} edus_problem;

/* Field schedule entry of one step: for every field evaluation of the step the
 * pump sum E_pump[xyz] followed by 0.5*E_probe[xyz] (RungeKutta.cpp:24-30).
 * RK4 uses 3 evaluations per step (t, t+dt/2, t+dt; stage 3 reuses the second),
 * Adams-Bashforth and Taylor use 1 (t_n).  `grad_on` is the E_non0 flag of the
 * step (RungeKutta.cpp:151-162; Linear_Multistep_ODE_solvers.cpp:14-26). */
#define EDUS_FIELDS_PER_EVAL 6

/* Creates a context on the current HIP device and uploads the problem.
 * Replaces: allocate_arr.cpp, the per-thread copies main_MPI.cpp:143-227,
 * CalculateIndicesMPI/Organize_kspace (one device holds the whole grid). */
int edus_create(const edus_problem *p, edus_ctx **out);

/* Frees every device buffer of the context. */
int edus_destroy(edus_ctx *ctx);

/* Advances `nsteps` RK4 steps.  fields: [nsteps][3][EDUS_FIELDS_PER_EVAL],
 * grad_on: [nsteps].  Replaces Source_Main/RungeKutta.cpp with
 * get_derivative_Df (headers/RungeKutta.h:7-297), Runge_Kutta_Ad/Ac
 * (RungeKutta.h:301-355) and Mpi_communication (MPIcommunication.h:1-20). */
int edus_step_rk4(edus_ctx *ctx, int32_t nsteps, const double *fields, const uint8_t *grad_on);

/* The field schedule of a whole run, uploaded once and kept on the device:
 * nsteps RK4 steps laid out as for edus_step_rk4 (fields [nsteps][3][6],
 * grad_on [nsteps]).  Resets the table position to its first step.  Part of the
 * captured-step path below; replaces the per-stage field evaluation of
 * RungeKutta.cpp:24-30 feeding get_derivative_Df.  RK4 contexts. */
int edus_set_field_table(edus_ctx *ctx, int32_t nsteps, const double *fields, const uint8_t *grad_on);

/* Advances `nsteps` RK4 steps reading their fields from the table, from its
 * current position on: one RK4 step (RungeKutta.cpp:11-137, the four fused
 * stage launches, Coulomb passes included) is captured once into a HIP graph
 * that reads its rows through a device step counter, then replayed once per
 * step.  The results are bit-identical to edus_step_rk4 with the same fields.
 * Fails with EDUS_ERR_INVALID past the end of the table. */
int edus_step_rk4_graph(edus_ctx *ctx, int32_t nsteps);

/* Advances `nsteps` Adams-Bashforth steps of the context's order.
 * fields: [nsteps][1][EDUS_FIELDS_PER_EVAL].  Replaces
 * Source_Main/Linear_Multistep_ODE_solvers.cpp with ODE_solver.h:66-98,152-179. */
int edus_step_ab(edus_ctx *ctx, int32_t nsteps, const double *fields, const uint8_t *grad_on);

/* Advances `nsteps` Taylor multistep steps of the context's order (constant dt,
 * tau^{-1} given at create).  Replaces Source_Main/Taylor_DE_solver_deriv_Polynom.cpp
 * with ODE_solver.h:3-33,107-123. */
int edus_step_taylor(edus_ctx *ctx, int32_t nsteps, const double *fields, const uint8_t *grad_on);

/* One step of the adaptive-dt loop with step size `dt`: RK4 for an RK4 context
 * (fields [3][6] at t, t+dt/2, t+dt), Taylor for a Taylor context (fields [6] at t,
 * `taylor_tau_inv` [(order-1)^2] for the current dt history).  Then the step-size
 * measure of Source_Main/dyn_dt_evolution.cpp:74-100,
 *   max_dp = max_{k,i} 2 |P_ii(before) - P_ii(after)| / (cond + |P_ii(after)| + 1e-7),
 * with cond = P_cond_max + n_cond from the last observables.  Replaces the step and
 * the max-reduction of dyn_dt_evolution.cpp (with RungeKutta.cpp:1-9,139-147 and
 * Taylor_DE_solver_deriv_Polynom.cpp:75-84); the dt update stays with the caller. */
int edus_step_dyn(edus_ctx *ctx, double dt, const double *fields, uint8_t grad_on, const double *taylor_tau_inv,
                  double cond, double *max_dp);

/* Observable reductions of the current state, FP64 with Kahan-compensated
 * block sums.  Replaces PrintLossesMPI_exciton (Observables_MPI.h:217-484, without
 * Coulomb), Current1 (668-718), Current2 (726-775) and CalculateTransientAbs (559-569).
 *   pop    [nb]  sum_k Re diag(U P U^+) / nktot   (before the hole transform 1-x;
 *                with the Coulomb term, U diagonalises the renormalised H)
 *   j1     [6]   (re,im) x,y,z of sum_k sum_ij P_ij gradH_ij * j_norm
 *   j2     [6]   (re,im) x,y,z of sum_k P_ij H_ik conj(D_jk) * j_norm
 *   tabs   [6]   (re,im) x,y,z of -2 sum_k sum_{i<nch<=j} D_ij P_ij * tabs_norm
 *   pminmax[2]   min over k,bands of the (hole) population, max |population|
 */
int edus_observables(edus_ctx *ctx, double *pop, double *j1, double *j2, double *tabs, double *pminmax);

/* Right-hand side dP/dt of the current state for one field evaluation
 * (field6 = pump sum xyz, half probe xyz) and gradient flag, written hermitised
 * in reference layout [nk][nb][nb].  Replaces get_derivative_Df
 * (headers/RungeKutta.h:7-297) as a standalone call. */
int edus_rhs(edus_ctx *ctx, const double *field6, uint8_t grad_on, double *pv);

/* Element (ic, jc) of the current state in the eigenbasis, conj(U) P U^T per k
 * (get_P_in_dia_vect, InitParameters.h:264-335; the diagonal forced real), as the
 * k-resolved snapshots of if_resolution.cpp:232-273 print it.  out [nk] complex.
 * U is the setup's (bare) eigenbasis, also with the Coulomb term. */
int edus_eigen_element(edus_ctx *ctx, int32_t ic, int32_t jc, double *out);

/* Eigenvalues (ascending) of the renormalised H per k at the last
 * edus_observables call, as PrintLossesMPI_exciton stores them for the Edia
 * files (Observables_MPI.h:286-304, 440-477).  Coulomb contexts only; out [nk][nb]. */
int edus_renorm_energies(edus_ctx *ctx, double *out);

/* Copies the current density matrix to/from the host, reference layout [nk][nb][nb].
 * Setting the state leaves an Adams-Bashforth or Taylor context's derivative
 * history as it was: a new trajectory of those steppers starts from a fresh
 * context, whose history is zero as at the reference's t_start. */
int edus_get_state(edus_ctx *ctx, double *rho);
int edus_set_state(edus_ctx *ctx, const double *rho);

/* Device time of the last edus_step_* call's kernels, in milliseconds, measured
 * with HIP events on the context's stream; *nlaunch receives the kernel count. */
int edus_last_step_time(edus_ctx *ctx, double *ms, int32_t *nlaunch);

/* The stage kernels' launch schedule of the context, out[5]: the padded tile of
 * the wave kernels (8, 16, 32; 0 for the thread-per-k kernel of 1..4 bands), the
 * strip width of their workgroup walk (0 = k order), 1 when a workgroup serves a
 * stencil pair of k (2q, 2q + 1: every pair's 12-point stencils overlap alike,
 * as on a uniform grid; the BF16x3 form and the FP32-storage Taylor step at 32
 * bands still run one k per workgroup), the stencil rows the second k of a pair
 * fetches itself, and the threads per workgroup of the thread-per-k kernel.  No
 * reference counterpart (a launch diagnostic). */
int edus_launch_schedule(edus_ctx *ctx, int32_t *out);

/* Picks the stage kernels' launch schedule by measurement on the context's own
 * data and keeps it for every later launch: for 5..32 bands the strip width of
 * the workgroup walk over the grid (0 = k order, or 16..256 columns narrower than
 * a grid row), for 1..4 bands the threads per workgroup (64, 128, 256).  Each
 * candidate runs `reps` right-hand-side launches with the field `field6` (as for
 * edus_rhs) into scratch memory; the state and the step results are unaffected
 * (every candidate forms the same sums for every k).  `chosen` (may be NULL)
 * receives the strip width or the workgroup size kept.  No reference
 * counterpart: the reference has no launch schedule. */
int edus_autotune(edus_ctx *ctx, const double *field6, uint8_t grad_on, int32_t reps, int32_t *chosen);

/*
 * Batches: n independent RK4 problems (different grids, band counts, models and
 * pulses; no Coulomb term) advanced together.  No reference counterpart: the
 * reference runs one problem per process.  This is for scans over laser
 * parameters and materials, where one small problem cannot fill the GPU.  Each
 * RK4 stage is one launch per kernel class present (small kernels by band
 * count, wave kernels by tile and precision).  A member steps bit-identically
 * to the same problem in its own context.
 */
typedef struct edus_batch edus_batch;
int edus_batch_create(const edus_problem *const *problems, int32_t n, edus_batch **out);
int edus_batch_destroy(edus_batch *b);
int edus_batch_size(const edus_batch *b);
/* fields[i] and grad_on[i]: member i's arrays, laid out as for edus_step_rk4. */
int edus_batch_step_rk4(edus_batch *b, int32_t nsteps, const double *const *fields, const uint8_t *const *grad_on);
int edus_batch_last_step_time(const edus_batch *b, double *ms, int32_t *nlaunch);
/* Member i's context, for edus_observables, edus_get_state, edus_set_state and
 * edus_rhs; it is owned by the batch (do not destroy it).  NULL if out of range. */
edus_ctx *edus_batch_member(edus_batch *b, int32_t i);

/* Eigen-decomposition of nmat Hermitian nb x nb matrices H [nmat][nb][nb] complex
 * (row-major, upper triangle read) for the setup's U tables: w [nmat][nb]
 * ascending eigenvalues, U [nmat][nb][nb] complex with row a = conj(v_a), each
 * eigenvector phased so that its largest-modulus component is real positive.
 * Replaces the per-k eig_sym of Wannier.h:496-542 (save_HUD.cpp:142-210). */
int edus_eigh_batched(int32_t nmat, int32_t nb, const double *H, double *w, double *U);

/* Parameters of the Coulomb coefficient generator (Source_Main/Fourier_Coulomb.cpp:95-500
 * with headers/Coulomb.h:60-267 and integrationWeight.h:5-76), "011" orientation:
 * the sample lies in the plane of crystal directions 1 and 2, q = b1 q1 + b2 q2. */
typedef struct edus_coulomb_gen {
    int32_t nb, ncut;
    int32_t n_sum;          /* Simpson points per direction of the main integral (reference: 4999) */
    int32_t n_peak;         /* points per direction of the q ~ 0 correction (reference: 39999)  */
    int32_t rytova_keldysh; /* 1: V = 2 pi / (eps sqrt(q^2+qTF^2) (|q| r0 + 1)); 0: without (|q| r0 + 1) */
    int32_t n_g;            /* lattice vectors G of the sum                                        */
    const int32_t *g;       /* [n_g][2] integer offsets of G along crystal directions 1 and 2     */
    double b1[3], b2[3];    /* cartesian reciprocal vectors of those directions                     */
    double eps_area;        /* epsilon_static times the in-plane cell area (Fourier_Coulomb.cpp:2) */
    double qtf, r0;         /* Thomas-Fermi wave vector, Rytova-Keldysh length (a.u.)              */
    double q_peak;          /* radius of the q ~ 0 correction: max(qTF, 0.02)                      */
    const double *disp;     /* [nb][3] orbital positions: real diagonal of D at k = (0, 0.3, 0.1)  */
    int32_t verbose;        /* progress lines on stderr                                            */
    int32_t reserved;
} edus_coulomb_gen;

/* The mean-field coefficients: A..D [ncut^2][nb^2], screen [nb^2] (Screen_const[0])
 * and vh [nb][nb] (V_Hartree), all complex, as Fourier_Coulomb.cpp:95-500 computes
 * and then writes them (A_coeff, B_coeff, C_coeff, D_coeff, Screen_const_,
 * V_Hartree_ files).  The lattice sum over G for each of the n_sum^2 Simpson points
 * runs on the GPU.  Independent of any context; uses the current HIP device. */
int edus_coulomb_coefficients(const edus_coulomb_gen *g, double *A, double *B, double *C, double *D, double *screen, double *vh);

/* Wannier phase contraction on the GPU: H(k) = sum_R e^{2 pi i k.R} H(R) and
 * r(k) likewise for the three position components, for every k of a grid
 * k = -0.5 + i/N (crystal units).  Replaces Wannier::energy and dipolex/y/z
 * (headers/Wannier.h:278-304, 336-471) as driven by save_HUD.cpp:19-68.
 *   nk3[3]   grid sizes (Nk0, Nk1, Nk2)
 *   R        [nR][3] integer lattice vectors (crystal)
 *   HR       [nR][nw][nw] complex   H_mn(R) in a.u. (already divided by ndegen)
 *   rR       [nR][nw][nw][3] complex r_mn(R) in a.u. (already divided by ndegen)
 *   H_out    [nk][nw][nw] complex, diagonal imaginary part zeroed (Wannier.h:489-492)
 *   D_out    [nk][nw][nw][3] complex, (D + D^T)/2 with zero diagonal imaginary part
 *            (Wannier.h:577-578, save_HUD.cpp:47-51); may be NULL
 * Phases are exact: k.R = -R/2 + (i R mod N)/N is reduced in integers. */
int edus_wannier_hk(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R,
                    const double *HR, const double *rR, double *H_out, double *D_out);

/* The same contraction with FP32 products on the matrix cores and FP32 sums
 * (the phases are formed in FP64 and rounded once, as are H(R) and r(R)); the
 * outputs are FP64 arrays holding the FP32 results.  The mixed-precision form of
 * Wannier::energy / dipolex/y/z (headers/Wannier.h:278-304, 336-471); each
 * element carries at most (2 nR + 4) 2^-24 sum_R |X(R)_mn| of error. */
int edus_wannier_hk_fp32(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R,
                         const double *HR, const double *rR, double *H_out, double *D_out);

/* The same contraction as BF16x3 split products on the bf16 matrix cores: each
 * FP32-rounded operand x is split into hi = bf16(x) and lo = bf16(x - hi) and every
 * real product is taken as hi*hi + hi*lo + lo*hi with FP32 accumulation (phases
 * formed in FP64).  The BF16x3 form of Wannier::energy / dipolex/y/z
 * (headers/Wannier.h:278-304, 336-471); each element carries at most
 * sqrt(2) (3 2^-18 + (2 nR + 8) 2^-24) sum_R |X(R)_mn| of error (the split keeps 16
 * of FP32's 24 mantissa bits).  No reference counterpart as a form: for setups
 * where that accuracy suffices. */
int edus_wannier_hk_bf16x3(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R, const double *HR, const double *rR,
                           double *H_out, double *D_out);

/* Size of the fixed device array behind edus_wannier_hk_noalloc (one copy per
 * device, held from the library's load until the process ends). */
#define EDUS_NOALLOC_SCRATCH_BYTES ((int64_t)128 << 20)

/* edus_wannier_hk without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * device buffer (lattice vectors, phase tables, H(R) and r(R), output chunks)
 * is carved from the library's fixed scratch array; the call allocates no device
 * memory and uses no allocating library routine.  Same arguments and the same
 * results as edus_wannier_hk, bit for bit.
 * Limit: 12 nR + 16 (Nk0+Nk1+Nk2)(nR+1) + 64 nR nw^2 bytes, plus one k of
 * output (128 nw^2 bytes with D_out, 64 nw^2 without) and 1 KiB of alignment,
 * must fit EDUS_NOALLOC_SCRATCH_BYTES (a 32-band model with 961 lattice vectors
 * needs 79 MB); a larger model returns EDUS_ERR_UNSUPPORTED before touching the
 * device.  The grid is walked in as many k chunks as the rest of the array
 * holds.  Calls run on the null stream and share the device's one array, so
 * calls that overlap on one device race. */
int edus_wannier_hk_noalloc(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R, const double *HR, const double *rR,
                            double *H_out, double *D_out);

/* Streaming rates of the current device measured with this library's own
 * kernels (16-byte accesses, grid-stride, the best of several workgroups-per-CU
 * choices, each the mean of `reps` launches after one untimed): copy moves
 * 2 x bytes per launch, triad a = b + s c 3 x bytes, read 1 x bytes; GB/s in
 * 1e9 bytes.  No reference counterpart: the streaming ceiling the stage kernels'
 * measured bandwidth is read against, and the known-size kernel the HBM
 * counters are calibrated with.  Allocates three arrays of `bytes` bytes. */
int edus_stream_probe(int64_t bytes, int32_t reps, double *copy_gbs, double *triad_gbs, double *read_gbs);

/* Message of the last failure on the calling thread ("" if none). */
const char *edus_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* EDUS_H */
