/*
 * edus_host.h -- C-ABI of libedus_host.so, the host side of the drop-in: it reads
 * the reference's input.txt, builds the crystal, k-grid, neighbour tables, model
 * matrices and laser schedule, drives the GPU engine (edus.h) through the time
 * loop and writes the reference's Output/<name>.txt observable files.
 *
 * Host code is C++ (as the reference is); these entry points exist so that other
 * languages (the Python mirror, the tests, bench.py) can call the same code.
 */
#ifndef EDUS_HOST_H
#define EDUS_HOST_H

#include <stdint.h>
#include "edus.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct edus_setup edus_setup;

/* Laser pulse in the reference's parametrisation (headers/Laser.h:1-51). */
typedef struct edus_pulse {
    int32_t gaussian, sin2, pw, envelope;
    double wl, period, intensity, A0, E0, t0, tf, w1, w2, ncycle, sigma, phase;
    double pol1[3], pol2[3];
} edus_pulse;

/* Scalars of a built setup. */
typedef struct edus_setup_info {
    int32_t nk0, nk1, nk2, nk;      /* grid (reference Nk[3]) and nk = product           */
    int32_t nb, nch, ncc, nbands[3];/* Ncv, Nb[0], Nb[0]+Nb[1], Nb[0..2]                  */
    int32_t n_nbr, n_shells;        /* gradient neighbours, shells (CalculateWeigths.h) */
    int32_t stepper, order;         /* EDUS_STEPPER_*, AB/Taylor order                    */
    int32_t it_resolution, i_current, i_tabs, n_pumps;
    int32_t model;                  /* 0 tight binding, 1 Wannier, 2 crystal interface     */
    int32_t n_diss;                 /* nonzero entries of T (main_MPI.cpp:66-79)          */
    int64_t itfi;                   /* last step index, loop runs it = 0..itfi            */
    double dt, t_fin, t_start;      /* a.u.                                               */
    double det_a, det_b, jacobian;  /* in-plane areas, |det M| (Coordinate.h:91)          */
    double j_norm, tabs_norm, nktot, wx2;
    double T1, T2, Tch;
    double shell_residual;          /* norm(A w - q) of the accepted shells               */
    int32_t ncut;                   /* Coulomb Fourier modes per direction (0: no Coulomb) */
    int32_t dyn_dt;                 /* 1: adaptive dt loop (dynamical_dt_evolution)       */
    double eps_step;                /* epsStepAbs of the adaptive loop                    */
    int32_t print_flags;            /* bit 0 PrintPopulation, bit 1 Print_new_band_dispersion */
    int32_t reserved2;
    double print_start, print_end, print_step;  /* snapshot window and spacing, a.u.   */
} edus_setup_info;

/* Parses `input_path` (reference grammar, headers/ReadInput.h:332-1234) and builds
 * every table.  Wannier seeds are resolved relative to the input file's directory.
 * flags: bit 0 = also verify the closed-form neighbour table against the
 * reference's brute-force search (O(Nk^2), small grids only); bit 1 = run the
 * Wannier contraction in FP32 (edus_wannier_hk_fp32), for mixed-precision runs. */
int edus_setup_from_input(const char *input_path, int32_t flags, edus_setup **out);
void edus_setup_free(edus_setup *s);
int edus_setup_get_info(const edus_setup *s, edus_setup_info *info);

/* Named arrays of the setup (read-only views, valid until edus_setup_free):
 *   "kpt"   [nk][3]            crystal k (BZ_parameters, InitParameters.h:24-38)
 *   "kcart" [nk][3]            cartesian k
 *   "a" "b" [3][3]             direct / reciprocal lattice vectors (rows, a.u.)
 *   "H" "U" "rho0"  [nk][nb][nb] complex
 *   "D" "gradH"     [nk][nb][nb][3] complex ("gradH" empty without Current)
 *   "T"     [nb][nb]
 *   "nbr_wb"[n_nbr][3]         weight * b-vector (cartesian)
 *   "bvec"  [n_nbr][3]         b-vectors (cartesian), "weights" [n_shells]
 *   "taylor_tau_inv" [(order-1)^2]
 *   "coul_A" .. "coul_D" [ncut^2][nb^2] complex, "coul_screen" [nb^2], "coul_vh" [nb][nb]
 * `count` receives the number of doubles (or int32 for "nbr"). */
const double *edus_setup_array(const edus_setup *s, const char *name, int64_t *count);
const int32_t *edus_setup_nbr(const edus_setup *s, int64_t *count);

/* Pulses: index 0..n_pumps-1 are the pumps, index n_pumps is the probe. */
int edus_setup_pulse(const edus_setup *s, int32_t index, edus_pulse *out);

/* Tight-binding model of the setup: name (NUL-terminated, up to len bytes) and
 * params[8] = a, g, b, t2, phi0, core, rchv, rchc (a.u.).  Returns 1 for a
 * Wannier setup (params untouched), 0 for tight binding. */
int edus_setup_model(const edus_setup *s, char *name, int32_t len, double *params);

/* Evaluates the setup's tight-binding model at n arbitrary crystal k points:
 * H, U [n][nb][nb] complex, D [n][nb][nb][3] complex (TightBinding.h energy,
 * unitary, dipole).  Not available for Wannier setups. */
int edus_setup_eval_tb(const edus_setup *s, const double *kcrys, int32_t n, double *H, double *U, double *D);

/* Fills an engine problem with views into the setup (edus_create copies them). */
int edus_setup_problem(const edus_setup *s, edus_problem *p);

/* Field schedule of steps [it0, it0+nsteps): per step `nfe` evaluations (3 for RK4,
 * 1 otherwise) of EDUS_FIELDS_PER_EVAL doubles, plus the E_non0 flag of each step,
 * with the reference's time accumulation (time_loop += dt). */
int edus_setup_fields(const edus_setup *s, int64_t it0, int64_t nsteps, double *fields, uint8_t *grad_on);

/* Time of step `it` exactly as the reference accumulates it. */
double edus_setup_time(const edus_setup *s, int64_t it);

/* Runs the full time loop (const_dt_evolution.cpp) on the GPU engine and writes
 * Output/{EF,Losses,J1,J2,TransientAbs}.txt and Output/k_grid.txt under `outdir`.
 * max_steps < 0 runs to itfi.  Returns the number of steps done or < 0. */
int64_t edus_run(const edus_setup *s, const char *outdir, int64_t max_steps, int32_t verbose);

/* Writes a wannier90 <seed>_tb.dat with seeded random Hermitian H(R), r(R) on a
 * hexagonal lattice (|R1|,|R2| <= rmax), H(R) ~ sigma_h exp(-|R|/decay) eV and
 * r(R) ~ sigma_r exp(-|R|/decay) Angstrom, H(-R) = H(R)^+.  a_ang: lattice
 * constant.  Returns 0 on success. */
int edus_write_synthetic_tb(const char *path, int32_t nw, int32_t rmax, double a_ang,
                            double sigma_h_ev, double sigma_r_ang, double decay_ang, uint64_t seed);

const char *edus_host_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* EDUS_HOST_H */
