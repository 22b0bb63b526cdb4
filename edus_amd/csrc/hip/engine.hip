// libedus_hip.so -- the MI355X engine of the EDUS time-stepping path.
//
// Device layout (all FP64):
//   * The density matrix is exactly Hermitian with a real diagonal at every step
//     of the reference's three steppers (the right-hand side is hermitised,
//     RungeKutta.h:283-294, and every update is a real-linear combination), so a
//     k point is stored packed: nb real diagonal entries, then (re, im) of the
//     strict upper triangle in row-major pair order -> nb*nb doubles.  This is
//     lossless and halves the HBM traffic of the reference's full complex arrays.
//   * nb <= 4: "SoA" -- element e of k at e*nkp + k (one thread per k, coalesced).
//     nb >= 5: "AoS" -- k-major, nb*nb contiguous doubles (one wavefront per k).
//   * H is packed like the state (upper triangle, as the RHS reads it);
//     D components are packed by their upper triangle, the lower one being D_ij
//     (symmetric tables) or conj(D_ij) (Hermitian tables); U is full complex.
//
// Kernels: fused RHS + stage update per RK4 stage / Adams-Bashforth step /
// Taylor step (RungeKutta.h:7-355, RungeKutta.cpp, Linear_Multistep_ODE_solvers.cpp,
// Taylor_DE_solver_deriv_Polynom.cpp), wave-per-k observable reductions with
// Kahan-compensated partials (Observables_MPI.h:217-775) and the Wannier phase
// contraction (Wannier.h:278-471).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/edus.h"

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string &m) {
    g_err = m;
    return code;
}

#define HIP_TRY(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return set_err(e_ == hipErrorOutOfMemory ? EDUS_ERR_NOMEM : EDUS_ERR_HIP,                \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                       \
    } while (0)

enum Mode { RK1 = 0, RK2 = 1, RK3 = 2, RK4S = 3, ABM = 4, TAY = 5, RHSO = 6 };

constexpr int kMaxNbr = 64;
constexpr int kMaxTinv = (EDUS_MAX_ORDER - 1) * (EDUS_MAX_ORDER - 1);

// Per-context constants of the right-hand side.
struct ConstArgs {
    const double *H;     // packed (layout of the state)
    const double *D;     // 3 packed components (upper triangle; lower = D_ij or conj(D_ij), see dherm)
    const double *U;     // full complex AoS [k][nb][nb] (dissipation)
    const double *T;     // [nb][nb]
    const float *Hf, *Df, *Uf;  // mixed precision: H, D, U in FP32 (same layouts)
    const int *nbr;      // [n_nbr][nkp]
    const int *eidx;     // packed element e -> i | j << 8 | comp << 16 (wave layout)
    int nk, nkp, nb, nch, ncc, n_nbr, diss, dherm;  // dherm: D_ji = conj(D_ij) (else D_ji = D_ij)
    int sw, scol;        // wave kernels: strip width and row length of the workgroup order (sw = 0: k order)
    // Wave kernels of one k per workgroup: pair = 1 when every pair of k (2q, 2q+1)
    // has the same stencil overlap (a uniform grid of even length along the
    // fastest axis); a workgroup then serves both and loads each shared
    // neighbour once.  pat[n] (n < n_nbr): the slot of 2q+1's stencil that is
    // 2q's slot n, or -1; pat[kMaxNbr + t] (t < npnew): 2q+1's slots found in
    // neither 2q's stencil nor 2q itself; pself: 2q+1's slot holding 2q, or -1.
    int pair, pself, npnew;
    const int *pat;
    double wx2;
    // Coulomb mean field (small-nb kernel only): the mean-field matrix per k,
    // peak correction against PW0, renormalised H out.
    int coul, N2, ncut, nu;
    double n_cond;
    const double *cxm, *sxm;   // [ncut][N1] cos/sin(2 pi m k1)
    const double *Tt;          // [2][ncut][nu][N2] complex
    const double *Xw;          // [2 nu][nkp] mean-field matrix (upper elements, re/im)
    const double *PW0;         // packed state at t = 0
    const double *EH;          // [nb] Hartree shift (minus equilibrium)
    const double *screen;      // [nb*nb] complex
    double *Hr;                // packed renormalised H (written)
    int coul_diag;             // Coulomb_diag_basis: the sums ran on PD, corrections in the eigenbasis
    const double *PD;          // packed state in the eigenbasis, conj(U) P U^T (diagonal basis)
};

// Per-launch stage data.  `tab` points at the stage's row of the field table:
// [0..2] pump sum, [3..5] half probe, [6..6+n_nbr) gradient coefficients
// c_n = sum_x E_pump[x] * w_shell b_n[x], [6+n_nbr] gradient flag.
struct StageArgs {
    const int *ctr;      // device step counter (field table path), or null: the row is `tab` itself
    long long ctr_stride;  // doubles per counted step of the field table
    const double *Y;     // stage input (neighbours gathered from it)
    const double *P0;    // RK: state at t_n
    double *ACC;         // RK: accumulator
    double *OUT;         // stage output
    double *HIST;        // AB/Taylor history slots
    const double *tab;
    int slot_out, order;
    int dmask;           // bit x: some field of the call has a nonzero x component (else D_x is not read)
    int pad1;
    int slots[EDUS_MAX_ORDER];
    double coef[EDUS_MAX_ORDER];
    double tinv[kMaxTinv];
    double dtp[EDUS_MAX_ORDER];
    double ca, cb, dt;
};

// The field-table row of a stage: `tab`, advanced by the device step counter when
// the launch belongs to a captured step (its arguments are fixed at capture, the
// counter moves on between replays).
__device__ __forceinline__ const double *stage_row(const StageArgs &a) {
    return a.ctr ? a.tab + (size_t)(*a.ctr) * (size_t)a.ctr_stride : a.tab;
}

// ---------------------------------------------------------------- complex helpers
struct cd {
    double x, y;
};
__device__ __forceinline__ cd mk(double a, double b) { return cd{a, b}; }
__device__ __forceinline__ cd cmul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cconj(cd a) { return cd{a.x, -a.y}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.x * s, a.y * s}; }

// Butterfly reductions over the 64 lanes of a wave (every lane gets the result;
// the order of the additions is fixed).
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Single-precision complex for the mixed-precision stage (operators and LDS products).
struct cf {
    float x, y;
};
template <typename R> struct CplxOf;
template <> struct CplxOf<double> { using T = cd; };
template <> struct CplxOf<float> { using T = cf; };

// Block index remapped so that each of the 8 XCDs (which receive workgroups
// round-robin) works on one contiguous range of k: the 12-point gradient stencil
// then finds its neighbours in the XCD's own L2.
__device__ __forceinline__ int xcd_block() {
    const int G = gridDim.x, b = blockIdx.x;
    const int x = b & 7, idx = b >> 3, per = G >> 3, rem = G & 7;
    return x * per + min(x, rem) + idx;
}

// k of the wave kernels' workgroup L when the grid is walked in strips: the
// nk / scol rows of scol points (k = row * scol + col) are cut into strips of
// sw columns (the last one narrower if sw does not divide scol), and each strip
// is walked row by row.  A strip's rows then meet their stencil neighbours on
// the rows above and below while those are still in L2, where whole rows of
// large matrices would not fit.  A bijection of [0, nk); sw = 0 keeps k = L.
constexpr long long kStripL2Bytes = 1 << 20;

__device__ __forceinline__ int strip_k(int L, int nk, int sw, int scol) {
    if (sw == 0) return L;
    const int nrow = nk / scol, nfull = scol / sw, per = nrow * sw, s = L / per;
    if (s < nfull) {
        const int u = L - s * per;
        return (u / sw) * scol + s * sw + u % sw;
    }
    const int wr = scol - nfull * sw, u = L - nfull * per;
    return (u / wr) * scol + nfull * sw + u % wr;
}

// Pair index of (i<j) in row-major upper-triangle order.
__host__ __device__ __forceinline__ int pair_index(int i, int j, int nb) { return i * nb - i * (i + 1) / 2 + (j - i - 1); }

// Streaming accesses: operators, accumulators, histories and stage outputs are
// touched once per launch, so they bypass the caches' normal retention and leave
// L2 to the stage input, whose 12 stencil neighbours are re-read from it.
__device__ __forceinline__ double ldnt(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ldnt(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stnt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void stnt(float *p, float v) { __builtin_nontemporal_store(v, p); }

// Two adjacent packed elements in one access (16 bytes in FP64, 8 in FP32), when
// `vec` (the pair lies inside the matrix and the array offset is even); else two
// single accesses at o0 and o1.  Halves the vector memory instructions of the
// wave kernels, whose address unit is their busiest memory stage.
template <typename T> struct Pair;
template <> struct Pair<double> { typedef double V __attribute__((ext_vector_type(2))); };
template <> struct Pair<float> { typedef float V __attribute__((ext_vector_type(2))); };
template <bool NT, typename T>
__device__ __forceinline__ void ld2(const T *p, size_t o0, size_t o1, bool vec, T &a, T &b) {
    if (vec) {
        const typename Pair<T>::V *q = reinterpret_cast<const typename Pair<T>::V *>(p + o0);
        const typename Pair<T>::V v = NT ? __builtin_nontemporal_load(q) : *q;
        a = v.x;
        b = v.y;
    } else {
        a = NT ? ldnt(p + o0) : p[o0];
        b = NT ? ldnt(p + o1) : p[o1];
    }
}
template <typename T>
__device__ __forceinline__ void st2(T *p, size_t o0, size_t o1, bool vec, T a, T b) {
    if (vec) {
        typename Pair<T>::V v;
        v.x = a;
        v.y = b;
        __builtin_nontemporal_store(v, reinterpret_cast<typename Pair<T>::V *>(p + o0));
    } else {
        stnt(p + o0, a);
        stnt(p + o1, b);
    }
}

// ---------------------------------------------------------------- stage update
// The elementwise tail shared by all kernels.  `kv` is one packed double of the
// right-hand side, `y` the same double of the stage input, `o` its offset in the
// state arrays and `h` the stride between history slots.
// S: storage type of the Adams-Bashforth / Taylor history (double, or float in
// the mixed form with FP32 stage storage); the state and its update stay FP64.
template <int MODE, typename S = double>
__device__ __forceinline__ void stage_update(const StageArgs &a, size_t o, size_t h, double kv, double y) {
    S *const hist = reinterpret_cast<S *>(a.HIST);
    if constexpr (MODE == RK1) {
        stnt(a.ACC + o, y + kv * a.ca);
        stnt(a.OUT + o, y + kv * a.cb);
    } else if constexpr (MODE == RK2 || MODE == RK3) {
        stnt(a.ACC + o, ldnt(a.ACC + o) + kv * a.ca);
        stnt(a.OUT + o, ldnt(a.P0 + o) + kv * a.cb);
    } else if constexpr (MODE == RK4S) {
        stnt(a.OUT + o, ldnt(a.ACC + o) + kv * a.ca);
    } else if constexpr (MODE == ABM) {
        // the order-1 history loads are issued together (unrolled to the largest
        // order, 5) before the sum, which keeps the order of the coefficients
        stnt(hist + (size_t)a.slot_out * h + o, (S)kv);
        double hv[5];
#pragma unroll
        for (int s = 1; s < 5; s++) hv[s] = (s < a.order) ? (double)ldnt(hist + (size_t)a.slots[s] * h + o) : 0.0;
        double delta = a.coef[0] * kv;
#pragma unroll
        for (int s = 1; s < 5; s++)
            if (s < a.order) delta += a.coef[s] * hv[s];
        stnt(a.OUT + o, y + a.dt * delta);
    } else if constexpr (MODE == RHSO) {
        a.OUT[o] = kv;
    } else {  // TAY
        stnt(hist + (size_t)a.slot_out * h + o, (S)kv);
        const int n = a.order - 1;
        double dp[EDUS_MAX_ORDER];
#pragma unroll
        for (int i = 0; i < EDUS_MAX_ORDER - 1; i++) dp[i] = (i < n) ? kv - (double)ldnt(hist + (size_t)a.slots[i + 1] * h + o) : 0.0;
        double delta = kv * a.dt;
        for (int i = 1; i <= n; i++) {
            double al = 0.0;
            for (int p = 0; p < n; p++) al += a.tinv[(i - 1) * n + p] * dp[p];
            delta += (a.dtp[i] * al) / (double)(i + 1);
        }
        stnt(a.OUT + o, y + delta);
    }
}

// stage_update of the Adams-Bashforth mode for two adjacent packed elements
// (o1 = o0 + 1 when vec): the same operations in the same order, with the
// history and state accessed as pairs.  (The Taylor mode keeps the single form:
// its runtime-order loops over two elements' differences cost 60 registers.)
template <typename S>
__device__ __forceinline__ void stage_update2_abm(const StageArgs &a, size_t o0, size_t o1, bool vec, size_t h, double kv0,
                                                  double kv1, double y0, double y1) {
    S *const hist = reinterpret_cast<S *>(a.HIST);
    const size_t so = (size_t)a.slot_out * h;
    st2(hist, so + o0, so + o1, vec, (S)kv0, (S)kv1);
    double h0[5], h1[5];
#pragma unroll
    for (int s = 1; s < 5; s++) {
        h0[s] = h1[s] = 0.0;
        if (s < a.order) {
            const size_t sl = (size_t)a.slots[s] * h;
            S u0, u1;
            ld2<true>(hist, sl + o0, sl + o1, vec, u0, u1);
            h0[s] = (double)u0;
            h1[s] = (double)u1;
        }
    }
    double d0 = a.coef[0] * kv0, d1 = a.coef[0] * kv1;
#pragma unroll
    for (int s = 1; s < 5; s++)
        if (s < a.order) {
            d0 += a.coef[s] * h0[s];
            d1 += a.coef[s] * h1[s];
        }
    st2(a.OUT, o0, o1, vec, y0 + a.dt * d0, y1 + a.dt * d1);
}

// The Taylor tail for two adjacent packed elements: the history is read as
// pairs, then each element's update is formed on its own (one set of
// differences live at a time), in the operation order of stage_update.
template <typename S>
__device__ __forceinline__ void stage_update2_tay(const StageArgs &a, size_t o0, size_t o1, bool vec, size_t h, double kv0,
                                                  double kv1, double y0, double y1) {
    S *const hist = reinterpret_cast<S *>(a.HIST);
    const size_t so = (size_t)a.slot_out * h;
    st2(hist, so + o0, so + o1, vec, (S)kv0, (S)kv1);
    const int n = a.order - 1;
    S u0[EDUS_MAX_ORDER - 1], u1[EDUS_MAX_ORDER - 1];
#pragma unroll
    for (int i = 0; i < EDUS_MAX_ORDER - 1; i++) {
        u0[i] = u1[i] = (S)0;
        if (i < n) {
            const size_t sl = (size_t)a.slots[i + 1] * h;
            ld2<true>(hist, sl + o0, sl + o1, vec, u0[i], u1[i]);
        }
    }
    double out[2];
#pragma unroll
    for (int w = 0; w < 2; w++) {
        const double kv = w ? kv1 : kv0;
        double dp[EDUS_MAX_ORDER];
#pragma unroll
        for (int i = 0; i < EDUS_MAX_ORDER - 1; i++) dp[i] = (i < n) ? kv - (double)(w ? u1[i] : u0[i]) : 0.0;
        double delta = kv * a.dt;
        for (int i = 1; i <= n; i++) {
            double al = 0.0;
            for (int q = 0; q < n; q++) al += a.tinv[(i - 1) * n + q] * dp[q];
            delta += (a.dtp[i] * al) / (double)(i + 1);
        }
        out[w] = (w ? y1 : y0) + delta;
    }
    st2(a.OUT, o0, o1, vec, out[0], out[1]);
}

// ================================================================ nb <= 4: thread per k
// DISS: the context has dephasing terms (c.diss); without them the U sandwich
// and its registers are compiled out.
template <int NB, int MODE, bool DISS>
__device__ __forceinline__ void stage_small(const ConstArgs &c, const StageArgs &a, const int k) {
    constexpr int NP = NB * NB;
    if (k >= c.nk) return;
    const size_t nkp = c.nkp;
    const double *tab = stage_row(a);
    const double Ep[3] = {tab[0], tab[1], tab[2]}, Eq[3] = {tab[3], tab[4], tab[5]};
    const int nn = c.n_nbr;
    const bool grad = nn > 0 && tab[6 + nn] != 0.0;
    // The first stencil indices are requested with the state, so the gather at
    // the end waits for one round trip, not two.
    constexpr int kPreNbr = NB <= 3 ? 12 : 0;  // a multiple of the batch size below
    int kpre[kPreNbr > 0 ? kPreNbr : 1];
    if (grad) {
#pragma unroll
        for (int n = 0; n < kPreNbr; n++) kpre[n] = c.nbr[(size_t)min(n, nn - 1) * nkp + k];
    }

    double y[NP], pre0[NP], pre1[NP];
#pragma unroll
    for (int e = 0; e < NP; e++) y[e] = a.Y[e * nkp + k];
    // the RK stage tails' reads of the accumulator and the step's base state,
    // requested with the stage input (up to 3 bands; at 4 the registers they
    // take cost more occupancy than the early request gains)
    constexpr bool kPreRK = NB <= 3;
#pragma unroll
    for (int e = 0; e < NP; e++) {
        pre0[e] = pre1[e] = 0.0;
        if constexpr (kPreRK && (MODE == RK2 || MODE == RK3)) {
            pre0[e] = ldnt(a.ACC + e * nkp + k);
            pre1[e] = ldnt(a.P0 + e * nkp + k);
        } else if constexpr (kPreRK && MODE == RK4S) {
            pre0[e] = ldnt(a.ACC + e * nkp + k);
        }
    }
    cd P[NB][NB], M[NB][NB];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        P[i][i] = mk(y[i], 0.0);
#pragma unroll
        for (int j = i + 1; j < NB; j++) {
            const int p = pair_index(i, j, NB);
            P[i][j] = mk(y[NB + 2 * p], y[NB + 2 * p + 1]);
            P[j][i] = cconj(P[i][j]);
        }
    }
    // Renormalised H with the Hartree-Fock mean field (RungeKutta.h:69-138):
    // X_W[i][j] = -sum_mn (X_mn)_{ji} trig(k), minus Screen_const * (P - P0)_{ji}
    // when n_cond > 1e-20, real diagonal, plus the Hartree shift.
    cd Hren[NB][NB];
    if (c.coul && c.coul_diag) {
        // Diagonal basis (RungeKutta.h:97-124): X_d from the sums over PD, minus
        // Screen_const[0][0] (PD - P_Bloch_0) on the diagonal (P_Bloch_0: the
        // filled bands) and conj(Screen_const[0][1]) PD_ji off it; back to the
        // orbital basis as U^+ X_d U (InitParameters.h:340-370).
        cd Xd[NB][NB], U[NB][NB];
        const double *uk = c.U + (size_t)k * NB * NB * 2;
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) {
                double re, im;
                ld2<false>(uk, (i * NB + j) * 2, (i * NB + j) * 2 + 1, true, re, im);
                U[i][j] = mk(re, im);
            }
        const cd scr0 = mk(c.screen[0], c.screen[1]), scr1c = mk(c.screen[2], -c.screen[3]);
        int u = 0;
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = i; j < NB; j++, u++) {
                cd xk = mk(c.Xw[(size_t)(2 * u) * nkp + k], c.Xw[(size_t)(2 * u + 1) * nkp + k]);
                if (i == j) {
                    xk = csub(xk, cscale(scr0, c.PD[i * nkp + k] - (i < c.ncc ? 1.0 : 0.0)));
                    Xd[i][i] = xk;
                } else {
                    const int p = pair_index(i, j, NB);
                    // PD_ji = conj(PD_ij)
                    const cd pdji = mk(c.PD[(NB + 2 * p) * nkp + k], -c.PD[(NB + 2 * p + 1) * nkp + k]);
                    xk = csub(xk, cmul(scr1c, pdji));
                    Xd[i][j] = xk;
                    Xd[j][i] = cconj(xk);
                }
            }
        cd T1[NB][NB];
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) {
                cd t = mk(0.0, 0.0);
#pragma unroll
                for (int l = 0; l < NB; l++) t = cadd(t, cmul(Xd[i][l], U[l][j]));
                T1[i][j] = t;
            }
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = i; j < NB; j++) {
                cd xw = mk(0.0, 0.0);
#pragma unroll
                for (int l = 0; l < NB; l++) xw = cadd(xw, cmul(cconj(U[l][i]), T1[l][j]));
                cd h;
                if (i == j) {
                    h = mk(c.H[i * nkp + k] + xw.x + c.EH[i], 0.0);
                    c.Hr[i * nkp + k] = h.x;
                } else {
                    const int p = pair_index(i, j, NB);
                    h = mk(c.H[(NB + 2 * p) * nkp + k] + xw.x, c.H[(NB + 2 * p + 1) * nkp + k] + xw.y);
                    c.Hr[(NB + 2 * p) * nkp + k] = h.x;
                    c.Hr[(NB + 2 * p + 1) * nkp + k] = h.y;
                }
                Hren[i][j] = h;
            }
    } else if (c.coul) {
        int u = 0;
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = i; j < NB; j++, u++) {
                cd xk = mk(c.Xw[(size_t)(2 * u) * nkp + k], c.Xw[(size_t)(2 * u + 1) * nkp + k]);
                if (c.n_cond > 1e-20) {
                    const int idx = (i != j) ? i * NB + j : 0;
                    cd scr = mk(c.screen[2 * idx], -c.screen[2 * idx + 1]);  // conj(Screen_const)
                    cd dp;
                    if (i == j) dp = mk(y[i] - c.PW0[i * nkp + k], 0.0);
                    else {
                        const int p = pair_index(i, j, NB);
                        // (P - P0)_{ji} = conj((P - P0)_{ij})
                        dp = mk(y[NB + 2 * p] - c.PW0[(NB + 2 * p) * nkp + k], -(y[NB + 2 * p + 1] - c.PW0[(NB + 2 * p + 1) * nkp + k]));
                    }
                    xk = csub(xk, cmul(scr, dp));
                }
                cd h;
                if (i == j) {
                    h = mk(c.H[i * nkp + k] + xk.x + c.EH[i], 0.0);
                    c.Hr[i * nkp + k] = h.x;
                } else {
                    const int p = pair_index(i, j, NB);
                    h = mk(c.H[(NB + 2 * p) * nkp + k] + xk.x, c.H[(NB + 2 * p + 1) * nkp + k] + xk.y);
                    c.Hr[(NB + 2 * p) * nkp + k] = h.x;
                    c.Hr[(NB + 2 * p + 1) * nkp + k] = h.y;
                }
                Hren[i][j] = h;
            }
    }
    // M = H + E.D with the probe on core<->non-core elements (RungeKutta.h:165-168).
#pragma unroll
    for (int i = 0; i < NB; i++) {
#pragma unroll
        for (int j = i; j < NB; j++) {
            const bool w = (i >= c.nch && j < c.nch) || (i < c.nch && j >= c.nch);
            const double E[3] = {w ? Eq[0] : Ep[0], w ? Eq[1] : Ep[1], w ? Eq[2] : Ep[2]};
            cd h, dx[3];
            if (i == j) {
                h = mk(ldnt(c.H + i * nkp + k), 0.0);
#pragma unroll
                for (int x = 0; x < 3; x++) dx[x] = (a.dmask >> x & 1) ? mk(ldnt(c.D + (x * NP + i) * nkp + k), 0.0) : mk(0.0, 0.0);
            } else {
                const int p = pair_index(i, j, NB);
                h = mk(ldnt(c.H + (NB + 2 * p) * nkp + k), ldnt(c.H + (NB + 2 * p + 1) * nkp + k));
#pragma unroll
                for (int x = 0; x < 3; x++)
                    dx[x] = (a.dmask >> x & 1) ? mk(ldnt(c.D + (x * NP + NB + 2 * p) * nkp + k), ldnt(c.D + (x * NP + NB + 2 * p + 1) * nkp + k))
                                               : mk(0.0, 0.0);
            }
            if (c.coul) h = Hren[i][j];
            cd mij = h, mji = cconj(h);
#pragma unroll
            for (int x = 0; x < 3; x++) {
                mij = cadd(mij, cscale(dx[x], E[x]));
                mji = cadd(mji, cscale(c.dherm ? cconj(dx[x]) : dx[x], E[x]));
            }
            M[i][j] = mij;
            M[j][i] = mji;
        }
    }
    // -i[M, rho] in the transposed storage (RungeKutta.h:158-212).
    double pv[NP];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        cd beta = mk(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < NB; j++) beta = cadd(beta, cmul(M[i][j], P[i][j]));
        pv[i] = 2. * beta.y;
    }
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = i + 1; j < NB; j++) {
            cd beta = mk(0.0, 0.0);
#pragma unroll
            for (int l = 0; l < NB; l++) beta = cadd(beta, csub(cmul(M[j][l], P[i][l]), cconj(cmul(M[i][l], P[j][l]))));
            cd v = mk(beta.y, -beta.x);                      // -i * beta
            if (i < c.nch && j >= c.nch) v = cadd(v, mk(-c.wx2 * P[i][j].y, c.wx2 * P[i][j].x));  // RWA term
            const int p = pair_index(i, j, NB);
            pv[NB + 2 * p] = v.x;
            pv[NB + 2 * p + 1] = v.y;
        }
    // Dephasing in the eigenbasis (RungeKutta.h:219-254) as
    // Pv -= conj(U) (T o (conj(U) P U^T)) conj(U) plus the hole source.
    if constexpr (DISS) {
        cd U[NB][NB];
        const double *u = c.U + (size_t)k * NB * NB * 2;
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) {  // one 16-byte load per complex entry (the k's U is 16-byte aligned)
                double re, im;
                ld2<true>(u, (i * NB + j) * 2, (i * NB + j) * 2 + 1, true, re, im);
                U[i][j] = mk(re, im);
            }
        cd G[NB][NB];
#pragma unroll
        for (int ia = 0; ia < NB; ia++)
#pragma unroll
            for (int ib = 0; ib < NB; ib++) {
                cd g = mk(0.0, 0.0);
#pragma unroll
                for (int i = 0; i < NB; i++) {
                    cd t = mk(0.0, 0.0);
#pragma unroll
                    for (int j = 0; j < NB; j++) t = cadd(t, cmul(U[ib][j], P[i][j]));
                    g = cadd(g, cmul(cconj(U[ia][i]), t));
                }
                G[ia][ib] = cscale(g, c.T[ia * NB + ib]);
            }
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = i; j < NB; j++) {
                cd w = mk(0.0, 0.0);
#pragma unroll
                for (int ia = 0; ia < NB; ia++) {
                    cd t = mk(0.0, 0.0);
#pragma unroll
                    for (int ib = 0; ib < NB; ib++) t = cadd(t, cmul(G[ia][ib], cconj(U[ib][j])));
                    w = cadd(w, cmul(cconj(U[i][ia]), t));
                }
                cd hole = mk(0.0, 0.0);
#pragma unroll
                for (int v = 0; v < NB; v++)  // static indices keep U in registers
                    if (v < c.ncc) hole = cadd(hole, cscale(cmul(cconj(U[v][j]), U[v][i]), c.T[v * NB + v]));
                if (i == j) {
                    pv[i] = pv[i] - w.x + hole.x;
                } else {
                    const int p = pair_index(i, j, NB);
                    pv[NB + 2 * p] = pv[NB + 2 * p] - w.x + hole.x;
                    pv[NB + 2 * p + 1] = pv[NB + 2 * p + 1] - w.y + hole.y;
                }
            }
    }
    // Field drift E.grad_k rho on the shell stencil (RungeKutta.h:258-276).  The
    // neighbour indices were requested at the start; the rows come GB neighbours
    // at a time, all in flight together, and the sum keeps the stencil order.
    if (grad) {
        constexpr int GB = NB <= 2 ? 6 : NB == 3 ? 4 : 2;
        double g[NP];
#pragma unroll
        for (int e = 0; e < NP; e++) g[e] = 0.0;
        auto batch = [&](int n0, const int (&kk)[GB]) {
            double v[GB][NP];
#pragma unroll
            for (int j = 0; j < GB; j++)
#pragma unroll
                for (int e = 0; e < NP; e++) v[j][e] = a.Y[e * nkp + kk[j]];
#pragma unroll
            for (int j = 0; j < GB; j++)
                if (n0 + j < nn) {
                    const double cn = tab[6 + n0 + j];
#pragma unroll
                    for (int e = 0; e < NP; e++) g[e] += cn * v[j][e];
                }
        };
#pragma unroll
        for (int n0 = 0; n0 < kPreNbr; n0 += GB) {
            if (n0 >= nn) break;
            int kk[GB];
#pragma unroll
            for (int j = 0; j < GB; j++) kk[j] = kpre[n0 + j];
            batch(n0, kk);
        }
        for (int n0 = kPreNbr; n0 < nn; n0 += GB) {
            int kk[GB];
#pragma unroll
            for (int j = 0; j < GB; j++) kk[j] = c.nbr[(size_t)min(n0 + j, nn - 1) * nkp + k];
            batch(n0, kk);
        }
#pragma unroll
        for (int e = 0; e < NP; e++) pv[e] += g[e];
    }
#pragma unroll
    for (int e = 0; e < NP; e++) {
        const size_t o = e * nkp + k;
        if constexpr (kPreRK && (MODE == RK2 || MODE == RK3)) {
            stnt(a.ACC + o, pre0[e] + pv[e] * a.ca);
            stnt(a.OUT + o, pre1[e] + pv[e] * a.cb);
        } else if constexpr (kPreRK && MODE == RK4S) {
            stnt(a.OUT + o, pre0[e] + pv[e] * a.ca);
        } else {
            stage_update<MODE>(a, o, (size_t)NP * nkp, pv[e], y[e]);
        }
    }
}

template <int NB, int MODE, bool DISS>
__global__ __launch_bounds__(256) void k_stage_small(ConstArgs c, StageArgs a) {
    stage_small<NB, MODE, DISS>(c, a, xcd_block() * blockDim.x + threadIdx.x);
}

// ================================================================ Coulomb mean field
// P in the eigenbasis for the Coulomb_diag_basis form (get_P_in_dia_vect,
// InitParameters.h:264-335): PD = conj(U) P U^T per k in the transposed storage,
// diagonal real, packed like the state (thread per k, nb <= 4).
template <int NB>
__global__ __launch_bounds__(256) void k_coul_pdiag(const double *__restrict__ Y, const double *__restrict__ Uall, int nk, int nkp,
                                                    double *__restrict__ PD) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    cd P[NB][NB], U[NB][NB];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        P[i][i] = mk(Y[(size_t)i * nkp + k], 0.0);
#pragma unroll
        for (int j = i + 1; j < NB; j++) {
            const int p = pair_index(i, j, NB);
            P[i][j] = mk(Y[(size_t)(NB + 2 * p) * nkp + k], Y[(size_t)(NB + 2 * p + 1) * nkp + k]);
            P[j][i] = cconj(P[i][j]);
        }
    }
    const double *uk = Uall + (size_t)k * NB * NB * 2;
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) U[i][j] = mk(uk[(i * NB + j) * 2], uk[(i * NB + j) * 2 + 1]);
#pragma unroll
    for (int a = 0; a < NB; a++)
#pragma unroll
        for (int b = a; b < NB; b++) {
            cd acc = mk(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < NB; i++) {
                cd t = mk(0.0, 0.0);
#pragma unroll
                for (int j = 0; j < NB; j++) t = cadd(t, cmul(P[i][j], U[b][j]));
                acc = cadd(acc, cmul(cconj(U[a][i]), t));
            }
            if (a == b) {
                PD[(size_t)a * nkp + k] = acc.x;
            } else {
                const int p = pair_index(a, b, NB);
                PD[(size_t)(NB + 2 * p) * nkp + k] = acc.x;
                PD[(size_t)(NB + 2 * p + 1) * nkp + k] = acc.y;
            }
        }
}

// Separable partial 2D Fourier sums of the packed state over the grid
// k = (i1, i2) (Coulomb.h:273-350):
//   S_y[i1][n][e] = sum_i2 y_n(i2) P_e(i1, i2)                     (y = cos, sin)
//   F_xy[mm][e]   = w (1 + corr_mn) sum_i1 x_m(i1) S_y[i1][n][e]    (4 combinations)
// then X_mn from A..D minus the equilibrium (Coulomb.h:413-470) and the column
// sums of the reconstruction (Coulomb.h:510-541), the row sum being done per k
// inside the RHS kernel.
struct CoulArgs {
    const double *Y;
    int nk, nkp, nb, NP, N1, N2, ncut, nu;
    double w;
    const double *cxm, *sxm, *cyn, *syn;
    double *S, *F;
    double *X;                 // [4][ncut^2][nb][nb] complex
    const double *X0;
    const double *A, *B, *C, *D, *vh;  // complex
    double *EH;
    const double *EH0;
    double *Tt;                // [2][ncut][nu][N2] complex
    double *Xw;                // [2 nu][nkp]: mean-field matrix per k (upper elements, re/im)
};

constexpr int kCoulThreads = 256;

// S_y[i1][n][e] = sum_i2 y_n(i2) P_e(i1, i2) as one product on the FP64 matrix
// cores (a thread per output with a sequential sum took 28 us at 300 x 300,
// bound by the address throughput of its uncoalesced loads): rows r = e N1 + i1
// (the state row Y[e][i1 N2 + i2], contiguous in i2), columns c = q ncut + n
// (q = 0: cos, 1: sin), K = N2.  One wave per 16 x 16 tile walks K four at a
// time with v_mfma_f64_16x16x4f64 (lane l feeds A[row l%16][k l/16] and
// B[k l/16][col l%16] and ends with C[l/16 + 4 i][l%16]); out-of-range rows,
// columns and k are fed zeros.
__global__ __launch_bounds__(64) void k_coul_s1_mfma(CoulArgs a) {
    typedef double v4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int nrow = a.NP * a.N1, ncol = 2 * a.ncut;
    const int ra = r0 + (lane & 15), cb = c0 + (lane & 15), kq = lane >> 4;
    const bool rok = ra < nrow, cok = cb < ncol;
    const double *yrow = a.Y + (rok ? (size_t)(ra / a.N1) * a.nkp + (size_t)(ra % a.N1) * a.N2 : 0);
    const int q = cok ? cb / a.ncut : 0, n = cok ? cb % a.ncut : 0;
    const double *trig = (q ? a.syn : a.cyn) + (size_t)n * a.N2;
    v4 acc = {0.0, 0.0, 0.0, 0.0};
    // K in blocks of 32: the block's 16 loads (at a clamped k, always in
    // bounds) are all issued before its 8 products; values past the edges are
    // zeroed after loading.
    for (int k0 = 0; k0 < a.N2; k0 += 32) {
        double yv[8], tv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int kc = min(k0 + 4 * u + kq, a.N2 - 1);
            yv[u] = yrow[kc];
            tv[u] = trig[kc];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const bool kok = k0 + 4 * u + kq < a.N2;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64((rok && kok) ? yv[u] : 0.0, (cok && kok) ? tv[u] : 0.0, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int r = r0 + (lane >> 4) + 4 * i, c = c0 + (lane & 15);
        if (r < nrow && c < ncol) {
            const int e = r / a.N1, i1 = r % a.N1, qq = c / a.ncut, nn = c % a.ncut;
            a.S[(((size_t)qq * a.N1 + i1) * a.ncut + nn) * a.NP + e] = acc[i];
        }
    }
}

// F_xy[mm][e] = w (1 + corr_mn) sum_i1 x_m(i1) S_y[i1][n][e]: one workgroup per
// mm; thread t takes element e = t % NP and every (256/NP)-th term of the sum,
// and the parts are added in a fixed order per element.  The same workgroup then
// forms X_mn from its F (coul_x_elem) and, for mm = 0, the Hartree shift: X_mn
// needs no other mode's F, so no launch of its own.
constexpr int kCoulMaxNP = 16;  // the mean field runs at nb <= 4

__device__ __forceinline__ void coul_x_elem(const CoulArgs &a, const double (*Fs)[kCoulMaxNP], int mm, int u);

__global__ __launch_bounds__(kCoulThreads) void k_coul_s2(CoulArgs a) {
    __shared__ double red[4][kCoulThreads];
    __shared__ double Fs[4][kCoulMaxNP];
    const int nmm = a.ncut * a.ncut, mm = blockIdx.x, m = mm / a.ncut, n = mm % a.ncut, t = threadIdx.x;
    const int NP = a.NP, nparts = kCoulThreads / NP, e = t % NP, part = t / NP;
    double cc = 0, cs = 0, sc = 0, ss = 0;
    if (part < nparts) {
        for (int i1 = part; i1 < a.N1; i1 += nparts) {
            const double xc = a.cxm[m * a.N1 + i1], xs = a.sxm[m * a.N1 + i1];
            const double yc = a.S[(((size_t)0 * a.N1 + i1) * a.ncut + n) * NP + e];
            const double ys = a.S[(((size_t)1 * a.N1 + i1) * a.ncut + n) * NP + e];
            cc += xc * yc;
            cs += xc * ys;
            sc += xs * yc;
            ss += xs * ys;
        }
    }
    red[0][t] = cc;
    red[1][t] = cs;
    red[2][t] = sc;
    red[3][t] = ss;
    __syncthreads();
    if (t < NP) {
        const double corr = -4 * M_PI * M_PI * (m * m + n * n) / ((double)a.nk * 27);
        const double f = a.w * (1. + corr);
        for (int qd = 0; qd < 4; qd++) {
            double v = 0.0;
            for (int q = 0; q < nparts; q++) v += red[qd][q * NP + t];
            a.F[((size_t)qd * nmm + mm) * NP + t] = v * f;
            Fs[qd][t] = v * f;
        }
    }
    __syncthreads();
    if (mm == 0 && t < a.nb) {  // Hartree shift from mode 0 (Coulomb.h:385-410)
        double e = 0.;
        for (int jc = 0; jc < a.nb; jc++) e += Fs[0][jc] * a.vh[2 * (t * a.nb + jc)];
        a.EH[t] = e - a.EH0[t];
    }
    if (t < a.nu) coul_x_elem(a, Fs, mm, t);
}

// X_mn for upper element u of mode mm from the mode's F (Coulomb.h:413-470).
__device__ __forceinline__ void coul_x_elem(const CoulArgs &a, const double (*Fs)[kCoulMaxNP], int mm, int u) {
    const int nb = a.nb, nmm = a.ncut * a.ncut, nb2 = nb * nb;
    int i = 0, j = 0;
    for (int q = u, r = 0; ; r++) {
        if (q < nb - r) { i = r; j = r + q; break; }
        q -= nb - r;
    }
    auto fget = [&](int qd) -> cd {
        const double *F = Fs[qd];
        if (i == j) return mk(F[i], 0.0);
        const int p = pair_index(i, j, nb);
        return mk(F[nb + 2 * p], F[nb + 2 * p + 1]);
    };
    const cd fcc = fget(0), fcs = fget(1), fsc = fget(2), fss = fget(3);
    const int d = i * nb + j, od = j * nb + i;
    const size_t base = (size_t)mm * nb2;
    const cd *A = reinterpret_cast<const cd *>(a.A), *B = reinterpret_cast<const cd *>(a.B);
    const cd *C = reinterpret_cast<const cd *>(a.C), *D = reinterpret_cast<const cd *>(a.D);
    const cd *X0 = reinterpret_cast<const cd *>(a.X0);
    cd *X = reinterpret_cast<cd *>(a.X);
    const cd Amn = A[base + (i == j ? 0 : d)], Dmn = D[base + (i == j ? 0 : d)];
    cd xcc = csub(cadd(cmul(Amn, fcc), cmul(Dmn, fss)), X0[0 * (size_t)nmm * nb2 + base + d]);
    cd xcs = csub(csub(cmul(Amn, fcs), cmul(Dmn, fsc)), X0[1 * (size_t)nmm * nb2 + base + d]);
    cd xsc = csub(csub(cmul(Amn, fsc), cmul(Dmn, fcs)), X0[2 * (size_t)nmm * nb2 + base + d]);
    cd xss = csub(cadd(cmul(Amn, fss), cmul(Dmn, fcc)), X0[3 * (size_t)nmm * nb2 + base + d]);
    if (i != j) {
        const cd Bmn = B[base + d], Cmn = C[base + d];
        xcc = cadd(xcc, cadd(cmul(Bmn, fcs), cmul(Cmn, fsc)));
        xcs = cadd(xcs, csub(cmul(Cmn, fss), cmul(Bmn, fcc)));
        xsc = cadd(xsc, csub(cmul(Bmn, fss), cmul(Cmn, fcc)));
        xss = csub(xss, cadd(cmul(Bmn, fsc), cmul(Cmn, fcs)));
        X[0 * (size_t)nmm * nb2 + base + od] = cconj(xcc);
        X[1 * (size_t)nmm * nb2 + base + od] = cconj(xcs);
        X[2 * (size_t)nmm * nb2 + base + od] = cconj(xsc);
        X[3 * (size_t)nmm * nb2 + base + od] = cconj(xss);
    }
    X[0 * (size_t)nmm * nb2 + base + d] = xcc;
    X[1 * (size_t)nmm * nb2 + base + d] = xcs;
    X[2 * (size_t)nmm * nb2 + base + d] = xsc;
    X[3 * (size_t)nmm * nb2 + base + d] = xss;
}

// Column sums of the reconstruction for upper element u = (i <= j), which reads
// (X_mn)_{ji} because the reference stores X_{jc,ic} (Coulomb.h:537).
__global__ void k_coul_s3(CoulArgs a) {
    const int nb = a.nb, nmm = a.ncut * a.ncut, nb2 = nb * nb;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.N2 * a.ncut * a.nu) return;
    const int i2 = t % a.N2, u = (t / a.N2) % a.nu, m = t / (a.N2 * a.nu);
    int i = 0, j = 0;
    for (int q = u, r = 0; ; r++) {
        if (q < nb - r) { i = r; j = r + q; break; }
        q -= nb - r;
    }
    const cd *X = reinterpret_cast<const cd *>(a.X);
    cd tc = mk(0, 0), ts = mk(0, 0);
    // n in blocks of 4: the block's loads (at a clamped n) are issued together,
    // the sums keep the order of n.
    for (int n0 = 0; n0 < a.ncut; n0 += 4) {
        cd xv[4][4];
        double ycv[4], ysv[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int n = min(n0 + b, a.ncut - 1);
            const size_t d = (size_t)(m * a.ncut + n) * nb2 + j * nb + i;
#pragma unroll
            for (int qd = 0; qd < 4; qd++) xv[b][qd] = X[qd * (size_t)nmm * nb2 + d];
            ycv[b] = a.cyn[n * a.N2 + i2];
            ysv[b] = a.syn[n * a.N2 + i2];
        }
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (n0 + b < a.ncut) {
                tc = cadd(tc, cadd(cscale(xv[b][0], ycv[b]), cscale(xv[b][1], ysv[b])));
                ts = cadd(ts, cadd(cscale(xv[b][2], ycv[b]), cscale(xv[b][3], ysv[b])));
            }
    }
    cd *T = reinterpret_cast<cd *>(a.Tt);
    T[(((size_t)0 * a.ncut + m) * a.nu + u) * a.N2 + i2] = tc;
    T[(((size_t)1 * a.ncut + m) * a.nu + u) * a.N2 + i2] = ts;
}

// Mean-field matrix of every k from the column sums (Coulomb.h:510-541):
// X_W(i1, i2)_u = -sum_m [cos_m(i1) Tc(m, u, i2) + sin_m(i1) Ts(m, u, i2)], stored
// SoA as Xw[2u + re/im][k], formed as products on the FP64 matrix cores: for each
// (u, re/im), Xw(i1, i2) = -sum_kk A[i1][kk] B[kk][i2] with A = [cos_m(i1) |
// sin_m(i1)] and B = [Tc(m, u, i2) ; Ts(m, u, i2)] (K = 2 Ncut).  Rows are i1,
// columns c = (2u + re/im) N2 + i2; one wave per 16 x 16 tile, its K loads
// issued before the products (lane l feeds A[row l%16][kk l/16] and
// B[kk l/16][col l%16] and ends with C[l/16 + 4 i][l%16]).
__global__ __launch_bounds__(64) void k_coul_w_mfma(CoulArgs a, int N1) {
    typedef double v4 __attribute__((ext_vector_type(4)));
    constexpr int KB = 16;  // K terms loaded per block (2 Ncut <= 64)
    const int lane = threadIdx.x, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int ncut = a.ncut, K = 2 * ncut, ncol = 2 * a.nu * a.N2;
    const int ra = r0 + (lane & 15), cb = c0 + (lane & 15), kq = lane >> 4;
    const bool rok = ra < N1, cok = cb < ncol;
    const int i1 = rok ? ra : 0, uc = cok ? cb / a.N2 : 0, i2 = cok ? cb % a.N2 : 0, u = uc >> 1, ri = uc & 1;
    v4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += KB) {
        double av[KB / 4], bv[KB / 4];
#pragma unroll
        for (int b = 0; b < KB / 4; b++) {
            const int kk = min(k0 + 4 * b + kq, K - 1), q = kk >= ncut, m = kk - q * ncut;
            av[b] = (q ? a.sxm : a.cxm)[m * N1 + i1];
            bv[b] = a.Tt[((((size_t)q * ncut + m) * a.nu + u) * a.N2 + i2) * 2 + ri];
        }
#pragma unroll
        for (int b = 0; b < KB / 4; b++) {
            const bool kok = k0 + 4 * b + kq < K;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64((rok && kok) ? av[b] : 0.0, (cok && kok) ? bv[b] : 0.0, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int r = r0 + (lane >> 4) + 4 * i, c = c0 + (lane & 15);
        if (r < N1 && c < ncol) a.Xw[(size_t)(c / a.N2) * a.nkp + (size_t)r * a.N2 + c % a.N2] = -acc[i];
    }
}

// Observables with the mean field (Observables_MPI.h:286-304, if_resolution.cpp:125-192):
// eigenvectors of the renormalised H (cyclic Jacobi), its shell gradient and
// H_ren D^+ for Current2.  One thread per k, nb <= 4.
template <int NB>
__global__ __launch_bounds__(256) void k_coul_obs(const double *Hr, const double *Dp, const int *nbr, const double *wb, int n_nbr,
                                                   int nk, int nkp, int dherm, double *Uexc, double *GE, double *Q, double *Ek) {
    constexpr int NP = NB * NB;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    cd A[NB][NB], V[NB][NB];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        A[i][i] = mk(Hr[i * (size_t)nkp + k], 0.0);
        V[i][i] = mk(1.0, 0.0);
#pragma unroll
        for (int j = i + 1; j < NB; j++) {
            const int p = pair_index(i, j, NB);
            A[i][j] = mk(Hr[(NB + 2 * p) * (size_t)nkp + k], Hr[(NB + 2 * p + 1) * (size_t)nkp + k]);
            A[j][i] = cconj(A[i][j]);
            V[i][j] = V[j][i] = mk(0.0, 0.0);
        }
    }
    for (int sweep = 0; sweep < 20; sweep++) {
        double off = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const double v = A[i][j].x * A[i][j].x + A[i][j].y * A[i][j].y;
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot) break;
#pragma unroll
        for (int p = 0; p < NB; p++)
#pragma unroll
            for (int q = p + 1; q < NB; q++) {
                const double mag = sqrt(A[p][q].x * A[p][q].x + A[p][q].y * A[p][q].y);
                if (mag == 0.0) continue;
                const cd e = mk(A[p][q].x / mag, A[p][q].y / mag);
                const double th = 0.5 * (A[q][q].x - A[p][p].x) / mag;
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    cd x = A[r][p], y = A[r][q];
                    A[r][p] = csub(cscale(x, cs), cscale(cmul(cconj(e), y), sn));
                    A[r][q] = cadd(cscale(cmul(e, x), sn), cscale(y, cs));
                    x = V[r][p]; y = V[r][q];
                    V[r][p] = csub(cscale(x, cs), cscale(cmul(cconj(e), y), sn));
                    V[r][q] = cadd(cscale(cmul(e, x), sn), cscale(y, cs));
                }
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    cd x = A[p][r], y = A[q][r];
                    A[p][r] = csub(cscale(x, cs), cscale(cmul(e, y), sn));
                    A[q][r] = cadd(cscale(cmul(cconj(e), x), sn), cscale(y, cs));
                }
            }
    }
    // ascending eigenvalues; Uexc rows are conj(v_a)
    int ord[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) ord[i] = i;
    for (int i = 0; i < NB; i++)
        for (int j = i + 1; j < NB; j++)
            if (A[ord[j]][ord[j]].x < A[ord[i]][ord[i]].x) { int tt = ord[i]; ord[i] = ord[j]; ord[j] = tt; }
    double *u = Uexc + (size_t)k * NB * NB * 2;
    for (int a2 = 0; a2 < NB; a2++) Ek[(size_t)k * NB + a2] = A[ord[a2]][ord[a2]].x;
    for (int a2 = 0; a2 < NB; a2++)
        for (int b = 0; b < NB; b++) {
            const cd v = V[b][ord[a2]];
            u[(a2 * NB + b) * 2] = v.x;
            u[(a2 * NB + b) * 2 + 1] = -v.y;
        }
    // gradient of the renormalised H on the shell stencil (packed Hermitian)
    for (int x = 0; x < 3; x++)
        for (int e = 0; e < NP; e++) {
            const double h0 = Hr[e * (size_t)nkp + k];
            double g = 0.0;
            for (int n = 0; n < n_nbr; n++) g += wb[n * 3 + x] * (Hr[e * (size_t)nkp + nbr[(size_t)n * nkp + k]] - h0);
            GE[((size_t)k * 3 + x) * NP + e] = g;
        }
    // Q_x = H_ren D_x^+ (D symmetric, packed)
    for (int x = 0; x < 3; x++)
        for (int i = 0; i < NB; i++)
            for (int j = 0; j < NB; j++) {
                cd acc = mk(0, 0);
                for (int kc = 0; kc < NB; kc++) {
                    cd h, d;
                    if (i == kc) h = mk(Hr[i * (size_t)nkp + k], 0.0);
                    else {
                        const int p = pair_index(min(i, kc), max(i, kc), NB);
                        h = mk(Hr[(NB + 2 * p) * (size_t)nkp + k], Hr[(NB + 2 * p + 1) * (size_t)nkp + k]);
                        if (i > kc) h = cconj(h);
                    }
                    if (j == kc) d = mk(Dp[(x * NP + j) * (size_t)nkp + k], 0.0);
                    else {
                        const int p = pair_index(min(j, kc), max(j, kc), NB);
                        d = mk(Dp[(x * NP + NB + 2 * p) * (size_t)nkp + k], Dp[(x * NP + NB + 2 * p + 1) * (size_t)nkp + k]);
                        if (dherm && j > kc) d = cconj(d);
                    }
                    acc = cadd(acc, cmul(h, cconj(d)));
                }
                double *q = Q + ((((size_t)k * 3 + x) * NB + i) * NB + j) * 2;
                q[0] = acc.x;
                q[1] = acc.y;
            }
}

// ================================================================ nb in 5..32: one workgroup per k
// Threads per k: 32 for nb <= 8 (eight k per workgroup), one wavefront for
// nb <= 16, four for nb <= 32.  The LDS GEMM of the VALU form (nb <= 8) runs on a
// GR x GC thread grid, each thread owning an RBR x RBC block of the NBT x NBT
// result.  OPA: conj(A) if 1.  OPB: 0 -> B[j][l] (result A B^T), 1 -> B[l][j],
// 2 -> conj(B[l][j]).
template <int NBT>
struct WaveGeom {
    // threads per k: 256 at 32 bands, 64 at 16, 32 at 8 (two packed elements
    // each, moved as one 16-byte pair)
    static constexpr int TPK = NBT >= 32 ? 256 : NBT >= 16 ? 64 : 32;
    // VALU product lane grid GR x GC, each thread an RBR x RBC block (interleaved)
    static constexpr int GR = TPK == 256 ? 16 : TPK == 64 ? 8 : 4;
    static constexpr int GC = TPK == 256 ? 16 : 8;
    static constexpr int RBR = NBT / GR, RBC = NBT / GC;
    static constexpr int PER = (NBT * NBT + TPK - 1) / TPK;  // packed doubles per thread
    static constexpr int LD = NBT + 1;  // LDS row length in complex elements (one pad)
    // k per workgroup: at nb <= 8 one k is little work, so eight k (256 threads)
    // share a workgroup and the dispatcher handles an eighth as many
    static constexpr int KPB = NBT == 8 ? 8 : 1;
    static constexpr int GATHER = NBT <= 8 ? 12 : NBT <= 16 ? 6 : 4;  // stencil neighbours per load batch
};

// The matrices are zero beyond nb (filled before use, and every product of
// zero-padded operands is zero there), so the inner loop has the compile-time
// trip count NBT; unrolled by 4 (a full unroll hoists every LDS read and
// costs twice the registers).
template <int NBT, int OPA, int OPB, typename T>
__device__ __forceinline__ void gemm_lds(const T *A, const T *B, int t,
                                         T (&acc)[WaveGeom<NBT>::RBR][WaveGeom<NBT>::RBC]) {
    constexpr int RBR = WaveGeom<NBT>::RBR, RBC = WaveGeom<NBT>::RBC, GR = WaveGeom<NBT>::GR, GC = WaveGeom<NBT>::GC;
    constexpr int LD = WaveGeom<NBT>::LD;
    // Lane (r0, c0) owns rows r0 + i GR and columns c0 + j GC (interleaved, not
    // blocked): the rows one LDS read instruction touches are then consecutive,
    // and rows NBT+1 complex apart start 4 banks apart, so the reads are free of
    // bank conflicts.
    const int r0 = t / GC, c0 = t % GC;
#pragma unroll
    for (int i = 0; i < RBR; i++)
#pragma unroll
        for (int j = 0; j < RBC; j++) acc[i][j] = T{0, 0};
#pragma unroll 4
    for (int l = 0; l < NBT; l++) {
        T av[RBR], bv[RBC];
#pragma unroll
        for (int i = 0; i < RBR; i++) {
            T v = A[(r0 + i * GR) * LD + l];
            av[i] = OPA ? T{v.x, -v.y} : v;
        }
#pragma unroll
        for (int j = 0; j < RBC; j++) {
            T v = (OPB == 0) ? B[(c0 + j * GC) * LD + l] : B[l * LD + c0 + j * GC];
            bv[j] = (OPB == 2) ? T{v.x, -v.y} : v;
        }
#pragma unroll
        for (int i = 0; i < RBR; i++)
#pragma unroll
            for (int j = 0; j < RBC; j++) {
                acc[i][j].x += av[i].x * bv[j].x - av[i].y * bv[j].y;
                acc[i][j].y += av[i].x * bv[j].y + av[i].y * bv[j].x;
            }
    }
}

template <int NBT, typename T>
__device__ __forceinline__ void store_block(T *C, int t, const T (&acc)[WaveGeom<NBT>::RBR][WaveGeom<NBT>::RBC]) {
    constexpr int RBR = WaveGeom<NBT>::RBR, RBC = WaveGeom<NBT>::RBC, GR = WaveGeom<NBT>::GR, GC = WaveGeom<NBT>::GC;
    constexpr int LD = WaveGeom<NBT>::LD;
    const int r0 = t / GC, c0 = t % GC;  // rows r0 + i GR, columns c0 + j GC
#pragma unroll
    for (int i = 0; i < RBR; i++)
#pragma unroll
        for (int j = 0; j < RBC; j++) C[(r0 + i * GR) * LD + c0 + j * GC] = acc[i][j];
}

// The same products on the matrix cores (NBT = 16: the k's one wave computes the
// 16x16 result; NBT = 32: wave w of the k's four computes tile (w / 2, w % 2)).
// A complex product is four real 16x16x4 MFMA chains per 4-wide slice of l:
// Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br.  Lane l feeds op(A)[row l%16][l/16]
// and op(B)[l/16][col l%16]; accumulator element i of lane l is C[row][l%16] with
// row = l/16 + 4 i for the FP64 instruction and 4 (l/16) + i for the FP32 one.
typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
template <typename R> struct MfmaAcc;
template <> struct MfmaAcc<double> { using T = v4d; };
template <> struct MfmaAcc<float> { using T = v4f; };
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

template <typename R>
__device__ __forceinline__ int mfma_row(int lane, int i) {
    return std::is_same<R, double>::value ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
}

template <int NBT>
__device__ __forceinline__ void mfma_tile(int t, int &r0, int &c0) {
    const int w = t >> 6;
    r0 = NBT == 32 ? (w >> 1) * 16 : 0;
    c0 = NBT == 32 ? (w & 1) * 16 : 0;
}

// LDS complex element access: an FP64 complex moves as one 16-byte access
// (ds_read_b128 / ds_write_b128; the matrices start 16-byte aligned and every
// element is 16 bytes) rather than two 8-byte halves, whose 32-lane passes put
// rows r and r + 1 of neighbouring K slices on one bank.
template <typename T>
__device__ __forceinline__ T lds_get(const T *p) { return *p; }
template <typename T>
__device__ __forceinline__ void lds_put(T *p, T v) { *p = v; }
template <>
__device__ __forceinline__ cd lds_get<cd>(const cd *p) {
    const v2d v = *reinterpret_cast<const v2d *>(p);
    return cd{v.x, v.y};
}
template <>
__device__ __forceinline__ void lds_put<cd>(cd *p, cd v) {
    *reinterpret_cast<v2d *>(p) = v2d{v.x, v.y};
}

template <int NBT, int OPA, int OPB, typename R, typename T = typename CplxOf<R>::T, typename V = typename MfmaAcc<R>::T>
__device__ __forceinline__ void gemm_mfma(const T *A, const T *B, int t, V &cre, V &cim) {
    constexpr int LD = WaveGeom<NBT>::LD;
    int r0, c0;
    mfma_tile<NBT>(t, r0, c0);
    const int lane = t & 63, ra = r0 + (lane & 15), cb = c0 + (lane & 15), kq = lane >> 4;
    cre = V{0, 0, 0, 0};
    cim = V{0, 0, 0, 0};
#pragma unroll
    for (int k0 = 0; k0 < NBT; k0 += 4) {
        const int l = k0 + kq;
        T a = lds_get(A + ra * LD + l);
        if (OPA) a.y = -a.y;
        T b = lds_get(B + ((OPB == 0) ? cb * LD + l : l * LD + cb));
        if (OPB == 2) b.y = -b.y;
        cre = mfma16(a.x, b.x, cre);
        cre = mfma16(-a.y, b.y, cre);
        cim = mfma16(a.x, b.y, cim);
        cim = mfma16(a.y, b.x, cim);
    }
}

template <int NBT, typename R, typename T = typename CplxOf<R>::T, typename V = typename MfmaAcc<R>::T>
__device__ __forceinline__ void store_mfma(T *C, int t, const V &cre, const V &cim) {
    constexpr int LD = WaveGeom<NBT>::LD;
    int r0, c0;
    mfma_tile<NBT>(t, r0, c0);
    const int lane = t & 63;
#pragma unroll
    for (int i = 0; i < 4; i++) lds_put(C + (r0 + mfma_row<R>(lane, i)) * LD + c0 + (lane & 15), T{cre[i], cim[i]});
}

// BF16x3 split products on the bf16 matrix cores: each FP32 operand x is split
// into hi = bf16(x) and lo = bf16(x - hi) and every real product is taken as
// lo*hi + hi*lo + hi*hi with FP32 accumulation (v_mfma_f32_16x16x32_bf16: lane l
// feeds op(A)[row l%16][8 (l/16) + q] and op(B)[8 (l/16) + q][col l%16], q < 8;
// the accumulator layout is the FP32 16x16x4 one).  The same operands, result
// and tile assignment as gemm_mfma<..., float>; K runs in slices of 32 (zero
// beyond NBT).
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void bf16_split(float x, __bf16 &hi, __bf16 &lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

__device__ __forceinline__ v4f prod3(const v8bf &xh, const v8bf &xl, const v8bf &yh, const v8bf &yl, v4f acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, yh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yl, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yh, acc, 0, 0, 0);
}

template <int NBT, int OPA, int OPB>
__device__ __forceinline__ void gemm_bf16x3(const cf *A, const cf *B, int t, v4f &cre, v4f &cim) {
    constexpr int LD = WaveGeom<NBT>::LD;
    int r0, c0;
    mfma_tile<NBT>(t, r0, c0);
    const int lane = t & 63, ra = r0 + (lane & 15), cb = c0 + (lane & 15), kq = 8 * (lane >> 4);
    cre = v4f{0.f, 0.f, 0.f, 0.f};
    cim = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < NBT; k0 += 32) {
        v8bf arh, arl, aih, ail, anh, anl, brh, brl, bih, bil;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int l = k0 + kq + q;
            cf a = cf{0.f, 0.f}, b = cf{0.f, 0.f};
            if (l < NBT) {
                a = A[ra * LD + l];
                b = (OPB == 0) ? B[cb * LD + l] : B[l * LD + cb];
            }
            if (OPA) a.y = -a.y;
            if (OPB == 2) b.y = -b.y;
            __bf16 h, lo;
            bf16_split(a.x, h, lo); arh[q] = h; arl[q] = lo;
            bf16_split(a.y, h, lo); aih[q] = h; ail[q] = lo;
            bf16_split(-a.y, h, lo); anh[q] = h; anl[q] = lo;
            bf16_split(b.x, h, lo); brh[q] = h; brl[q] = lo;
            bf16_split(b.y, h, lo); bih[q] = h; bil[q] = lo;
        }
        // Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br
        cre = prod3(arh, arl, brh, brl, cre);
        cre = prod3(anh, anl, bih, bil, cre);
        cim = prod3(arh, arl, bih, bil, cim);
        cim = prod3(aih, ail, brh, brl, cim);
    }
}

// The stage kernels' products: native MFMA of R, or BF16x3 (B3, R = float).
template <int NBT, int OPA, int OPB, typename R, bool B3>
__device__ __forceinline__ void gemm_prod(const typename CplxOf<R>::T *A, const typename CplxOf<R>::T *B, int t,
                                          typename MfmaAcc<R>::T &cre, typename MfmaAcc<R>::T &cim) {
    if constexpr (B3) {
        static_assert(std::is_same<R, float>::value, "BF16x3 products take FP32 operands");
        gemm_bf16x3<NBT, OPA, OPB>(A, B, t, cre, cim);
    } else {
        gemm_mfma<NBT, OPA, OPB, R>(A, B, t, cre, cim);
    }
}

// Row and column of accumulator element i of thread t (for the dephasing rates).
template <int NBT, typename R>
__device__ __forceinline__ void mfma_elem(int t, int i, int &row, int &col) {
    int r0, c0;
    mfma_tile<NBT>(t, r0, c0);
    row = r0 + mfma_row<R>(t & 63, i);
    col = c0 + (t & 15);
}

// Number of LDS matrices of one wave: P, M (the product X = P M^T overwrites M)
// and, with dephasing, U.
__host__ __device__ constexpr int wave_mats(int diss) { return diss ? 3 : 2; }

// One k per workgroup.  Every global operand of the stage (own state, H, D, the
// stage's accumulator and base state, the 12 stencil neighbours) is requested at
// the start, so their latencies overlap each other and the LDS work; with one
// k per workgroup the barriers only order that k's own LDS traffic.
// R is the precision of the operators and of the LDS products: double for the
// drop-in, float for the mixed-precision mode (state, stencil and increments stay
// FP64 in both).
// LDS of one k of the wave kernel, in doubles.
template <int NBT, typename R>
__host__ __device__ constexpr int wave_lds_doubles(int diss) {
    return wave_mats(diss) * NBT * WaveGeom<NBT>::LD * (int)sizeof(typename CplxOf<R>::T) / 8;
}

// k: the k this group of TPK threads serves (a valid index); lane: thread within
// the group; smem: the group's LDS; store: false for a group past the end of the
// grid, which computes its clamped k again but writes nothing (some stage modes
// update their accumulator in place).
// ROLE 0: k alone.  A stencil pair (k, kb = k + 1, see ConstArgs::pair) is served
// by two calls: ROLE 1 steps k and forms kb's stencil sum in gb, ROLE 2 steps kb
// with that sum.  The sum of k runs in its own slot order (as alone); the one of
// kb adds the neighbours it shares with k in k's slot order, then k itself, then
// its other neighbours in its own slot order.
// S: storage type of the RK stage inputs/outputs (YA, YB) and of the multistep
// histories: double, or float in the mixed form with FP32 stage storage, where
// the state at t_n (P0), the RK accumulator and every update stay FP64.
template <int NBT, int MODE, typename R, typename S = double, bool B3 = false, int ROLE = 0>
__device__ __forceinline__ void stage_wave(const ConstArgs &c, const StageArgs &a, const int k, const int kb, const int lane,
                                           double *smem, const bool store, double (&gb)[WaveGeom<NBT>::PER]) {
    using CT = typename CplxOf<R>::T;
    // stage input in S for the RK stages 2-4 (they read YA / YB); the FP64 state otherwise
    constexpr bool kInS = std::is_same<S, float>::value && (MODE == RK2 || MODE == RK3 || MODE == RK4S);
    using YT = typename std::conditional<kInS, S, double>::type;
    const YT *const Yin = reinterpret_cast<const YT *>(a.Y);
    constexpr int TPK = WaveGeom<NBT>::TPK, RBR = WaveGeom<NBT>::RBR, RBC = WaveGeom<NBT>::RBC, PER = WaveGeom<NBT>::PER;
    constexpr int LD = WaveGeom<NBT>::LD, MAT = NBT * LD;
    const int nb = c.nb, NP = nb * nb, nmat = wave_mats(c.diss);
    constexpr bool kMfma = NBT >= 16;  // 16x16 matrix-core tiles; NBT = 8 stays on the VALU
    CT *A = reinterpret_cast<CT *>(smem), *B = A + MAT, *C = c.diss ? B + MAT : B;
    const R *Hop = reinterpret_cast<const R *>(sizeof(R) == 8 ? (const void *)c.H : (const void *)c.Hf);
    const R *Dop = reinterpret_cast<const R *>(sizeof(R) == 8 ? (const void *)c.D : (const void *)c.Df);
    const R *Uop = reinterpret_cast<const R *>(sizeof(R) == 8 ? (const void *)c.U : (const void *)c.Uf);
    const double *tab = stage_row(a);
    const R Ep[3] = {(R)tab[0], (R)tab[1], (R)tab[2]}, Eq[3] = {(R)tab[3], (R)tab[4], (R)tab[5]};
    const int nn = c.n_nbr;
    const bool grad = tab[6 + nn] != 0.0;

    // Lanes past the last packed element (nb < NBT) repeat element NP-1: every
    // load stays in bounds without a branch, and their LDS and global stores
    // write the same value to the same address as the lane that owns it.
    // With an even number of doubles per thread, a lane owns adjacent pairs of
    // elements (2q, 2q+1), q = lane + TPK m', read and written 16 bytes at a time
    // when nb is even (NP even keeps every pair inside one k and aligned); the
    // pair is a real diagonal pair or the (re, im) of one upper element.
    constexpr bool kPairs = PER % 2 == 0;
    // A stencil pair's calls load the RK2/RK3 accumulator and state at t_n after
    // the products, not with the other inputs: the registers then hold kb's
    // stencil sum instead through the gather and the products.
    constexpr bool kLatePre = ROLE != 0 && kPairs;
    bool pv2[PER > 1 ? PER / 2 : 1];
    int eo[PER], ei[PER], ej[PER], ec[PER];
    double y[PER], pv[PER], g[PER], pre0[PER], pre1[PER];
    R hv[PER], dv[PER][3];
    const size_t base = (size_t)k * NP;
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const int e = kPairs ? min(2 * (lane + TPK * (m >> 1)) + (m & 1), NP - 1) : min(lane + TPK * m, NP - 1);
        eo[m] = e;
        const int code = c.eidx[e];
        ei[m] = code & 255;
        ej[m] = (code >> 8) & 255;
        ec[m] = code >> 16;
        g[m] = pre0[m] = pre1[m] = 0.0;
        if constexpr (ROLE == 1) gb[m] = 0.0;
    }
#pragma unroll
    for (int h = 0; h < (PER > 1 ? PER / 2 : 1); h++) pv2[h] = kPairs && (NP % 2 == 0) && 2 * (lane + TPK * h) + 1 < NP;
    // Neighbour data of the stage input, GB neighbours per call, their loads all
    // issued before any is used (k's or kb's stencil slots, by `slot`).
    constexpr int GB = WaveGeom<NBT>::GATHER;
    auto gather = [&](const int kk, auto slot, const int cnt, auto &v) {
        constexpr int NV = std::extent<std::remove_reference_t<decltype(v)>>::value;
        size_t kn[NV];
#pragma unroll
        for (int j = 0; j < NV; j++) kn[j] = (size_t)c.nbr[(size_t)slot(min(j, cnt - 1)) * c.nkp + kk] * NP;
#pragma unroll
        for (int j = 0; j < NV; j++) {
            if constexpr (kPairs) {
#pragma unroll
                for (int h = 0; h < PER / 2; h++) {
                    YT v0, v1;
                    ld2<false>(Yin, kn[j] + eo[2 * h], kn[j] + eo[2 * h + 1], pv2[h], v0, v1);
                    v[j][2 * h] = (double)v0;
                    v[j][2 * h + 1] = (double)v1;
                }
            } else {
#pragma unroll
                for (int m = 0; m < PER; m++) v[j][m] = (double)Yin[kn[j] + eo[m]];
            }
        }
    };
    if constexpr (kPairs) {
#pragma unroll
        for (int h = 0; h < PER / 2; h++) {
            const int m = 2 * h;
            const bool v2 = pv2[h];
            const size_t o0 = base + eo[m], o1 = base + eo[m + 1];
            YT y0, y1;
            ld2<false>(Yin, o0, o1, v2, y0, y1);
            y[m] = (double)y0;
            y[m + 1] = (double)y1;
            ld2<true>(Hop, o0, o1, v2, hv[m], hv[m + 1]);
#pragma unroll
            for (int x = 0; x < 3; x++) {
                if (a.dmask >> x & 1) ld2<true>(Dop + (size_t)x * c.nk * NP, o0, o1, v2, dv[m][x], dv[m + 1][x]);
                else dv[m][x] = dv[m + 1][x] = (R)0;
            }
            if constexpr ((MODE == RK2 || MODE == RK3) && !kLatePre) {
                ld2<true>(a.ACC, o0, o1, v2, pre0[m], pre0[m + 1]);
                ld2<true>(a.P0, o0, o1, v2, pre1[m], pre1[m + 1]);
            } else if constexpr (MODE == RK4S) {
                ld2<true>(a.ACC, o0, o1, v2, pre0[m], pre0[m + 1]);
            }
        }
    } else {
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int e = eo[m];
            y[m] = (double)Yin[base + e];
            hv[m] = ldnt(Hop + base + e);
#pragma unroll
            for (int x = 0; x < 3; x++) dv[m][x] = (a.dmask >> x & 1) ? ldnt(Dop + (size_t)x * c.nk * NP + base + e) : (R)0;
            if constexpr (MODE == RK2 || MODE == RK3) {
                pre0[m] = ldnt(a.ACC + base + e);
                pre1[m] = ldnt(a.P0 + base + e);
            } else if constexpr (MODE == RK4S) {
                pre0[m] = ldnt(a.ACC + base + e);
            }
        }
    }
    // Field drift E.grad_k rho on the shell stencil (RungeKutta.h:258-276).  The
    // neighbours are fetched GB at a time: their index and data loads are all in
    // flight together, and the sum keeps the stencil order.  First of a pair
    // (ROLE 1): every neighbour of k that is also one of kb's adds to kb's sum gb
    // as well, and kb's remaining neighbours are fetched after k's.  Second of a
    // pair (ROLE 2): the sum is gb.
    if constexpr (ROLE == 2) {
#pragma unroll
        for (int m = 0; m < PER; m++) g[m] = gb[m];
    } else if (grad) {
        for (int n0 = 0; n0 < nn; n0 += GB) {
            double v[GB][PER];
            gather(k, [&](int j) { return n0 + j; }, nn - n0, v);
#pragma unroll
            for (int j = 0; j < GB; j++)
                if (n0 + j < nn) {
                    const double cn = tab[6 + n0 + j];
#pragma unroll
                    for (int m = 0; m < PER; m++) g[m] += cn * v[j][m];
                    if constexpr (ROLE == 1) {
                        const int pj = c.pat[n0 + j];
                        if (pj >= 0) {
                            const double cb = tab[6 + pj];
#pragma unroll
                            for (int m = 0; m < PER; m++) gb[m] += cb * v[j][m];
                        }
                    }
                }
        }
        if constexpr (ROLE == 1) {
            if (c.pself >= 0) {
                const double cb = tab[6 + c.pself];
#pragma unroll
                for (int m = 0; m < PER; m++) gb[m] += cb * y[m];
            }
            // kb's own rows (5 on the hexagonal two-shell stencil) in one batch
            constexpr int GBN = GB < 5 ? 5 : GB;
            for (int t0 = 0; t0 < c.npnew; t0 += GBN) {
                double v[GBN][PER];
                gather(kb, [&](int j) { return c.pat[kMaxNbr + t0 + j]; }, c.npnew - t0, v);
#pragma unroll
                for (int j = 0; j < GBN; j++)
                    if (t0 + j < c.npnew) {
                        const double cb = tab[6 + c.pat[kMaxNbr + t0 + j]];
#pragma unroll
                        for (int m = 0; m < PER; m++) gb[m] += cb * v[j][m];
                    }
            }
        }
    }
    if (nb < NBT)
        for (int t = lane; t < nmat * MAT; t += TPK) A[t] = CT{0, 0};
    __syncthreads();
    // A <- P (full Hermitian), B <- M = H + E.D.  Element (i, j, comp) writes its
    // upper value and its mirror; a diagonal element writes (y, 0).  A lane's pair
    // (re, im) of one upper element writes both as one complex, upper and mirror.
    {
        R *Ad = reinterpret_cast<R *>(A), *Bd = reinterpret_cast<R *>(B);
#pragma unroll
        for (int m = 0; m < PER; m++) {
            if (kPairs && (m & 1) == 0 && pv2[m >> 1] && eo[m] >= nb) {
                const int i = ei[m], j = ej[m];
                const bool w = (i >= c.nch && j < c.nch) || (i < c.nch && j >= c.nch);
                const R E[3] = {w ? Eq[0] : Ep[0], w ? Eq[1] : Ep[1], w ? Eq[2] : Ep[2]};
                R dr = hv[m], di = hv[m + 1], ddi = 0;
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    dr += E[x] * dv[m][x];
                    di += E[x] * dv[m + 1][x];
                    ddi += E[x] * dv[m + 1][x];
                }
                const R yr = (R)y[m], yi = (R)y[m + 1];
                lds_put(A + i * LD + j, CT{yr, yi});
                lds_put(A + j * LD + i, CT{yr, -yi});
                lds_put(B + i * LD + j, CT{dr, di});
                lds_put(B + j * LD + i, CT{dr, c.dherm ? -di : -hv[m + 1] + ddi});
                continue;
            }
            if (kPairs && (m & 1) == 1 && pv2[m >> 1] && eo[m - 1] >= nb) continue;  // written with m - 1
            const int i = ei[m], j = ej[m], comp = ec[m];
            const bool w = (i >= c.nch && j < c.nch) || (i < c.nch && j >= c.nch);
            const R E[3] = {w ? Eq[0] : Ep[0], w ? Eq[1] : Ep[1], w ? Eq[2] : Ep[2]};  // selects, no pointer
            R dsum = hv[m], dd = 0;
#pragma unroll
            for (int x = 0; x < 3; x++) {
                dsum += E[x] * dv[m][x];
                dd += E[x] * dv[m][x];
            }
            const bool diag = i == j;
            const R yr = (R)y[m];
            const int up = 2 * (i * LD + j) + comp;
            const int lo = diag ? up + 1 : 2 * (j * LD + i) + comp;
            const R mlo = comp ? (c.dherm ? -dsum : -hv[m] + dd) : dsum;
            Ad[up] = yr;
            Ad[lo] = diag ? (R)0 : (comp ? -yr : yr);
            Bd[up] = dsum;
            Bd[lo] = diag ? (R)0 : mlo;
        }
    }
    __syncthreads();
    // X = P M^T; -i[M,rho] (transposed storage) = -i (X - X^+)
    if constexpr (kMfma) {
        typename MfmaAcc<R>::T cre, cim;
        gemm_prod<NBT, 0, 0, R, B3>(A, B, lane, cre, cim);
        if (C == B) __syncthreads();  // every lane has finished reading M
        store_mfma<NBT, R>(C, lane, cre, cim);
    } else {
        CT acc[RBR][RBC];
        gemm_lds<NBT, 0, 0>(A, B, lane, acc);
        if (C == B) __syncthreads();  // every lane has finished reading M
        store_block<NBT>(C, lane, acc);
    }
    __syncthreads();
    if constexpr (kLatePre && (MODE == RK2 || MODE == RK3)) {
#pragma unroll
        for (int h = 0; h < PER / 2; h++) {
            const int m = 2 * h;
            const size_t o0 = base + eo[m], o1 = base + eo[m + 1];
            ld2<true>(a.ACC, o0, o1, pv2[h], pre0[m], pre0[m + 1]);
            ld2<true>(a.P0, o0, o1, pv2[h], pre1[m], pre1[m + 1]);
        }
    }
    // pv = -i (X_ij - conj(X_ji)) plus the RWA term on core -> non-core pairs; on
    // the diagonal this is 2 Im X_ii.
#pragma unroll
    for (int m = 0; m < PER; m++) {
        if (kPairs && (m & 1) == 1 && pv2[m >> 1] && eo[m - 1] >= nb) continue;  // formed with m - 1
        const int i = ei[m], j = ej[m];
        const CT xij = lds_get(C + i * LD + j), xji = lds_get(C + j * LD + i), pij = lds_get(A + i * LD + j);
        const double bre = (double)xij.x - (double)xji.x, bim = (double)xij.y + (double)xji.y;
        const double wr = (i < c.nch && j >= c.nch) ? c.wx2 : 0.0;
        const double vr = bim + -wr * (double)pij.y, vi = -bre + wr * (double)pij.x;
        if (kPairs && (m & 1) == 0 && pv2[m >> 1] && eo[m] >= nb) {
            pv[m] = vr;
            pv[m + 1] = vi;
        } else {
            pv[m] = ec[m] ? vi : vr;
        }
    }
    if (c.diss) {
        __syncthreads();
        // C <- U
        const R *u = Uop + (size_t)k * nb * nb * 2;
        for (int t = lane; t < nb * nb; t += TPK) C[(t / nb) * LD + (t % nb)] = CT{ldnt(u + 2 * t), ldnt(u + 2 * t + 1)};
        __syncthreads();
        if constexpr (kMfma) {
            typename MfmaAcc<R>::T cre, cim;
            gemm_prod<NBT, 0, 0, R, B3>(A, C, lane, cre, cim);     // T1 = P U^T
            __syncthreads();
            store_mfma<NBT, R>(B, lane, cre, cim);
            __syncthreads();
            gemm_prod<NBT, 1, 1, R, B3>(C, B, lane, cre, cim);     // G = conj(U) T1
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int ia, ib;
                mfma_elem<NBT, R>(lane, i, ia, ib);
                const R t = (ia < nb && ib < nb) ? (R)c.T[ia * nb + ib] : (R)0;
                cre[i] *= t;
                cim[i] *= t;
            }
            __syncthreads();
            store_mfma<NBT, R>(A, lane, cre, cim);               // A <- K = T o G
            __syncthreads();
            gemm_prod<NBT, 0, 2, R, B3>(A, C, lane, cre, cim);     // T2 = K conj(U)
            __syncthreads();
            store_mfma<NBT, R>(B, lane, cre, cim);
            __syncthreads();
            gemm_prod<NBT, 1, 1, R, B3>(C, B, lane, cre, cim);     // W = conj(U) T2
            __syncthreads();
            store_mfma<NBT, R>(A, lane, cre, cim);
            __syncthreads();
        } else {
            CT acc[RBR][RBC];
            gemm_lds<NBT, 0, 0>(A, C, lane, acc);     // T1 = P U^T
            __syncthreads();
            store_block<NBT>(B, lane, acc);
            __syncthreads();
            gemm_lds<NBT, 1, 1>(C, B, lane, acc);     // G = conj(U) T1
            {
                constexpr int GR = WaveGeom<NBT>::GR, GC = WaveGeom<NBT>::GC;
                const int r0 = lane / GC, c0 = lane % GC;
    #pragma unroll
                for (int i = 0; i < RBR; i++)
    #pragma unroll
                    for (int j = 0; j < RBC; j++) {
                        const int ia = r0 + i * GR, ib = c0 + j * GC;
                        const R t = (ia < nb && ib < nb) ? (R)c.T[ia * nb + ib] : (R)0;
                        acc[i][j] = CT{acc[i][j].x * t, acc[i][j].y * t};
                    }
            }
            __syncthreads();
            store_block<NBT>(A, lane, acc);               // A <- K = T o G
            __syncthreads();
            gemm_lds<NBT, 0, 2>(A, C, lane, acc);     // T2 = K conj(U)
            __syncthreads();
            store_block<NBT>(B, lane, acc);
            __syncthreads();
            gemm_lds<NBT, 1, 1>(C, B, lane, acc);     // W = conj(U) T2
            __syncthreads();
            store_block<NBT>(A, lane, acc);
            __syncthreads();
        }
#pragma unroll
        for (int m = 0; m < PER; m++) {
            if (kPairs && (m & 1) == 1 && pv2[m >> 1] && eo[m - 1] >= nb) continue;  // formed with m - 1
            const int i = ei[m], j = ej[m];
            const CT w = A[i * LD + j];
            CT hole = CT{0, 0};
            for (int v = 0; v < c.ncc; v++) {
                const CT uj = C[v * LD + j], ui = C[v * LD + i];
                const R tv = (R)c.T[v * nb + v];
                // conj(U[v][j]) U[v][i] T[v][v]
                hole.x += (uj.x * ui.x + uj.y * ui.y) * tv;
                hole.y += (uj.x * ui.y - uj.y * ui.x) * tv;
            }
            if (kPairs && (m & 1) == 0 && pv2[m >> 1] && eo[m] >= nb) {
                pv[m] = pv[m] - (double)w.x + (double)hole.x;
                pv[m + 1] = pv[m + 1] - (double)w.y + (double)hole.y;
            } else {
                pv[m] = ec[m] ? pv[m] - (double)w.y + (double)hole.y : pv[m] - (double)w.x + (double)hole.x;
            }
        }
    }
    if (!store) return;  // after the last barrier (a group past the grid serves no pair)
    if constexpr (kPairs && MODE != RHSO) {
        S *const outs = reinterpret_cast<S *>(a.OUT);  // RK stages 1-3 write the next stage input
#pragma unroll
        for (int h = 0; h < PER / 2; h++) {
            const int m = 2 * h;
            const double k0 = pv[m] + g[m], k1 = pv[m + 1] + g[m + 1];
            const size_t o0 = base + eo[m], o1 = base + eo[m + 1];
            const bool v2 = pv2[h];
            if constexpr (MODE == ABM) {
                stage_update2_abm<S>(a, o0, o1, v2, (size_t)c.nk * NP, k0, k1, y[m], y[m + 1]);
            } else if constexpr (MODE == TAY) {
                stage_update2_tay<S>(a, o0, o1, v2, (size_t)c.nk * NP, k0, k1, y[m], y[m + 1]);
            } else if constexpr (MODE == RK1) {
                st2(a.ACC, o0, o1, v2, y[m] + k0 * a.ca, y[m + 1] + k1 * a.ca);
                st2(outs, o0, o1, v2, (S)(y[m] + k0 * a.cb), (S)(y[m + 1] + k1 * a.cb));
            } else if constexpr (MODE == RK2 || MODE == RK3) {
                st2(a.ACC, o0, o1, v2, pre0[m] + k0 * a.ca, pre0[m + 1] + k1 * a.ca);
                st2(outs, o0, o1, v2, (S)(pre1[m] + k0 * a.cb), (S)(pre1[m + 1] + k1 * a.cb));
            } else {
                st2(a.OUT, o0, o1, v2, pre0[m] + k0 * a.ca, pre0[m + 1] + k1 * a.ca);
            }
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const double kv = pv[m] + g[m];
        const size_t o = base + eo[m];
        S *const outs = reinterpret_cast<S *>(a.OUT);  // RK stages 1-3 write the next stage input
        if constexpr (MODE == RK1) {
            stnt(a.ACC + o, y[m] + kv * a.ca);
            stnt(outs + o, (S)(y[m] + kv * a.cb));
        } else if constexpr (MODE == RK2 || MODE == RK3) {
            stnt(a.ACC + o, pre0[m] + kv * a.ca);
            stnt(outs + o, (S)(pre1[m] + kv * a.cb));
        } else if constexpr (MODE == RK4S) {
            stnt(a.OUT + o, pre0[m] + kv * a.ca);
        } else {
            stage_update<MODE, S>(a, o, (size_t)c.nk * NP, kv, y[m]);
        }
    }
}

// One k per workgroup, or with PAIR (ConstArgs::pair, one k per workgroup only)
// the stencil pair (2q, 2q + 1).
template <int NBT, int MODE, typename R, typename S = double, bool B3 = false, bool PAIR = false>
__global__ __launch_bounds__(WaveGeom<NBT>::TPK * WaveGeom<NBT>::KPB) void k_stage_wave(ConstArgs c, StageArgs a) {
    extern __shared__ double smem_d[];
    constexpr int TPK = WaveGeom<NBT>::TPK;
    double gb[WaveGeom<NBT>::PER];
    if constexpr (PAIR) {
        static_assert(WaveGeom<NBT>::KPB == 1, "stencil pairs: one k per workgroup");
        // xcd_block and strip_k: bijections of the nk / 2 pairs (the pair grid has
        // half as many columns: strips of sw / 2 pairs)
        const int k = 2 * strip_k(xcd_block(), c.nk >> 1, c.sw >> 1, c.scol >> 1);
        stage_wave<NBT, MODE, R, S, B3, 1>(c, a, k, k + 1, threadIdx.x, smem_d, true, gb);
        __syncthreads();  // every lane is done with k's matrices
        // the second call forms its lane-derived indices and addresses anew
        // rather than keeping the first call's live across both
        int lane2 = threadIdx.x;
        asm volatile("" : "+v"(lane2));
        stage_wave<NBT, MODE, R, S, B3, 2>(c, a, k + 1, -1, lane2, smem_d, true, gb);
    } else if constexpr (WaveGeom<NBT>::KPB == 1) {
        // xcd_block and strip_k: bijections of [0, nk)
        stage_wave<NBT, MODE, R, S, B3>(c, a, strip_k(xcd_block(), c.nk, c.sw, c.scol), -1, threadIdx.x, smem_d, true, gb);
    } else {
        const int sub = threadIdx.x / TPK;
        const int kk = xcd_block() * WaveGeom<NBT>::KPB + sub;
        stage_wave<NBT, MODE, R, S, B3>(c, a, strip_k(min(kk, c.nk - 1), c.nk, c.sw, c.scol), -1, threadIdx.x % TPK,
                                 smem_d + sub * wave_lds_doubles<NBT, R>(c.diss), kk < c.nk, gb);
    }
}

// ---------------------------------------------------------------- batched instances
// One launch per RK4 stage and kernel class advances several independent
// problems: workgroup b serves (instance, block of that instance) from a device
// table, and every instance's arguments sit in a device array.  The stage body
// is the single-problem kernel's, so a batched step is bit-identical to
// stepping each problem on its own.
// Threads per workgroup of the batched small kernel: an instance of a few
// thousand k then spreads over more CUs (the k past nk, up to the padded nkp,
// are computed and never read).
constexpr int kSmallBatchThreads = 64;

struct BatchEntry {
    ConstArgs c;
    StageArgs a[4];  // RK4 stages; tab is the instance's row of the call's first step
};

template <int MODE>
__host__ __device__ constexpr int rk4_stage_of() { return MODE == RK1 ? 0 : MODE == RK2 ? 1 : MODE == RK3 ? 2 : 3; }

template <int NB, int MODE, bool DISS>
__global__ __launch_bounds__(256) void k_stage_small_batch(const BatchEntry *E, const int2 *blk, const int *dmask, size_t tab_step) {
    const int2 w = blk[xcd_block()];
    const BatchEntry &be = E[w.x];
    StageArgs a = be.a[rk4_stage_of<MODE>()];
    a.tab += tab_step;
    a.dmask = dmask[w.x];
    stage_small<NB, MODE, DISS>(be.c, a, w.y * blockDim.x + threadIdx.x);
}

template <int NBT, int MODE, typename R, bool PAIR = false>
__global__ __launch_bounds__(WaveGeom<NBT>::TPK * WaveGeom<NBT>::KPB) void k_stage_wave_batch(const BatchEntry *E, const int2 *blk,
                                                                                         const int *dmask, size_t tab_step) {
    extern __shared__ double smem_d[];
    constexpr int TPK = WaveGeom<NBT>::TPK;
    const int2 w = blk[xcd_block()];
    const BatchEntry &be = E[w.x];
    StageArgs a = be.a[rk4_stage_of<MODE>()];
    a.tab += tab_step;
    a.dmask = dmask[w.x];
    double gb[WaveGeom<NBT>::PER];
    if constexpr (PAIR) {  // every instance of the class is paired: block w.y serves pair w.y
        static_assert(WaveGeom<NBT>::KPB == 1, "stencil pairs: one k per workgroup");
        stage_wave<NBT, MODE, R, double, false, 1>(be.c, a, 2 * w.y, 2 * w.y + 1, threadIdx.x, smem_d, true, gb);
        __syncthreads();  // every lane is done with the first k's matrices
        stage_wave<NBT, MODE, R, double, false, 2>(be.c, a, 2 * w.y + 1, -1, threadIdx.x, smem_d, true, gb);
    } else if constexpr (WaveGeom<NBT>::KPB == 1) {
        stage_wave<NBT, MODE, R>(be.c, a, w.y, -1, threadIdx.x, smem_d, true, gb);
    } else {
        const int sub = threadIdx.x / TPK, kk = w.y * WaveGeom<NBT>::KPB + sub;
        stage_wave<NBT, MODE, R>(be.c, a, min(kk, be.c.nk - 1), -1, threadIdx.x % TPK,
                                 smem_d + sub * wave_lds_doubles<NBT, R>(be.c.diss), kk < be.c.nk, gb);
    }
}

// ================================================================ observables
// Values per partial: pop[nb], j1 re x3, j2 (re,im) x3, tabs (re,im) x3 -> nb + 15,
// plus min and max of the population diagnostics.
struct ObsArgs {
    const double *P;     // state
    const double *U;     // full complex AoS
    const double *GE;    // [k][3][NP] packed Hermitian gradient of H (may be null)
    const double *Q;     // [k][3][nb][nb] complex: H D^+ (Current2)
    const double *DT;    // [k][3][nch][nb-nch] complex: D on the core<->non-core block
    double *part;        // [nblocks][nv + 2] (sums with Kahan compensation collapsed)
    int nk, nb, nch, ncc;
    size_t sk, se;       // state strides: element e of k at k*sk + e*se
};

struct Kahan {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double v) {
        double yv = v - c;
        double t = s + yv;
        c = (t - s) - yv;
        s = t;
    }
};

__device__ __forceinline__ cd load_p(const ObsArgs &o, int k, int i, int j) {
    const int nb = o.nb;
    if (i == j) return mk(o.P[k * o.sk + i * o.se], 0.0);
    const int a = min(i, j), b = max(i, j);
    const int p = pair_index(a, b, nb);
    cd v = mk(o.P[k * o.sk + (nb + 2 * p) * o.se], o.P[k * o.sk + (nb + 2 * p + 1) * o.se]);
    return (i < j) ? v : cconj(v);
}

// One workgroup of T threads walks its k points and stages each k's full P (from
// the packed state) and U in its LDS, so the population quadratic form and the
// current / absorption sums read their matrix operands from LDS.  T = 256 for
// large band counts: the k's loads and sums are spread over four waves, and the
// population form of band c over S = T / 32 lanes (terms ii = s, s + S, ...,
// summed by shuffles), so no lane walks a long chain of dependent loads.
constexpr int kObsMaxBlocks = 8192;

template <int T>
__global__ __launch_bounds__(T) void k_observables(ObsArgs o) {
    extern __shared__ double obs_lds[];
    constexpr int S = T / 32, NW = T / 64;
    const int lane = threadIdx.x, wave = lane >> 6;
    const int nb = o.nb, NP = nb * nb, NN = nb * nb, nch = o.nch;
    // U rows padded to nb + 1: lanes of consecutive bands start 4 banks apart
    const int LU = nb + 1;
    cd *Pl = reinterpret_cast<cd *>(obs_lds), *Ul = Pl + NN;
    double *red = reinterpret_cast<double *>(Ul + nb * LU);  // [NW][17] cross-wave partials, then [32] populations
    Kahan pop, j1[3], j2[6], ta[6];
    double pmin = 1.0, pmax = 0.0;
    const int pc = lane / S, ps = lane % S;  // band and term slice of the population form
    for (int k = blockIdx.x; k < o.nk; k += gridDim.x) {
        __syncthreads();  // the previous k's LDS reads are done
        for (int t = lane; t < NN; t += T) {
            Pl[t] = load_p(o, k, t / nb, t % nb);
            double ur, ui;
            ld2<false>(o.U, ((size_t)k * NN + t) * 2, ((size_t)k * NN + t) * 2 + 1, true, ur, ui);
            Ul[(t / nb) * LU + t % nb] = mk(ur, ui);
        }
        __syncthreads();
        // population of band c = Re (conj(U) P_sym U^T)_cc (InitParameters.h:264-335)
        {
            double sre = 0.0;
            if (pc < nb) {
                const cd *uc = Ul + pc * LU;
                cd sacc = mk(0.0, 0.0);
                for (int ii = ps; ii < nb; ii += S) {
                    cd t = mk(0.0, 0.0);
#pragma unroll 4
                    for (int jj = 0; jj < nb; jj++) t = cadd(t, cmul(Pl[ii * nb + jj], uc[jj]));
                    sacc = cadd(sacc, cmul(cconj(uc[ii]), t));
                }
                sre = sacc.x;
            }
#pragma unroll
            for (int off = S / 2; off > 0; off >>= 1) sre += __shfl_xor(sre, off, 64);
            if (pc < nb && ps == 0) {
                pop.add(sre);
                const double hv = (pc < o.ncc) ? 1. - sre : sre;
                pmax = fmax(pmax, fabs(hv));
                pmin = fmin(pmin, hv);
            }
        }
        // Current1: sum_ij P_ij gradH_ij (real: P and gradH are Hermitian)
        if (o.GE) {
            for (int e = lane; e < NP; e += T) {
                const double pe = o.P[k * o.sk + e * o.se];
                const double wgt = (e < nb) ? 1.0 : 2.0;   // a pair carries (i,j) and (j,i)
                // Re(P_ij G_ij): re*re - im*im over the (re, im) doubles of the pair
                const double sgn = (e >= nb && ((e - nb) & 1)) ? -1.0 : 1.0;
                for (int x = 0; x < 3; x++) j1[x].add(wgt * sgn * pe * o.GE[((size_t)k * 3 + x) * NP + e]);
            }
        }
        // Current2: sum_ij P_ij (H D^+)_ij
        for (int t = lane; t < NN; t += T) {
            const cd pij = Pl[t];
            for (int x = 0; x < 3; x++) {
                double qr, qi;
                ld2<false>(o.Q, (((size_t)k * 3 + x) * NN + t) * 2, (((size_t)k * 3 + x) * NN + t) * 2 + 1, true, qr, qi);
                cd v = cmul(pij, mk(qr, qi));
                j2[2 * x].add(v.x);
                j2[2 * x + 1].add(v.y);
            }
        }
        // Transient absorption: -2 sum_{i<nch<=j} D_ij P_ij (Observables_MPI.h:559-569)
        const int nblk = nch * (nb - nch);
        for (int t = lane; t < nblk; t += T) {
            const int i = t / (nb - nch), j = nch + t % (nb - nch);
            const cd pij = Pl[i * nb + j];
            for (int x = 0; x < 3; x++) {
                const double *d = o.DT + (((size_t)k * 3 + x) * nblk + t) * 2;
                cd v = cmul(mk(d[0], d[1]), pij);
                ta[2 * x].add(-2. * v.x);
                ta[2 * x + 1].add(-2. * v.y);
            }
        }
    }
    // Lane sums by shuffle, then across the waves in wave order, in a fixed order.
    __syncthreads();
    double *pw = red + wave * 17;
    for (int x = 0; x < 3; x++) {
        const double v = wave_sum(j1[x].s);
        if ((lane & 63) == 0) pw[x] = v;
    }
    for (int x = 0; x < 6; x++) {
        const double v = wave_sum(j2[x].s), w = wave_sum(ta[x].s);
        if ((lane & 63) == 0) { pw[3 + x] = v; pw[9 + x] = w; }
    }
    const double mn = wave_min(pmin), mx = wave_max(pmax);
    if ((lane & 63) == 0) { pw[15] = mn; pw[16] = mx; }
    double *pops = red + NW * 17;
    if (pc < nb && ps == 0) pops[pc] = pop.s;
    __syncthreads();
    const int nv = nb + 15;
    double *out = o.part + (size_t)blockIdx.x * (nv + 2);
    if (lane < nb) out[lane] = pops[lane];
    if (lane < 17) {
        double v = red[lane];
        for (int w = 1; w < NW; w++) {
            const double u = red[w * 17 + lane];
            v = lane == 15 ? fmin(v, u) : lane == 16 ? fmax(v, u) : v + u;
        }
        out[nb + lane] = v;  // nb+0..2 J1, +3..8 J2, +9..14 TAbs, nv and nv+1 min and max
    }
}

// Final reduction of the per-block partials: one workgroup per value.  Each
// thread sums a strided subset with Kahan compensation, and the 256 (sum,
// compensation) pairs are combined pairwise with an error-free two-sum, so the
// result keeps the compensated accuracy; the last two values are the
// population diagnostics' min and max.
__global__ __launch_bounds__(256) void k_obs_final(const double *part, int nparts, int nv, double *out) {
    __shared__ double shi[256], slo[256];
    const int c = blockIdx.x, t = threadIdx.x;
    const size_t stride = (size_t)nv + 2;
    if (c < nv) {
        double s = 0.0, comp = 0.0;
        for (int b = t; b < nparts; b += 256) {
            const double yv = part[(size_t)b * stride + c] - comp;
            const double q = s + yv;
            comp = (q - s) - yv;
            s = q;
        }
        shi[t] = s;
        slo[t] = -comp;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (t < h) {
                const double a = shi[t], b = shi[t + h];
                const double sum = a + b, bp = sum - a;
                const double err = (a - (sum - bp)) + (b - bp);
                shi[t] = sum;
                slo[t] = slo[t] + slo[t + h] + err;
            }
            __syncthreads();
        }
        if (t == 0) out[c] = shi[0] + slo[0];
    } else {
        const bool mn = c == nv;
        double v = mn ? 1.0 : 0.0;
        for (int b = t; b < nparts; b += 256) {
            const double x = part[(size_t)b * stride + c];
            v = mn ? fmin(v, x) : fmax(v, x);
        }
        shi[t] = v;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (t < h) shi[t] = mn ? fmin(shi[t], shi[t + h]) : fmax(shi[t], shi[t + h]);
            __syncthreads();
        }
        if (t == 0) out[c] = shi[0];
    }
}

// ================================================================ setup helpers
// Q_x = H D_x^+ per k (Current2's sum_kc H_ik conj(D_jk), Observables_MPI.h:726-775)
// from the packed operators: full complex [k][3][nb][nb].  Element e of k sits at
// k*sk + e*se, component x of D at x*dxs.
// FP32 copy of an operator table (mixed-precision mode), grid-stride.
__global__ __launch_bounds__(256) void k_to_float(const double *src, float *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = (float)src[i];
}

// Engine creation packs the host's reference-layout tables on the device: a
// chunk of kc k is copied as is, then one thread per (k, i, j) writes the upper
// triangle of component x (complex stride cs per element) to dst[k sk + e se]
// (e = i on the diagonal, real part only; nb + 2 pair(i, j) and + 1 above it).
__global__ void k_pack(const double *__restrict__ src, long long n, int nb, int cs, int x, size_t sk, size_t se,
                       double *__restrict__ dst) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int nb2 = nb * nb, ij = (int)(t % nb2), i = ij / nb, j = ij % nb;
    if (j < i) return;
    const size_t k = (size_t)(t / nb2);
    const double re = src[((size_t)t * cs + x) * 2], im = src[((size_t)t * cs + x) * 2 + 1];
    if (i == j) {
        dst[k * sk + (size_t)i * se] = re;
    } else {
        const size_t e = nb + 2 * (size_t)pair_index(i, j, nb);
        dst[k * sk + e * se] = re;
        dst[k * sk + (e + 1) * se] = im;
    }
}

// D on the core <-> non-core block (rows i < nch, columns j >= nch) for the
// transient absorption, [k][3][nch (nb - nch)] complex, from a chunk of the
// reference [k][nb][nb][3] table.
__global__ void k_pack_dt(const double *__restrict__ src, int kc, int nb, int nch, double *__restrict__ dst) {
    const int nblk = nch * (nb - nch);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)kc * 3 * nblk) return;
    const int r = (int)(t % (3 * nblk)), x = r / nblk, q = r % nblk, i = q / (nb - nch), j = nch + q % (nb - nch);
    const size_t k = (size_t)(t / (3 * nblk));
    const double *d = src + (((k * nb + i) * nb + j) * 3 + x) * 2;
    dst[2 * t] = d[0];
    dst[2 * t + 1] = d[1];
}

__global__ void k_build_q(const double *Hp, const double *Dp, int nk, int nb, size_t sk, size_t se, size_t dxs, int dherm, double *Q) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)nk * 3 * nb * nb) return;
    const int j = (int)(id % nb), i = (int)((id / nb) % nb), x = (int)((id / ((long long)nb * nb)) % 3);
    const long long k = id / (3LL * nb * nb);
    auto el = [&](const double *base, int a, int b, bool herm) -> cd {
        if (a == b) return mk(base[k * sk + a * se], 0.0);
        const int p = pair_index(min(a, b), max(a, b), nb);
        const cd v = mk(base[k * sk + (nb + 2 * p) * se], base[k * sk + (nb + 2 * p + 1) * se]);
        return (a < b || !herm) ? v : cconj(v);
    };
    cd acc = mk(0.0, 0.0);
    for (int kc = 0; kc < nb; kc++) acc = cadd(acc, cmul(el(Hp, i, kc, true), cconj(el(Dp + x * dxs, j, kc, dherm != 0))));
    Q[2 * id] = acc.x;
    Q[2 * id + 1] = acc.y;
}

// ================================================================ captured steps
// The step counter of the field-table path: set by the host's table upload (in
// stream order) and advanced by one at the end of every captured step.
__global__ void k_counter_set(int *ctr, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *ctr = v;
}
__global__ void k_counter_step(int *ctr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *ctr += 1;
}

// ================================================================ adaptive dt
// Diagonal of the state before a step (P_W_prev, RungeKutta.cpp:1-9,139-147 and
// Taylor_DE_solver_deriv_Polynom.cpp:75-84).  Element e of k sits at k*sk + e*se.
__global__ void k_diag_save(const double *Y, double *DG, int nk, int nb, size_t sk, size_t se) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    for (int i = 0; i < nb; i++) DG[(size_t)i * nk + k] = Y[k * sk + i * se];
}

// Block maxima of the step-size measure of dyn_dt_evolution.cpp:74-87:
// 2 |P_ii(before) - P_ii(after)| / (cond + |P_ii(after)| + 1e-7), cond = P_cond_max + n_cond.
__global__ __launch_bounds__(256) void k_diag_dp(const double *Y, const double *DG, int nk, int nb, size_t sk, size_t se,
                                                 double cond, double *part) {
    __shared__ double red[4];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (k < nk)
        for (int i = 0; i < nb; i++) {
            const double y = Y[k * sk + i * se];
            m = fmax(m, 2. * fabs(DG[(size_t)i * nk + k] - y) / (cond + fabs(y) + 1.0e-7));
        }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// ================================================================ k-resolved snapshots
// Element (ic, jc) of conj(U) P_sym U^T per k, P_sym from the stored upper
// triangle, diagonal forced real (get_P_in_dia_vect, InitParameters.h:264-335).
__global__ void k_eigen_element(const double *Y, const double *U, int nk, int nb, size_t sk, size_t se, int ic, int jc, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double *u = U + (size_t)k * nb * nb * 2;
    auto pel = [&](int i, int j) -> cd {
        if (i == j) return mk(Y[k * sk + i * se], 0.0);
        const int a = min(i, j), b = max(i, j), p = pair_index(a, b, nb);
        const cd v = mk(Y[k * sk + (nb + 2 * p) * se], Y[k * sk + (nb + 2 * p + 1) * se]);
        return (i < j) ? v : cconj(v);
    };
    cd acc = mk(0.0, 0.0);
    for (int ii = 0; ii < nb; ii++) {
        cd m = mk(0.0, 0.0);
        for (int jj = 0; jj < nb; jj++) m = cadd(m, cmul(pel(ii, jj), mk(u[(jc * nb + jj) * 2], u[(jc * nb + jj) * 2 + 1])));
        acc = cadd(acc, cmul(mk(u[(ic * nb + ii) * 2], -u[(ic * nb + ii) * 2 + 1]), m));
    }
    if (ic == jc) acc.y = 0.0;
    out[2 * (size_t)k] = acc.x;
    out[2 * (size_t)k + 1] = acc.y;
}

// ================================================================ Wannier contraction
// Out[k][m] = sum_R phase(k,R) Val[R][m], Val = [H | r_x | r_y | r_z] per R.
// phase(k,R) = prod_l (-1)^{R_l} w_l[(i_l R_l) mod N_l] for N_l > 1 (k_l = -1/2 + i_l/N_l).
// Tiled complex GEMM: 64 k x 64 m per block, 16x16 threads with 4x4 outputs.
constexpr int WT = 64, WK = 8;

// Per-direction phase tables of the contraction, built once per call:
// PT[Off_l + r N_l + i] = tw[Off'_l + (i R_l mod N_l)] (-1)^{R_l}, the factor of
// direction l in e^{2 pi i k.R} for grid index i (exact integer reduction; the
// sign flip for odd R_l is the shift of the grid to k = -0.5 + i/N); 1 for a
// direction with a single k point.
__global__ void k_wannier_ptab(const int *R, const cd *tw, int nR, int N0, int N1, int N2, int toff1, int toff2, cd *PT) {
    const int Nn[3] = {N0, N1, N2}, Toff[3] = {0, toff1, toff2};
    const long long o1 = (long long)nR * N0, o2 = o1 + (long long)nR * N1, tot = o2 + (long long)nR * N2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
        const int l = t < o1 ? 0 : t < o2 ? 1 : 2;
        const long long u = t - (l == 0 ? 0 : l == 1 ? o1 : o2);
        const int r = (int)(u / Nn[l]), i = (int)(u % Nn[l]);
        const long long Rl = R[r * 3 + l];
        long long q = ((long long)i * Rl) % Nn[l];
        if (q < 0) q += Nn[l];
        cd w = tw[Toff[l] + q];
        if (Rl & 1) w = mk(-w.x, -w.y);
        PT[t] = Nn[l] == 1 ? mk(1.0, 0.0) : w;  // a direction without k points contributes no phase
    }
}

// Phase of grid point ka and lattice vector r from the tables above.
__device__ __forceinline__ cd wannier_phase(const cd *PT, int nR, int ka, int r, int N0, int N1, int N2) {
    const int i0 = ka / (N1 * N2), i1 = (ka / N2) % N1, i2 = ka % N2;
    const size_t o1 = (size_t)nR * N0, o2 = o1 + (size_t)nR * N1;
    cd ph = PT[(size_t)r * N0 + i0];
    ph = cmul(ph, PT[o1 + (size_t)r * N1 + i1]);
    return cmul(ph, PT[o2 + (size_t)r * N2 + i2]);
}

__global__ __launch_bounds__(256) void k_wannier(const cd *__restrict__ PT, const cd *__restrict__ Val, int N0, int N1, int N2, int nR,
                                                 int nm, int nk, int kbase, cd *__restrict__ Out) {
    __shared__ cd Ph[WT][WK + 1];
    __shared__ cd Vs[WK][WT + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int k0 = blockIdx.y * WT, m0 = blockIdx.x * WT;
    cd acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = mk(0.0, 0.0);
    for (int r0 = 0; r0 < nR; r0 += WK) {
        for (int t = threadIdx.x; t < WT * WK; t += 256) {
            const int kk = t / WK, rr = t % WK, k = k0 + kk, r = r0 + rr;
            Ph[kk][rr] = (k < nk && r < nR) ? wannier_phase(PT, nR, kbase + k, r, N0, N1, N2) : mk(0.0, 0.0);
        }
        for (int t = threadIdx.x; t < WK * WT; t += 256) {
            const int rr = t / WT, mm = t % WT, r = r0 + rr, m = m0 + mm;
            Vs[rr][mm] = (r < nR && m < nm) ? Val[(size_t)r * nm + m] : mk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < WK; rr++) {
            cd pv[4], vv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) pv[i] = Ph[ty * 4 + i][rr];
#pragma unroll
            for (int j = 0; j < 4; j++) vv[j] = Vs[rr][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    acc[i][j].x += pv[i].x * vv[j].x - pv[i].y * vv[j].y;
                    acc[i][j].y += pv[i].x * vv[j].y + pv[i].y * vv[j].x;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + ty * 4 + i, m = m0 + tx * 4 + j;
            if (k < nk && m < nm) Out[(size_t)k * nm + m] = acc[i][j];
        }
}

// The same contraction on the FP64 matrix cores: Out = Ph Val as four real
// products (Re = PhR ValR - PhI ValI, Im = PhR ValI + PhI ValR) of 16x16x4
// v_mfma_f64 tiles.  A workgroup of 4 waves owns 64 k x 32 columns (each wave
// 16 k x 32 columns: two 16x16 tiles with real and imaginary accumulators); R
// is consumed 16 at a time through LDS.  Lane l feeds A[row l%16][kk l/16] and
// B[kk l/16][col l%16] and holds D[l/16 + 4 i][l%16], i = 0..3 (the layout of
// the gfx950 instruction, checked against a host product).
constexpr int WMK = 64, WMN = 32, WMR = 16;

__global__ __launch_bounds__(256) void k_wannier_mfma(const cd *__restrict__ PT, const cd *__restrict__ Val, int N0, int N1, int N2,
                                                      int nR, int nm, int nk, int kbase, cd *__restrict__ Out) {
    __shared__ cd Ph[WMK][WMR + 1];
    __shared__ cd Vs[WMR][WMN + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k0 = blockIdx.y * WMK, m0 = blockIdx.x * WMN;
    v4d are[2], aim[2];
#pragma unroll
    for (int t = 0; t < 2; t++) are[t] = aim[t] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int r0 = 0; r0 < nR; r0 += WMR) {
        for (int t = threadIdx.x; t < WMK * WMR; t += 256) {
            const int kk = t / WMR, rr = t % WMR, k = k0 + kk, r = r0 + rr;
            Ph[kk][rr] = (k < nk && r < nR) ? wannier_phase(PT, nR, kbase + k, r, N0, N1, N2) : mk(0.0, 0.0);
        }
        for (int t = threadIdx.x; t < WMR * WMN; t += 256) {
            const int rr = t / WMN, mm = t % WMN, r = r0 + rr, m = m0 + mm;
            Vs[rr][mm] = (r < nR && m < nm) ? Val[(size_t)r * nm + m] : mk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < WMR / 4; ks++) {
            const cd a = Ph[wave * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const cd b = Vs[ks * 4 + (lane >> 4)][t * 16 + (lane & 15)];
                are[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, are[t], 0, 0, 0);
                are[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, are[t], 0, 0, 0);
                aim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, aim[t], 0, 0, 0);
                aim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, aim[t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + wave * 16 + (lane >> 4) + 4 * i, m = m0 + t * 16 + (lane & 15);
            if (k < nk && m < nm) Out[(size_t)k * nm + m] = mk(are[t][i], aim[t][i]);
        }
}

// The contraction in FP32 on the matrix cores (v_mfma_f32_16x16x4f32, FP32
// accumulation): the phases are formed in FP64 from the exact per-direction
// tables and rounded once into LDS, H(R) and r(R) likewise; the sums are widened
// to FP64 on the way out.  Same tiling as k_wannier_mfma; the FP32 instruction
// holds D[4 (l/16) + i][l%16] in lane l.
__global__ __launch_bounds__(256) void k_wannier_mfma32(const cd *__restrict__ PT, const cd *__restrict__ Val, int N0, int N1, int N2,
                                                        int nR, int nm, int nk, int kbase, cd *__restrict__ Out) {
    __shared__ cf Ph[WMK][WMR + 1];
    __shared__ cf Vs[WMR][WMN + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k0 = blockIdx.y * WMK, m0 = blockIdx.x * WMN;
    v4f are[2], aim[2];
#pragma unroll
    for (int t = 0; t < 2; t++) are[t] = aim[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int r0 = 0; r0 < nR; r0 += WMR) {
        for (int t = threadIdx.x; t < WMK * WMR; t += 256) {
            const int kk = t / WMR, rr = t % WMR, k = k0 + kk, r = r0 + rr;
            const cd ph = (k < nk && r < nR) ? wannier_phase(PT, nR, kbase + k, r, N0, N1, N2) : mk(0.0, 0.0);
            Ph[kk][rr] = cf{(float)ph.x, (float)ph.y};
        }
        for (int t = threadIdx.x; t < WMR * WMN; t += 256) {
            const int rr = t / WMN, mm = t % WMN, r = r0 + rr, m = m0 + mm;
            const cd v = (r < nR && m < nm) ? Val[(size_t)r * nm + m] : mk(0.0, 0.0);
            Vs[rr][mm] = cf{(float)v.x, (float)v.y};
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < WMR / 4; ks++) {
            const cf a = Ph[wave * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const cf b = Vs[ks * 4 + (lane >> 4)][t * 16 + (lane & 15)];
                are[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, are[t], 0, 0, 0);
                are[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.y, b.y, are[t], 0, 0, 0);
                aim[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, aim[t], 0, 0, 0);
                aim[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, aim[t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + wave * 16 + 4 * (lane >> 4) + i, m = m0 + t * 16 + (lane & 15);
            if (k < nk && m < nm) Out[(size_t)k * nm + m] = mk((double)are[t][i], (double)aim[t][i]);
        }
}

// The contraction as BF16x3 split products on the bf16 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation): every FP32 operand x is split
// into hi = bf16(x) and lo = bf16(x - hi), and each real product a b is taken as
// a_hi b_hi + a_hi b_lo + a_lo b_hi (the lo lo term, ~2^-16 |a b|, dropped).
// Phases and H(R), r(R) are formed in FP64 and rounded once to FP32.  Tiling as
// k_wannier_mfma32 with R consumed 32 at a time (one K slice of the
// instruction); lane l supplies A[row l%16][8 (l/16) .. +7] and B[same k][col l%16]
// and holds D[4 (l/16) + i][l%16].
constexpr int WBR = 32;

__global__ __launch_bounds__(256) void k_wannier_mfma_bf16x3(const cd *__restrict__ PT, const cd *__restrict__ Val, int N0, int N1,
                                                             int N2, int nR, int nm, int nk, int kbase, cd *__restrict__ Out) {
    // [hi, lo] planes of the real and imaginary parts, K (R) innermost with one pad
    __shared__ __bf16 Ph[2][2][WMK][WBR + 8];
    __shared__ __bf16 Vs[2][2][WMN][WBR + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k0 = blockIdx.y * WMK, m0 = blockIdx.x * WMN;
    v4f are[2], aim[2];
#pragma unroll
    for (int t = 0; t < 2; t++) are[t] = aim[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int r0 = 0; r0 < nR; r0 += WBR) {
        for (int t = threadIdx.x; t < WMK * WBR; t += 256) {
            const int kk = t / WBR, rr = t % WBR, k = k0 + kk, r = r0 + rr;
            const cd ph = (k < nk && r < nR) ? wannier_phase(PT, nR, kbase + k, r, N0, N1, N2) : mk(0.0, 0.0);
            bf16_split((float)ph.x, Ph[0][0][kk][rr], Ph[0][1][kk][rr]);
            bf16_split((float)ph.y, Ph[1][0][kk][rr], Ph[1][1][kk][rr]);
        }
        for (int t = threadIdx.x; t < WBR * WMN; t += 256) {
            const int rr = t / WMN, mm = t % WMN, r = r0 + rr, m = m0 + mm;
            const cd v = (r < nR && m < nm) ? Val[(size_t)r * nm + m] : mk(0.0, 0.0);
            bf16_split((float)v.x, Vs[0][0][mm][rr], Vs[0][1][mm][rr]);
            bf16_split((float)v.y, Vs[1][0][mm][rr], Vs[1][1][mm][rr]);
        }
        __syncthreads();
        const int ka = 8 * (lane >> 4);
        v8bf a[2][2], an[2];  // [re/im][hi/lo]; an: -Im (hi, lo)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int q = 0; q < 8; q++) a[c][h][q] = Ph[c][h][wave * 16 + (lane & 15)][ka + q];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int q = 0; q < 8; q++) an[h][q] = (__bf16)(-(float)a[1][h][q]);
#pragma unroll
        for (int t = 0; t < 2; t++) {
            v8bf b[2][2];
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int h = 0; h < 2; h++)
#pragma unroll
                    for (int q = 0; q < 8; q++) b[c][h][q] = Vs[c][h][t * 16 + (lane & 15)][ka + q];
            // Re += PhR VR - PhI VI, Im += PhR VI + PhI VR, each as hi hi + hi lo + lo hi
            auto prod3 = [&](const v8bf &xh, const v8bf &xl, const v8bf &yh, const v8bf &yl, v4f acc) {
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, yh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yl, acc, 0, 0, 0);
                return __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yh, acc, 0, 0, 0);
            };
            are[t] = prod3(a[0][0], a[0][1], b[0][0], b[0][1], are[t]);
            are[t] = prod3(an[0], an[1], b[1][0], b[1][1], are[t]);
            aim[t] = prod3(a[0][0], a[0][1], b[1][0], b[1][1], aim[t]);
            aim[t] = prod3(a[1][0], a[1][1], b[0][0], b[0][1], aim[t]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + wave * 16 + 4 * (lane >> 4) + i, m = m0 + t * 16 + (lane & 15);
            if (k < nk && m < nm) Out[(size_t)k * nm + m] = mk((double)are[t][i], (double)aim[t][i]);
        }
}

// Contraction output [k][c][i][j] (c = 0 H, 1..3 r) to the caller's layouts,
// one thread per (k, i, j): H with zero diagonal imaginary parts
// (Wannier.h:489-492) and D = (r + r^T)/2 with zero diagonal imaginary parts
// (Wannier.h:577-578, save_HUD.cpp:187-191), [k][i][j][x].  rtx: the r part of
// each output row is [i][j][x] (the caller's own order) instead of [x][i][j].
__global__ void k_wannier_finish(const cd *__restrict__ Out, int kc, int nw, double *__restrict__ Hd, double *__restrict__ Dd,
                                 int rtx) {
    const int nw2 = nw * nw;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)kc * nw2) return;
    const int kk = (int)(t / nw2), ij = (int)(t % nw2), i = ij / nw, j = ij % nw;
    const cd *src = Out + (size_t)kk * 4 * nw2;
    const cd h = src[ij];
    Hd[2 * t] = h.x;
    Hd[2 * t + 1] = (i == j) ? 0.0 : h.y;
    if (!Dd) return;
#pragma unroll
    for (int x = 0; x < 3; x++) {
        const cd a = rtx ? src[nw2 + (i * nw + j) * 3 + x] : src[(x + 1) * nw2 + i * nw + j];
        const cd b = rtx ? src[nw2 + (j * nw + i) * 3 + x] : src[(x + 1) * nw2 + j * nw + i];
        const double ai = (i == j) ? 0.0 : a.y, bi = (i == j) ? 0.0 : b.y;
        Dd[(3 * t + x) * 2] = 0.5 * (a.x + b.x);
        Dd[(3 * t + x) * 2 + 1] = 0.5 * (ai + bi);
    }
}

// ================================================================ Coulomb coefficient generator
// W(q) = sum_G V(|q+G|) e^{i (q+G).(d_i - d_j)} on a grid of q (Create_Wq_with_G_sum,
// Coulomb.h:87-175), one thread per q point, the G list streamed through LDS in
// tiles (an all-pairs sum: every q meets every G).  The phase is split as
// e^{i q.beta} e^{i G.beta} with beta_p = b.(d_i - d_j) per band pair p.
constexpr int kCgTile = 256;
constexpr int kCgMaxPair = 6;  // nb <= 4

struct CGenArgs {
    int npair, n_g, N, mode;    // mode 0: grid q = -0.5 + i h (N per direction); 1: list of (ix, iy) on that grid
    double h, qtf2, eps, r0, rk;
    double b1[3], b2[3];
    double beta1[kCgMaxPair], beta2[kCgMaxPair];
    const double *g;            // [n_g][2] crystal offsets
    const double *gph;          // [n_g][npair][2] e^{i G.beta_p}
    const int *qlist;           // mode 1: [nq][2]
};

__global__ __launch_bounds__(256) void k_cgen_w(CGenArgs a, long long iq0, int nq, double *W) {
    __shared__ double sg[kCgTile * 2];
    __shared__ double sph[kCgTile * kCgMaxPair * 2];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < nq;
    int ix = 0, iy = 0;
    if (valid) {
        if (a.mode == 0) {
            const long long iq = iq0 + t;
            ix = (int)(iq / a.N);
            iy = (int)(iq % a.N);
        } else {
            ix = a.qlist[2 * t];
            iy = a.qlist[2 * t + 1];
        }
    }
    // Grid_and_Simpson (integrationWeight.h:5-8): q_i = -0.5 + i h
    const double qx = -0.5 + ix * a.h, qy = -0.5 + iy * a.h;
    double pr[kCgMaxPair], pi[kCgMaxPair], wr[kCgMaxPair], wi[kCgMaxPair];
    for (int p = 0; p < a.npair; p++) {
        double s, c;
        sincos(qx * a.beta1[p] + qy * a.beta2[p], &s, &c);
        pr[p] = c;
        pi[p] = s;
        wr[p] = wi[p] = 0.0;
    }
    double w00 = 0.0;
    const double twopi = 2 * M_PI;
    for (int g0 = 0; g0 < a.n_g; g0 += kCgTile) {
        const int nt = min(kCgTile, a.n_g - g0);
        __syncthreads();
        for (int i = threadIdx.x; i < nt * 2; i += blockDim.x) sg[i] = a.g[(size_t)g0 * 2 + i];
        for (int i = threadIdx.x; i < nt * a.npair * 2; i += blockDim.x) sph[i] = a.gph[(size_t)g0 * a.npair * 2 + i];
        __syncthreads();
        for (int j = 0; j < nt; j++) {
            const double c1 = qx + sg[2 * j], c2 = qy + sg[2 * j + 1];
            const double x = a.b1[0] * c1 + a.b2[0] * c2, y = a.b1[1] * c1 + a.b2[1] * c2, z = a.b1[2] * c1 + a.b2[2] * c2;
            const double q2 = x * x + y * y + z * z;
            // Coloumb_THh_F (Coulomb.h:60-84)
            const double den = (a.rk != 0.0) ? a.eps * sqrt(q2 + a.qtf2) * (sqrt(q2) * a.r0 + 1.0) : a.eps * sqrt(q2 + a.qtf2);
            const double V = twopi / den;
            w00 += V;
            for (int p = 0; p < a.npair; p++) {
                const double gr = sph[(j * a.npair + p) * 2], gi = sph[(j * a.npair + p) * 2 + 1];
                const double er = pr[p] * gr - pi[p] * gi, ei = pr[p] * gi + pi[p] * gr;
                wr[p] += V * er;
                wi[p] += V * ei;
            }
        }
    }
    if (!valid) return;
    const int nw = 1 + 2 * a.npair;
    double *o = W + (size_t)t * nw;
    o[0] = w00;
    for (int p = 0; p < a.npair; p++) {
        o[1 + 2 * p] = wr[p];
        o[2 + 2 * p] = wi[p];
    }
}

// T[t][iy][c] = sum_ix w[ix] trig_t(ix) W[ix][iy][c], trig_t = cos(2 pi m q), t < ncut,
// sin(2 pi (t - ncut) q) otherwise (first pass of the separable Simpson integral,
// Fourier_Coulomb.cpp:230-265 with Coulomb.h:182-232).
__global__ void k_cgen_pass1(const double *W, const double *wtrig, int N, int ntrig, int nw, double *T) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)ntrig * nw * N) return;
    const int iy = (int)(id % N), c = (int)((id / N) % nw), t = (int)(id / ((long long)N * nw));
    double s = 0.0;
    for (int ix = 0; ix < N; ix++) s += wtrig[(size_t)t * N + ix] * W[((size_t)ix * N + iy) * nw + c];
    T[((size_t)t * N + iy) * nw + c] = s;
}

// S[t1][t2][c] = sum_iy w[iy] trig_t2(iy) T[t1][iy][c].
__global__ void k_cgen_pass2(const double *T, const double *wtrig, int N, int ntrig, int nw, double *S) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ntrig * ntrig * nw) return;
    const int c = id % nw, t2 = (id / nw) % ntrig, t1 = id / (nw * ntrig);
    double s = 0.0;
    for (int iy = 0; iy < N; iy++) s += wtrig[(size_t)t2 * N + iy] * T[((size_t)t1 * N + iy) * nw + c];
    S[id] = s;
}

// Correction around q = 0 (Fourier_Coulomb.cpp:375-456): per listed point,
// (W0 - Fourier series) * dk2 for the entries ic <= jc, block partial sums.
__global__ __launch_bounds__(256) void k_cgen_peak(const double *W0, const int *qlist, int np, int nb, int ncut, double h, double dk2,
                                                   const double *A, const double *B, const double *C, const double *D, double *part) {
    __shared__ double red[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb2 = nb * nb, npair = nb * (nb - 1) / 2, nw = 1 + 2 * npair;
    double qx = 0, qy = 0;
    if (t < np) {
        qx = -0.5 + qlist[2 * t] * h;
        qy = -0.5 + qlist[2 * t + 1] * h;
    }
    int p = 0;
    for (int ic = 0; ic < nb; ic++)
        for (int jc = ic; jc < nb; jc++) {
            const int idx = ic * nb + jc;
            double vr = 0, vi = 0;
            if (t < np) {
                // Fourier_series_Coloumb (Coulomb.h:235-267)
                for (int m = 0; m < ncut; m++) {
                    double sm, cm;
                    sincos(2 * M_PI * m * qx, &sm, &cm);
                    for (int n = 0; n < ncut; n++) {
                        double sn, cn;
                        sincos(2 * M_PI * n * qy, &sn, &cn);
                        const size_t r = ((size_t)(n + ncut * m) * nb2 + idx) * 2;
                        vr += A[r] * cm * cn + B[r] * cm * sn + C[r] * sm * cn + D[r] * sm * sn;
                        vi += A[r + 1] * cm * cn + B[r + 1] * cm * sn + C[r + 1] * sm * cn + D[r + 1] * sm * sn;
                    }
                }
                const double *w = W0 + (size_t)t * nw;
                double wr = w[0], wi = 0.0;
                if (jc != ic) { wr = w[1 + 2 * p]; wi = w[2 + 2 * p]; }
                vr = (wr - vr) * dk2;
                vi = (wi - vi) * dk2;
            }
            if (jc != ic) p++;
            for (int comp = 0; comp < 2; comp++) {
                red[threadIdx.x] = comp ? vi : vr;
                __syncthreads();
                for (int s = 128; s > 0; s >>= 1) {
                    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                    __syncthreads();
                }
                if (threadIdx.x == 0) part[((size_t)blockIdx.x * nb2 + idx) * 2 + comp] = red[0];
                __syncthreads();
            }
        }
}

}  // namespace

// ==================================================================== context
struct edus_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nk = 0, nkp = 0, nb = 0, np = 0, nch = 0, ncc = 0, n_nbr = 0, stepper = 0, order = 1, diss = 0, dherm = 0;
    int sw = 0, scol = 0;  // strip order of the wave kernels (strip_k)
    int pair = 0, pself = -1, npnew = 0;  // stencil pairs of the wave kernels (ConstArgs::pair)
    int tpb = 0;           // threads per workgroup of the nb <= 4 kernel (0: by grid size)
    int precision = EDUS_PRECISION_FP64;
    float *Hf = nullptr, *Df = nullptr, *Uf = nullptr;  // FP32 operator copies (mixed precision)
    int *eidx = nullptr;  // packed element -> (i, j, comp) of the wave layout
    int dmask = 7;        // D components some field of the current call needs (see upload_table)
    bool soa = true;
    bool own_stream = true;  // false when the stream belongs to a batch
    double dt = 0, wx2 = 0, j_norm = 0, tabs_norm = 0, nktot = 0;
    std::vector<double> tinv, coef;
    std::vector<double> wb;  // [n_nbr][3]
    // device buffers
    double *P0 = nullptr, *ACC = nullptr, *YA = nullptr, *YB = nullptr, *HIST = nullptr;
    double *H = nullptr, *D = nullptr, *U = nullptr, *T = nullptr, *GE = nullptr, *Q = nullptr, *DT = nullptr;
    int *nbr = nullptr;
    double *tab = nullptr;
    size_t tab_cap = 0;
    double *part = nullptr, *obs = nullptr;
    int obs_blocks = 0;
    int cur = 0;          // AB/Taylor: state in P0 (0) or YA (1)
    long long nstep = 0;  // steps done (history ring position)
    float last_ms = 0;
    int last_launches = 0;
    std::vector<void *> allocs;
    // Coulomb mean field
    int coul = 0, coul_diag = 0, N1 = 0, N2 = 0, ncut = 0, nu = 0;
    double *cPD = nullptr;  // diagonal basis: packed state in the eigenbasis
    double n_cond = 0.0;
    double *cxm = nullptr, *sxm = nullptr, *cyn = nullptr, *syn = nullptr, *cS = nullptr, *cF = nullptr;
    double *cX = nullptr, *cX0 = nullptr, *cA = nullptr, *cB = nullptr, *cC = nullptr, *cD = nullptr, *cvh = nullptr;
    double *cEH = nullptr, *cEH0 = nullptr, *cT = nullptr, *cPW0 = nullptr, *cHr = nullptr, *cscreen = nullptr, *cXw = nullptr;
    double *cUexc = nullptr, *cGE = nullptr, *cQ = nullptr, *dwb = nullptr, *cEk = nullptr;
    bool cEk_valid = false;
    double *snap = nullptr;  // [nk] complex, k-resolved snapshot element
    // adaptive dt: diagonal before the step, block maxima
    double *DG = nullptr, *dpart = nullptr;
    // device-resident field table of a run and the captured RK4 step reading it
    double *gtab = nullptr;
    size_t gtab_cap = 0;
    int *gctr = nullptr;          // rows consumed, in steps (advanced by the captured step)
    long long gtab_steps = 0, gpos = 0;
    int gdmask = 7;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace {

template <class T>
int dalloc(edus_ctx *c, T **p, size_t n) {
    void *q = nullptr;
    HIP_TRY(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    c->allocs.push_back(q);
    *p = static_cast<T *>(q);
    return 0;
}

// Packs reference [k][nb][nb] complex (upper triangle) into the state layout.
void pack_state(const edus_ctx *c, const double *full, std::vector<double> &out) {
    const int nb = c->nb, np = c->np;
    out.assign(c->soa ? (size_t)np * c->nkp : (size_t)np * c->nk, 0.0);
    for (int k = 0; k < c->nk; k++) {
        const double *m = full + (size_t)k * nb * nb * 2;
        for (int i = 0; i < nb; i++) {
            const size_t e = i;
            out[c->soa ? e * c->nkp + k : (size_t)k * np + e] = m[(i * nb + i) * 2];
            for (int j = i + 1; j < nb; j++) {
                const int p = pair_index(i, j, nb);
                const size_t er = nb + 2 * p;
                out[c->soa ? er * c->nkp + k : (size_t)k * np + er] = m[(i * nb + j) * 2];
                out[c->soa ? (er + 1) * c->nkp + k : (size_t)k * np + er + 1] = m[(i * nb + j) * 2 + 1];
            }
        }
    }
}

void unpack_state(const edus_ctx *c, const std::vector<double> &in, double *full) {
    const int nb = c->nb, np = c->np;
    for (int k = 0; k < c->nk; k++) {
        double *m = full + (size_t)k * nb * nb * 2;
        auto at = [&](size_t e) { return in[c->soa ? e * c->nkp + k : (size_t)k * np + e]; };
        for (int i = 0; i < nb; i++) {
            m[(i * nb + i) * 2] = at(i);
            m[(i * nb + i) * 2 + 1] = 0.0;
            for (int j = i + 1; j < nb; j++) {
                const int p = pair_index(i, j, nb);
                const double re = at(nb + 2 * p), im = at(nb + 2 * p + 1);
                m[(i * nb + j) * 2] = re;
                m[(i * nb + j) * 2 + 1] = im;
                m[(j * nb + i) * 2] = re;
                m[(j * nb + i) * 2 + 1] = -im;
            }
        }
    }
}

// A host table of nk rows (each nb x nb elements of cs complex values, the
// reference layout) copied in chunks of stage_k rows through the device buffer
// `stage`; pack(k0, kc) launches the packing of each chunk, in stream order.
template <class F>
int upload_packed(edus_ctx *c, const double *src, int cs, double *stage, int stage_k, F pack) {
    const size_t per = (size_t)c->nb * c->nb * cs * 2;
    for (int k0 = 0; k0 < c->nk; k0 += stage_k) {
        const int kc = std::min(stage_k, c->nk - k0);
        HIP_TRY(hipMemcpyAsync(stage, src + (size_t)k0 * per, (size_t)kc * per * sizeof(double), hipMemcpyHostToDevice, c->stream));
        pack(k0, kc);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

void launch_pack(const edus_ctx *c, const double *stage, int kc, int cs, int x, size_t sk, size_t se, double *dst) {
    const long long n = (long long)kc * c->nb * c->nb;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, stage, n, c->nb, cs, x, sk, se, dst);
}

struct DeviceFree {
    void operator()(double *p) const { (void)hipFree(p); }
};

int upload(edus_ctx *c, double *dst, const std::vector<double> &v) {
    HIP_TRY(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return 0;
}

int state_elems(const edus_ctx *c) { return c->soa ? c->np * c->nkp : c->np * c->nk; }

// ---------------------------------------------------------------- launches
CoulArgs coul_args(edus_ctx *c, const double *Y) {
    CoulArgs a;
    a.Y = Y; a.nk = c->nk; a.nkp = c->nkp; a.nb = c->nb; a.NP = c->np; a.N1 = c->N1; a.N2 = c->N2; a.ncut = c->ncut; a.nu = c->nu;
    a.w = 1.0 / (double)c->nk;
    a.cxm = c->cxm; a.sxm = c->sxm; a.cyn = c->cyn; a.syn = c->syn; a.S = c->cS; a.F = c->cF; a.X = c->cX; a.X0 = c->cX0;
    a.A = c->cA; a.B = c->cB; a.C = c->cC; a.D = c->cD; a.vh = c->cvh; a.EH = c->cEH; a.EH0 = c->cEH0; a.Tt = c->cT;
    a.Xw = c->cXw;
    return a;
}

// Mean-field coefficients of state Y: reduction, X_mn, column sums (4 launches).
int launch_coulomb(edus_ctx *c, const double *Y) {
    CoulArgs a = coul_args(c, Y);
    const int t3 = c->N2 * c->ncut * c->nu;
    // 64-thread workgroups: the passes have a few 10^4 outputs each, and small
    // workgroups spread them over all CUs.  Pass 1 runs a wave per 16 x 16 tile.
    hipLaunchKernelGGL(k_coul_s1_mfma, dim3((c->np * c->N1 + 15) / 16, (2 * c->ncut + 15) / 16), dim3(64), 0, c->stream, a);
    hipLaunchKernelGGL(k_coul_s2, dim3(c->ncut * c->ncut), dim3(kCoulThreads), 0, c->stream, a);
    hipLaunchKernelGGL(k_coul_s3, dim3((t3 + 63) / 64), dim3(64), 0, c->stream, a);
    hipLaunchKernelGGL(k_coul_w_mfma, dim3((c->N1 + 15) / 16, (2 * c->nu * c->N2 + 15) / 16), dim3(64), 0, c->stream, a,
                       c->N1);
    HIP_TRY(hipGetLastError());
    return 0;
}

// The mean-field sums of the stage input Y: over Y itself (Wannier basis) or over
// Y in the eigenbasis (Coulomb_diag_basis, RungeKutta.h:41-55).
int launch_coulomb_of(edus_ctx *c, const double *Y) {
    if (!c->coul_diag) return launch_coulomb(c, Y);
    const dim3 grid((c->nk + 255) / 256), block(256);
    switch (c->nb) {
        case 1: hipLaunchKernelGGL(k_coul_pdiag<1>, grid, block, 0, c->stream, Y, (const double *)c->U, c->nk, c->nkp, c->cPD); break;
        case 2: hipLaunchKernelGGL(k_coul_pdiag<2>, grid, block, 0, c->stream, Y, (const double *)c->U, c->nk, c->nkp, c->cPD); break;
        case 3: hipLaunchKernelGGL(k_coul_pdiag<3>, grid, block, 0, c->stream, Y, (const double *)c->U, c->nk, c->nkp, c->cPD); break;
        default: hipLaunchKernelGGL(k_coul_pdiag<4>, grid, block, 0, c->stream, Y, (const double *)c->U, c->nk, c->nkp, c->cPD); break;
    }
    HIP_TRY(hipGetLastError());
    return launch_coulomb(c, c->cPD);
}

ConstArgs const_args(const edus_ctx *c) {
    ConstArgs ca;
    std::memset(&ca, 0, sizeof ca);
    ca.H = c->H; ca.D = c->D; ca.U = c->U; ca.T = c->T; ca.nbr = c->nbr;
    ca.Hf = c->Hf; ca.Df = c->Df; ca.Uf = c->Uf; ca.eidx = c->eidx;
    ca.nk = c->nk; ca.nkp = c->nkp; ca.nb = c->nb; ca.nch = c->nch; ca.ncc = c->ncc; ca.n_nbr = c->n_nbr;
    ca.diss = c->diss; ca.dherm = c->dherm; ca.wx2 = c->wx2; ca.sw = c->sw; ca.scol = c->scol;
    ca.pair = c->pair; ca.pself = c->pself; ca.npnew = c->npnew; ca.pat = c->nbr + (size_t)std::max(c->n_nbr, 1) * c->nkp;
    ca.coul = c->coul; ca.N2 = c->N2; ca.ncut = c->ncut; ca.nu = c->nu; ca.n_cond = c->n_cond;
    ca.cxm = c->cxm; ca.sxm = c->sxm; ca.Tt = c->cT; ca.PW0 = c->cPW0; ca.EH = c->cEH; ca.screen = c->cscreen; ca.Hr = c->cHr;
    ca.Xw = c->cXw;
    ca.coul_diag = c->coul_diag;
    ca.PD = c->cPD;
    return ca;
}

template <int MODE>
int launch_stage(edus_ctx *c, StageArgs a) {
    a.dmask = c->dmask;
    const ConstArgs ca = const_args(c);
    if (c->coul) {
        int rc = launch_coulomb_of(c, a.Y);
        if (rc) return rc;
    }
    if (c->soa) {
        // A grid of fewer k than 256 per CU runs 64-thread workgroups, so that
        // more CUs share it (the k past nk, up to the padded nkp, are computed
        // and never read).
        const int tpb = c->tpb ? c->tpb : (c->nk < 256 * 256 ? 64 : 256), blocks = (c->nk + tpb - 1) / tpb;
        auto go = [&](auto dtag) {
            constexpr bool DISS = decltype(dtag)::value;
            switch (c->nb) {
                case 1: hipLaunchKernelGGL((k_stage_small<1, MODE, DISS>), dim3(blocks), dim3(tpb), 0, c->stream, ca, a); break;
                case 2: hipLaunchKernelGGL((k_stage_small<2, MODE, DISS>), dim3(blocks), dim3(tpb), 0, c->stream, ca, a); break;
                case 3: hipLaunchKernelGGL((k_stage_small<3, MODE, DISS>), dim3(blocks), dim3(tpb), 0, c->stream, ca, a); break;
                default: hipLaunchKernelGGL((k_stage_small<4, MODE, DISS>), dim3(blocks), dim3(tpb), 0, c->stream, ca, a); break;
            }
        };
        if (c->nb < 1 || c->nb > 4) return set_err(EDUS_ERR_UNSUPPORTED, "no small kernel for this band count");
        if (c->diss) go(std::true_type{});
        else go(std::false_type{});
    } else {
        auto go = [&](auto tag, auto ptag) {
            constexpr int NBT = decltype(tag)::value, KPB = WaveGeom<NBT>::KPB;
            constexpr bool PAIR = decltype(ptag)::value;
            const dim3 grid(PAIR ? c->nk / 2 : (c->nk + KPB - 1) / KPB), block(WaveGeom<NBT>::TPK * KPB);
            const size_t lds_f = (size_t)KPB * wave_lds_doubles<NBT, float>(c->diss) * 8;
            const size_t lds_d = (size_t)KPB * wave_lds_doubles<NBT, double>(c->diss) * 8;
            if constexpr (NBT >= 16 && !PAIR) {
                if (c->precision == EDUS_PRECISION_BF16X3) {
                    hipLaunchKernelGGL((k_stage_wave<NBT, MODE, float, double, true>), grid, block, lds_f, c->stream, ca, a);
                    return;
                }
            }
            if (c->precision == EDUS_PRECISION_MIXED32)
                hipLaunchKernelGGL((k_stage_wave<NBT, MODE, float, float, false, PAIR>), grid, block, lds_f, c->stream, ca, a);
            else if (c->precision == EDUS_PRECISION_MIXED)
                hipLaunchKernelGGL((k_stage_wave<NBT, MODE, float, double, false, PAIR>), grid, block, lds_f, c->stream, ca, a);
            else
                hipLaunchKernelGGL((k_stage_wave<NBT, MODE, double, double, false, PAIR>), grid, block, lds_d, c->stream, ca, a);
        };
        // stencil pairs only where a workgroup serves one k (nb > 8; see
        // stencil_pairs), and where they measured faster: not in the BF16x3 form,
        // whose split products leave no registers for the second k's stencil
        // sum, nor in the 32-band Taylor step with FP32 storage (3.73 ms paired,
        // 3.53 one k per workgroup; its RK4 stages gain: 2.21 against 2.26 ms)
        const bool pair = c->pair && c->precision != EDUS_PRECISION_BF16X3 &&
                          !(MODE == TAY && c->nb > 16 && c->precision == EDUS_PRECISION_MIXED32);
        if (c->nb <= 8) go(std::integral_constant<int, 8>{}, std::false_type{});
        else if (c->nb <= 16 && pair) go(std::integral_constant<int, 16>{}, std::true_type{});
        else if (c->nb <= 16) go(std::integral_constant<int, 16>{}, std::false_type{});
        else if (pair) go(std::integral_constant<int, 32>{}, std::true_type{});
        else go(std::integral_constant<int, 32>{}, std::false_type{});
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Field table rows for n stages: E (6), gradient coefficients (n_nbr), flag.
// A field component that is zero in every row of the call multiplies its
// dipole component by an exact zero, so the stage kernels skip loading it (the
// sums are bit-identical).  Rows are [E_pump(3), E_probe/2(3), ...] of length rs.
int upload_table(edus_ctx *c, const std::vector<double> &rows) {
    const size_t rs = 7 + (size_t)c->n_nbr;
    int mask = 0;
    for (size_t r = 0; r + rs <= rows.size(); r += rs)
        for (int x = 0; x < 3; x++)
            if (rows[r + x] != 0.0 || rows[r + 3 + x] != 0.0) mask |= 1 << x;
    c->dmask = mask;
    if (rows.size() > c->tab_cap) {
        if (c->tab) {
            HIP_TRY(hipStreamSynchronize(c->stream));
            HIP_TRY(hipFree(c->tab));
            c->tab = nullptr;
        }
        size_t cap = std::max<size_t>(rows.size(), 4096);
        HIP_TRY(hipMalloc(&c->tab, cap * sizeof(double)));
        c->tab_cap = cap;
    }
    HIP_TRY(hipMemcpyAsync(c->tab, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return 0;
}

void table_row(const edus_ctx *c, const double *E, uint8_t on, double *row) {
    for (int i = 0; i < 6; i++) row[i] = E[i];
    for (int n = 0; n < c->n_nbr; n++)
        row[6 + n] = E[0] * c->wb[n * 3] + E[1] * c->wb[n * 3 + 1] + E[2] * c->wb[n * 3 + 2];
    row[6 + c->n_nbr] = on ? 1.0 : 0.0;
}

int check_ctx(edus_ctx *c) {
    if (!c) return set_err(EDUS_ERR_INVALID, "null context");
    hipError_t e = hipSetDevice(c->dev);
    if (e != hipSuccess) return set_err(EDUS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return 0;
}

// Stencil overlap of the pairs (ka, kb) = (2q, 2q + 1) of a neighbour table
// nbr[k][nn]: each slot j of kb's stencil is matched to the first unmatched slot
// of ka's holding the same point (pat[i] = j for ka's slot i), else to ka itself
// (pself = j), else listed as new (pat[kMaxNbr + t] = j, t < npnew).  True when
// every pair has the same pattern, as on a uniform grid of even length along the
// fastest axis; the wave kernels then serve a pair per workgroup.
bool stencil_pairs(const int32_t *nbr, int nk, int nn, int *pat, int &pself, int &npnew) {
    if (nn > kMaxNbr || nk < 2 || nk % 2) return false;
    struct Pattern {
        int pm[kMaxNbr], pn[kMaxNbr], self, cnt;
    };
    auto pattern = [&](int q, Pattern &P) {
        const int32_t *A = nbr + (size_t)(2 * q) * nn, *B = A + nn;
        uint64_t used = 0;
        P.self = -1;
        P.cnt = 0;
        for (int i = 0; i < nn; i++) P.pm[i] = -1;
        for (int j = 0; j < nn; j++) {
            int hit = -1;
            for (int i = 0; i < nn && hit < 0; i++)
                if (!(used >> i & 1) && A[i] == B[j]) hit = i;
            if (hit >= 0) {
                P.pm[hit] = j;
                used |= 1ull << hit;
            } else if (B[j] == 2 * q && P.self < 0) {
                P.self = j;
            } else {
                P.pn[P.cnt++] = j;
            }
        }
    };
    Pattern P0, P;
    pattern(0, P0);
    for (int q = 1; q < nk / 2; q++) {
        pattern(q, P);
        if (P.self != P0.self || P.cnt != P0.cnt || std::memcmp(P.pm, P0.pm, nn * sizeof(int)) ||
            std::memcmp(P.pn, P0.pn, P.cnt * sizeof(int)))
            return false;
    }
    for (int i = 0; i < nn; i++) pat[i] = P0.pm[i];
    for (int t = 0; t < P0.cnt; t++) pat[kMaxNbr + t] = P0.pn[t];
    pself = P0.self;
    npnew = P0.cnt;
    return true;
}

}  // namespace

// ==================================================================== fixed scratch
// Device buffers of edus_wannier_hk_noalloc: one copy of this array per device,
// held from the library's load until the process ends.  Host code never uses
// the name itself as an address (on the host it is only a stand-in): the
// device address comes from hipGetSymbolAddress, per device (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[EDUS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call with that device current and kept.
int noalloc_scratch(void **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_err(EDUS_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

// ==================================================================== C-ABI
extern "C" {

const char *edus_last_error(void) { return g_err.c_str(); }

int edus_create(const edus_problem *p, edus_ctx **out) {
    if (!p || !out) return set_err(EDUS_ERR_INVALID, "null argument");
    *out = nullptr;
    const int nb = p->nb;
    if (p->nk < 1 || nb < 1 || nb > EDUS_MAX_NB) return set_err(EDUS_ERR_INVALID, "nk/nb out of range");
    if (p->nch < 0 || p->nch > nb || p->ncc < p->nch || p->ncc > nb) return set_err(EDUS_ERR_INVALID, "core/filled band counts out of range");
    if (p->n_nbr < 0 || p->n_nbr > kMaxNbr) return set_err(EDUS_ERR_INVALID, "neighbour count out of range (0..64)");
    if (!p->H || !p->D || !p->U || !p->T || !p->rho0 || (p->n_nbr > 0 && (!p->nbr || !p->nbr_wb)))
        return set_err(EDUS_ERR_INVALID, "missing table");
    if (p->stepper < 0 || p->stepper > 2) return set_err(EDUS_ERR_INVALID, "unknown stepper");
    if (p->stepper != EDUS_STEPPER_RK4 && (p->order < 1 || p->order > EDUS_MAX_ORDER))
        return set_err(EDUS_ERR_INVALID, "order out of range");
    if (p->stepper == EDUS_STEPPER_TAYLOR && (p->order < 2 || !p->taylor_tau_inv))
        return set_err(EDUS_ERR_INVALID, "Taylor needs order >= 2 and tau^-1");
    // Neighbour indices must address the grid: the gradient gather trusts them.
    for (size_t i = 0; i < (size_t)p->nk * p->n_nbr; i++)
        if (p->nbr[i] < 0 || p->nbr[i] >= p->nk) return set_err(EDUS_ERR_INVALID, "neighbour index outside the grid");
    // The packed operator layout needs D symmetric (tight binding, the Wannier
    // (D + D^T)/2) or Hermitian (the crystal interface), with a real diagonal.
    bool dsym = true, dhermit = true;
    for (int k = 0; k < p->nk; k++)
        for (int x = 0; x < 3; x++)
            for (int i = 0; i < nb; i++) {
                const double *dii = p->D + (((size_t)k * nb + i) * nb + i) * 6 + 2 * x;
                if (dii[1] != 0.0) return set_err(EDUS_ERR_UNSUPPORTED, "dipole diagonal must be real");
                for (int j = i + 1; j < nb; j++) {
                    const double *a = p->D + (((size_t)k * nb + i) * nb + j) * 6 + 2 * x;
                    const double *b = p->D + (((size_t)k * nb + j) * nb + i) * 6 + 2 * x;
                    if (a[0] != b[0] || a[1] != b[1]) dsym = false;
                    if (a[0] != b[0] || a[1] != -b[1]) dhermit = false;
                }
            }
    if (!dsym && !dhermit) return set_err(EDUS_ERR_UNSUPPORTED, "dipole table must be symmetric or Hermitian");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(EDUS_ERR_NODEVICE, "no HIP device visible");

    auto *c = new edus_ctx();
    c->dherm = (!dsym && dhermit) ? 1 : 0;
    auto fail = [&](int rc) { edus_destroy(c); return rc; };
    if (hipGetDevice(&c->dev) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "stream create failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "event create failed"));
    c->nk = p->nk; c->nb = nb; c->np = nb * nb; c->nch = p->nch; c->ncc = p->ncc; c->n_nbr = p->n_nbr;
    c->stepper = p->stepper; c->order = (p->stepper == EDUS_STEPPER_RK4) ? 4 : p->order;
    c->soa = nb <= 4;
    c->precision = p->precision;
    if (c->precision != EDUS_PRECISION_FP64 && c->precision != EDUS_PRECISION_MIXED && c->precision != EDUS_PRECISION_MIXED32 &&
        c->precision != EDUS_PRECISION_BF16X3)
        return fail(set_err(EDUS_ERR_INVALID, "precision must be EDUS_PRECISION_FP64, _MIXED, _MIXED32 or _BF16X3"));
    if (c->precision != EDUS_PRECISION_FP64 && c->soa)
        return fail(set_err(EDUS_ERR_UNSUPPORTED, "mixed precision is implemented for nb >= 5"));
    if (c->precision == EDUS_PRECISION_BF16X3 && nb <= 8)
        return fail(set_err(EDUS_ERR_UNSUPPORTED, "EDUS_PRECISION_BF16X3 runs on the matrix-core tiles of nb >= 9"));
    c->nkp = c->soa ? ((p->nk + 255) / 256) * 256 : p->nk;
    // Wave kernels: walk the grid in strips (strip_k) when a strip's two previous
    // rows of packed state, the stencil's reach, fit in kStripL2Bytes while whole
    // rows would not (512 x 512 k at 32 bands: 64-column strips).
    if (!c->soa) {
        const long long n0 = p->nk_dims[0], n1 = p->nk_dims[1], n2 = p->nk_dims[2];
        if (n0 > 0 && n1 > 0 && n2 > 0 && n0 * n1 * n2 == p->nk) {
            const long long row_k = 2LL * 8 * nb * nb;  // two rows' bytes per column
            long long sw = 16;
            while (2 * sw * row_k <= kStripL2Bytes) sw *= 2;
            if (sw < n2) {
                c->sw = (int)sw;
                c->scol = (int)n2;
            }
        }
    }
    c->dt = p->dt; c->wx2 = p->wx2; c->j_norm = p->j_norm; c->tabs_norm = p->tabs_norm; c->nktot = p->nktot;
    c->wb.assign(p->nbr_wb, p->nbr_wb + (size_t)p->n_nbr * 3);
    for (int i = 0; i < nb * nb; i++)
        if (std::fabs(p->T[i]) > 1e-24) c->diss = 1;
    if (p->stepper == EDUS_STEPPER_TAYLOR) c->tinv.assign(p->taylor_tau_inv, p->taylor_tau_inv + (size_t)(p->order - 1) * (p->order - 1));
    // Adams-Bashforth coefficients (ODE_solver.h:152-179).
    if (p->stepper == EDUS_STEPPER_AB) {
        static const double ab[5][5] = {{1., 0, 0, 0, 0},
                                        {3.0 / 2.0, -1.0 / 2.0, 0, 0, 0},
                                        {23.0 / 12.0, -16.0 / 12.0, 5.0 / 12.0, 0, 0},
                                        {55.0 / 24.0, -59.0 / 24.0, 37.0 / 24.0, -9.0 / 24.0, 0},
                                        {1901.0 / 720.0, -2774.0 / 720.0, 2616.0 / 720.0, -1274.0 / 720.0, 251.0 / 720.0}};
        if (p->order > 5) return fail(set_err(EDUS_ERR_UNSUPPORTED, "Adams-Bashforth order > 5"));
        c->coef.assign(ab[p->order - 1], ab[p->order - 1] + p->order);
    }

    const size_t se = state_elems(c);
    int rc = 0;
    if ((rc = dalloc(c, &c->P0, se)) || (rc = dalloc(c, &c->ACC, se)) || (rc = dalloc(c, &c->YA, se)) || (rc = dalloc(c, &c->YB, se)))
        return fail(rc);
    if (c->stepper != EDUS_STEPPER_RK4) {
        if ((rc = dalloc(c, &c->HIST, se * c->order))) return fail(rc);
        if (hipMemsetAsync(c->HIST, 0, se * c->order * sizeof(double), c->stream) != hipSuccess)
            return fail(set_err(EDUS_ERR_HIP, "memset history"));
    }
    if (!c->soa) {
        std::vector<int> ex(c->np);
        for (int i = 0; i < nb; i++) ex[i] = i | (i << 8);
        for (int i = 0; i < nb; i++)
            for (int j = i + 1; j < nb; j++) {
                const int pp = pair_index(i, j, nb);
                ex[nb + 2 * pp] = i | (j << 8);
                ex[nb + 2 * pp + 1] = i | (j << 8) | (1 << 16);
            }
        if ((rc = dalloc(c, &c->eidx, ex.size())) ||
            hipMemcpyAsync(c->eidx, ex.data(), ex.size() * sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(rc ? rc : set_err(EDUS_ERR_HIP, "element table upload"));
    }
    // Operators in the state layout, packed on the device from chunks of the
    // host tables (sk, sek: k and element strides of the state layout).
    const size_t sk = c->soa ? 1 : (size_t)c->np, sek = c->soa ? (size_t)c->nkp : 1;
    const int stage_k = (int)std::max<long long>(1, std::min<long long>(p->nk, (1LL << 25) / ((long long)nb * nb * 6)));
    std::unique_ptr<double, DeviceFree> stage;
    {
        double *q = nullptr;
        if (hipMalloc(&q, (size_t)stage_k * nb * nb * 6 * sizeof(double)) != hipSuccess)
            return fail(set_err(EDUS_ERR_NOMEM, "staging buffer"));
        stage.reset(q);
    }
    double *st = stage.get();
    if ((rc = dalloc(c, &c->H, se)) || (rc = dalloc(c, &c->D, 3 * se))) return fail(rc);
    if (c->soa && (hipMemsetAsync(c->H, 0, se * sizeof(double), c->stream) != hipSuccess ||
                   hipMemsetAsync(c->D, 0, 3 * se * sizeof(double), c->stream) != hipSuccess))
        return fail(set_err(EDUS_ERR_HIP, "memset operators"));  // padding columns of the SoA layout
    if ((rc = upload_packed(c, p->H, 1, st, stage_k, [&](int k0, int kc) { launch_pack(c, st, kc, 1, 0, sk, sek, c->H + k0 * sk); })))
        return fail(rc);
    {
        // D: three packed components [x][state], and its core <-> non-core block.
        const int nblk = c->nch * (nb - c->nch);
        const size_t ndt = (size_t)p->nk * 3 * std::max(nblk, 1) * 2;
        if ((rc = dalloc(c, &c->DT, ndt))) return fail(rc);
        if (hipMemsetAsync(c->DT, 0, ndt * sizeof(double), c->stream) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "memset DT"));
        rc = upload_packed(c, p->D, 3, st, stage_k, [&](int k0, int kc) {
            for (int x = 0; x < 3; x++) launch_pack(c, st, kc, 3, x, sk, sek, c->D + x * se + k0 * sk);
            if (nblk > 0) {
                const long long n = (long long)kc * 3 * nblk;
                hipLaunchKernelGGL(k_pack_dt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (const double *)st, kc, nb, c->nch,
                                   c->DT + (size_t)k0 * 3 * nblk * 2);
            }
        });
        if (rc) return fail(rc);
        // Q_x = H D_x^+ for Current2, built on the device from the packed operators.
        const size_t nq = (size_t)p->nk * 3 * nb * nb;
        if ((rc = dalloc(c, &c->Q, nq * 2))) return fail(rc);
        hipLaunchKernelGGL(k_build_q, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, c->stream, (const double *)c->H,
                           (const double *)c->D, p->nk, nb, sk, sek, se, c->dherm, c->Q);
        if (hipGetLastError() != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "k_build_q launch"));
    }
    {
        const size_t nu = (size_t)p->nk * nb * nb * 2;
        if ((rc = dalloc(c, &c->U, nu)) || (rc = dalloc(c, &c->T, (size_t)nb * nb))) return fail(rc);
        if (hipMemcpyAsync(c->U, p->U, nu * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->T, p->T, (size_t)nb * nb * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return fail(set_err(EDUS_ERR_HIP, "upload U, T"));
    }
    if (c->precision != EDUS_PRECISION_FP64) {
        const size_t nh = se, nd = 3 * se, nu = (size_t)p->nk * nb * nb * 2;
        if ((rc = dalloc(c, &c->Hf, nh)) || (rc = dalloc(c, &c->Df, nd)) || (rc = dalloc(c, &c->Uf, nu))) return fail(rc);
        const double *src[3] = {c->H, c->D, c->U};
        float *dst[3] = {c->Hf, c->Df, c->Uf};
        const size_t n[3] = {nh, nd, nu};
        for (int q = 0; q < 3; q++)
            hipLaunchKernelGGL(k_to_float, dim3((unsigned)std::min<size_t>((n[q] + 255) / 256, 65536)), dim3(256), 0, c->stream, src[q], dst[q], n[q]);
        if (hipGetLastError() != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "k_to_float launch"));
    }
    if (p->gradH) {
        // packed Hermitian gradient of H, [k][3][NP] (always k-major)
        const size_t np3 = (size_t)3 * c->np;
        if ((rc = dalloc(c, &c->GE, (size_t)p->nk * np3))) return fail(rc);
        rc = upload_packed(c, p->gradH, 3, st, stage_k, [&](int k0, int kc) {
            for (int x = 0; x < 3; x++) launch_pack(c, st, kc, 3, x, np3, 1, c->GE + (size_t)k0 * np3 + (size_t)x * c->np);
        });
        if (rc) return fail(rc);
    }
    {
        // [n_nbr][nkp] table, then the stencil-pair pattern (2 kMaxNbr ints)
        const size_t ntab = (size_t)std::max(p->n_nbr, 1) * c->nkp;
        std::vector<int> nbr(ntab + 2 * kMaxNbr, -1);
        for (int k = 0; k < p->nk; k++)
            for (int n = 0; n < p->n_nbr; n++) nbr[(size_t)n * c->nkp + k] = p->nbr[(size_t)k * p->n_nbr + n];
        if (!c->soa && nb > 8 && p->n_nbr > 0 && p->nk % 2 == 0 && c->scol % 2 == 0 &&
            stencil_pairs(p->nbr, p->nk, p->n_nbr, nbr.data() + ntab, c->pself, c->npnew))
            c->pair = 1;
        if ((rc = dalloc(c, &c->nbr, nbr.size()))) return fail(rc);
        if (hipMemcpyAsync(c->nbr, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return fail(set_err(EDUS_ERR_HIP, "upload nbr"));
    }
    c->obs_blocks = std::max(1, std::min(kObsMaxBlocks, p->nk));
    if ((rc = dalloc(c, &c->part, (size_t)c->obs_blocks * (nb + 17))) || (rc = dalloc(c, &c->obs, nb + 17))) return fail(rc);
    {
        if (c->soa && hipMemsetAsync(c->P0, 0, se * sizeof(double), c->stream) != hipSuccess)
            return fail(set_err(EDUS_ERR_HIP, "memset state"));
        if ((rc = upload_packed(c, p->rho0, 1, st, stage_k, [&](int k0, int kc) { launch_pack(c, st, kc, 1, 0, sk, sek, c->P0 + k0 * sk); })))
            return fail(rc);
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "table packing"));
        stage.reset();
        if (c->soa) {  // padding columns stay zero in every state buffer
            if (hipMemsetAsync(c->ACC, 0, se * sizeof(double), c->stream) != hipSuccess ||
                hipMemsetAsync(c->YA, 0, se * sizeof(double), c->stream) != hipSuccess ||
                hipMemsetAsync(c->YB, 0, se * sizeof(double), c->stream) != hipSuccess)
                return fail(set_err(EDUS_ERR_HIP, "memset state"));
        }
    }
    if (p->coulomb) {
        const edus_coulomb *q = p->coulomb;
        if (!c->soa) return fail(set_err(EDUS_ERR_UNSUPPORTED, "the Coulomb term is implemented for nb <= 4"));
        if (p->nk_dims[0] != 1 || p->nk_dims[1] * p->nk_dims[2] != p->nk)
            return fail(set_err(EDUS_ERR_UNSUPPORTED, "the Coulomb term needs a 1 x n1 x n2 grid (sample in yz)"));
        if (q->ncut < 1 || !q->A || !q->B || !q->C || !q->D || !q->screen || !q->v_hartree)
            return fail(set_err(EDUS_ERR_INVALID, "incomplete Coulomb tables"));
        if (q->ncut > 32) return fail(set_err(EDUS_ERR_UNSUPPORTED, "Ncut above 32"));
        c->coul = 1; c->N1 = p->nk_dims[1]; c->N2 = p->nk_dims[2]; c->ncut = q->ncut; c->nu = nb * (nb + 1) / 2;
        const int nmm = c->ncut * c->ncut, nb2 = nb * nb;
        // trig tables cos/sin(2 pi m k) with k = -0.5 + i/N (integrationWeight.h:139-175)
        std::vector<double> cx((size_t)c->ncut * c->N1), sx(cx.size()), cy((size_t)c->ncut * c->N2), sy(cy.size());
        for (int m = 0; m < c->ncut; m++) {
            for (int i = 0; i < c->N1; i++) {
                const double kk = (1. / c->N1 != 1.) ? -0.5 + (1. / c->N1) * i : 0.;
                const double arg = 2 * M_PI * m * (kk);
                cx[(size_t)m * c->N1 + i] = cos(arg);
                sx[(size_t)m * c->N1 + i] = sin(arg);
            }
            for (int i = 0; i < c->N2; i++) {
                const double kk = (1. / c->N2 != 1.) ? -0.5 + (1. / c->N2) * i : 0.;
                const double arg = 2 * M_PI * m * (kk);
                cy[(size_t)m * c->N2 + i] = cos(arg);
                sy[(size_t)m * c->N2 + i] = sin(arg);
            }
        }
        auto up = [&](double **dst, const std::vector<double> &v) -> int {
            int r = dalloc(c, dst, v.size());
            return r ? r : upload(c, *dst, v);
        };
        auto upc = [&](double **dst, const double *src, size_t n) -> int { return up(dst, std::vector<double>(src, src + n)); };
        if ((rc = up(&c->cxm, cx)) || (rc = up(&c->sxm, sx)) || (rc = up(&c->cyn, cy)) || (rc = up(&c->syn, sy))) return fail(rc);
        if ((rc = upc(&c->cA, q->A, (size_t)nmm * nb2 * 2)) || (rc = upc(&c->cB, q->B, (size_t)nmm * nb2 * 2)) ||
            (rc = upc(&c->cC, q->C, (size_t)nmm * nb2 * 2)) || (rc = upc(&c->cD, q->D, (size_t)nmm * nb2 * 2)) ||
            (rc = upc(&c->cscreen, q->screen, (size_t)nb2 * 2)) || (rc = upc(&c->cvh, q->v_hartree, (size_t)nb2 * 2)))
            return fail(rc);
        if ((rc = dalloc(c, &c->cS, (size_t)2 * c->N1 * c->ncut * c->np)) || (rc = dalloc(c, &c->cF, (size_t)4 * nmm * c->np)) ||
            (rc = dalloc(c, &c->cX, (size_t)4 * nmm * nb2 * 2)) || (rc = dalloc(c, &c->cT, (size_t)2 * c->N2 * c->ncut * c->nu * 2)) ||
            (rc = dalloc(c, &c->cEH, nb)) || (rc = dalloc(c, &c->cPW0, se)) || (rc = dalloc(c, &c->cHr, se)) ||
            (rc = dalloc(c, &c->cUexc, (size_t)p->nk * nb2 * 2)) || (rc = dalloc(c, &c->cGE, (size_t)p->nk * 3 * c->np)) ||
            (rc = dalloc(c, &c->cQ, (size_t)p->nk * 3 * nb2 * 2)) || (rc = dalloc(c, &c->cEk, (size_t)p->nk * nb)) ||
            (rc = dalloc(c, &c->cXw, (size_t)2 * c->nu * c->nkp)))
            return fail(rc);
        if ((rc = up(&c->cX0, std::vector<double>((size_t)4 * nmm * nb2 * 2, 0.0))) || (rc = up(&c->cEH0, std::vector<double>(nb, 0.0))) ||
            (rc = upc(&c->dwb, p->nbr_wb, (size_t)std::max(p->n_nbr, 1) * 3)))
            return fail(rc);
        // P_Wannier_0 = P0, renormalised H starts as H (main_MPI.cpp:214, 330)
        if (hipMemcpyAsync(c->cPW0, c->P0, se * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->cHr, c->H, se * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return fail(set_err(EDUS_ERR_HIP, "copy P0/H"));
        // equilibrium: X0 and the Hartree shift of P0 (main_MPI.cpp:276-321); in the
        // diagonal basis of the filled bands, occupation 1 (main_MPI.cpp:291-303)
        if (q->diag_basis) {
            c->coul_diag = 1;
            if ((rc = dalloc(c, &c->cPD, se))) return fail(rc);
            std::vector<double> peq(se, 0.0);
            for (int i = 0; i < c->ncc; i++)
                for (int k = 0; k < p->nk; k++) peq[(size_t)i * c->nkp + k] = 1.0;
            if ((rc = upload(c, c->cPD, peq))) return fail(rc);
            if ((rc = launch_coulomb(c, c->cPD))) return fail(rc);
            if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "diagonal-basis equilibrium"));
        } else if ((rc = launch_coulomb(c, c->P0))) {
            return fail(rc);
        }
        if (hipMemcpyAsync(c->cX0, c->cX, (size_t)4 * nmm * nb2 * 2 * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->cEH0, c->cEH, nb * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return fail(set_err(EDUS_ERR_HIP, "copy equilibrium"));
        c->n_cond = 0.0;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "sync after upload"));
    *out = c;
    g_err.clear();
    return 0;
}

int edus_destroy(edus_ctx *c) {
    if (!c) return 0;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->gtab) (void)hipFree(c->gtab);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->tab) (void)hipFree(c->tab);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

static int finish_step(edus_ctx *c, int launches) {
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = launches;
    return 0;
}

// Arguments of RK4 stage st (0..3) of size dt reading the field row `row`
// (RungeKutta.cpp:11-150): y_n in P0, stages through YA/YB, k-sum in ACC.
static StageArgs rk4_stage_args(const edus_ctx *c, int st, const double *row, double dt) {
    StageArgs a;
    std::memset(&a, 0, sizeof a);
    a.dt = dt;
    a.tab = row;
    a.dmask = c->dmask;
    a.P0 = c->P0; a.ACC = c->ACC;
    const double dt6 = dt / 6.0, dt3 = dt / 3.0, dt2 = dt / 2.0;
    switch (st) {
        case 0: a.Y = c->P0; a.OUT = c->YA; a.ca = dt6; a.cb = dt2; break;
        case 1: a.Y = c->YA; a.OUT = c->YB; a.ca = dt3; a.cb = dt2; break;
        case 2: a.Y = c->YB; a.OUT = c->YA; a.ca = dt3; a.cb = dt; break;
        default: a.Y = c->YA; a.OUT = c->P0; a.ca = dt6; break;
    }
    return a;
}

// One RK4 step of size dt from the four table rows at `tab`.
static int rk4_step(edus_ctx *c, const double *tab, int rs, double dt) {
    int rc;
    if ((rc = launch_stage<RK1>(c, rk4_stage_args(c, 0, tab, dt)))) return rc;
    if ((rc = launch_stage<RK2>(c, rk4_stage_args(c, 1, tab + rs, dt)))) return rc;
    if ((rc = launch_stage<RK3>(c, rk4_stage_args(c, 2, tab + 2 * rs, dt)))) return rc;
    return launch_stage<RK4S>(c, rk4_stage_args(c, 3, tab + 3 * rs, dt));
}

// Fields [nsteps][3][6] -> table rows [nsteps][4] (stage 3 reuses the midpoint field).
static void rk4_rows(const edus_ctx *c, int32_t nsteps, const double *fields, const uint8_t *grad_on, std::vector<double> &rows) {
    const int rs = 7 + c->n_nbr;
    rows.assign((size_t)nsteps * 4 * rs, 0.0);
    static const int evalof[4] = {0, 1, 1, 2};
    for (int s = 0; s < nsteps; s++)
        for (int st = 0; st < 4; st++)
            table_row(c, fields + ((size_t)s * 3 + evalof[st]) * 6, grad_on[s], &rows[((size_t)s * 4 + st) * rs]);
}

int edus_step_rk4(edus_ctx *c, int32_t nsteps, const double *fields, const uint8_t *grad_on) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->stepper != EDUS_STEPPER_RK4) return set_err(EDUS_ERR_INVALID, "context was created for another stepper");
    if (nsteps < 0 || (nsteps > 0 && (!fields || !grad_on))) return set_err(EDUS_ERR_INVALID, "bad step arguments");
    const int rs = 7 + c->n_nbr;
    std::vector<double> rows;
    rk4_rows(c, nsteps, fields, grad_on, rows);
    if ((rc = upload_table(c, rows))) return rc;
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; s++)
        if ((rc = rk4_step(c, c->tab + (size_t)s * 4 * rs, rs, c->dt))) return rc;
    c->nstep += nsteps;
    return finish_step(c, 4 * nsteps);
}

static void drop_graph(edus_ctx *c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

int edus_set_field_table(edus_ctx *c, int32_t nsteps, const double *fields, const uint8_t *grad_on) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->stepper != EDUS_STEPPER_RK4) return set_err(EDUS_ERR_UNSUPPORTED, "the field-table path runs RK4 contexts");
    if (nsteps < 1 || !fields || !grad_on) return set_err(EDUS_ERR_INVALID, "bad field table");
    std::vector<double> rows;
    rk4_rows(c, nsteps, fields, grad_on, rows);
    const size_t rs = 7 + (size_t)c->n_nbr;
    int mask = 0;
    for (size_t r = 0; r + rs <= rows.size(); r += rs)
        for (int x = 0; x < 3; x++)
            if (rows[r + x] != 0.0 || rows[r + 3 + x] != 0.0) mask |= 1 << x;
    if (rows.size() > c->gtab_cap || mask != c->gdmask) drop_graph(c);  // the graph holds the table address and the mask
    if (rows.size() > c->gtab_cap) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->gtab) HIP_TRY(hipFree(c->gtab));
        c->gtab = nullptr;
        c->gtab_cap = 0;
        HIP_TRY(hipMalloc(&c->gtab, rows.size() * sizeof(double)));
        c->gtab_cap = rows.size();
    }
    if (!c->gctr) {
        int *q = nullptr;
        HIP_TRY(hipMalloc(&q, sizeof(int)));
        c->allocs.push_back(q);
        c->gctr = q;
    }
    c->gdmask = mask;
    HIP_TRY(hipMemcpyAsync(c->gtab, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_counter_set, dim3(1), dim3(1), 0, c->stream, c->gctr, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));  // the host rows are released on return
    c->gtab_steps = nsteps;
    c->gpos = 0;
    return 0;
}

// One RK4 step whose stage rows come from the field table at the device step
// counter, captured once into a graph (four stage launches and the counter
// advance; Coulomb passes included) and replayed for every step.
static int build_step_graph(edus_ctx *c) {
    const int rs = 7 + c->n_nbr;
    const int saved = c->dmask;
    c->dmask = c->gdmask;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int st = 0; st < 4 && rc == 0; st++) {
        StageArgs a = rk4_stage_args(c, st, c->gtab + (size_t)st * rs, c->dt);
        a.ctr = c->gctr;
        a.ctr_stride = 4LL * rs;
        rc = st == 0 ? launch_stage<RK1>(c, a) : st == 1 ? launch_stage<RK2>(c, a) : st == 2 ? launch_stage<RK3>(c, a) : launch_stage<RK4S>(c, a);
    }
    if (rc == 0) hipLaunchKernelGGL(k_counter_step, dim3(1), dim3(1), 0, c->stream, c->gctr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->dmask = saved;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_err(EDUS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    c->graph = g;
    HIP_TRY(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    return 0;
}

int edus_step_rk4_graph(edus_ctx *c, int32_t nsteps) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->stepper != EDUS_STEPPER_RK4) return set_err(EDUS_ERR_UNSUPPORTED, "the field-table path runs RK4 contexts");
    if (!c->gtab) return set_err(EDUS_ERR_INVALID, "no field table: call edus_set_field_table first");
    if (nsteps < 0 || c->gpos + nsteps > c->gtab_steps) return set_err(EDUS_ERR_INVALID, "steps beyond the end of the field table");
    if (!c->gexec && (rc = build_step_graph(c))) return rc;
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; s++) HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
    c->gpos += nsteps;
    c->nstep += nsteps;
    return finish_step(c, 4 * nsteps);
}

// One AB or Taylor step of size dt from the table row at `tab`
// (Linear_Multistep_ODE_solvers.cpp:1-93, Taylor_DE_solver_deriv_Polynom.cpp:1-177).
static int multistep_one(edus_ctx *c, const double *tab, bool taylor, double dt, const double *tinv) {
    const int o = c->order;
    StageArgs a;
    std::memset(&a, 0, sizeof a);
    a.dt = dt;
    a.order = o;
    a.tab = tab;
    a.HIST = c->HIST;
    // The derivative of step n lives in slot (-n-1) mod order (ODE_solver.h:66-75
    // shifts the ring before each step); slots[s] holds f_{n-s}.
    const long long n = c->nstep;
    for (int q = 0; q < o; q++) a.slots[q] = (int)((((q - n - 1) % o) + o) % o);
    a.slot_out = a.slots[0];
    if (taylor) {
        std::copy(tinv, tinv + (size_t)(o - 1) * (o - 1), a.tinv);
        for (int i = 0; i < o; i++) a.dtp[i] = std::pow(dt, i + 1);
    } else {
        std::copy(c->coef.begin(), c->coef.end(), a.coef);
    }
    double *in = c->cur ? c->YA : c->P0, *out = c->cur ? c->P0 : c->YA;
    a.Y = in; a.P0 = in; a.OUT = out;
    const int rc = taylor ? launch_stage<TAY>(c, a) : launch_stage<ABM>(c, a);
    if (rc) return rc;
    c->cur ^= 1;
    c->nstep++;
    return 0;
}

static int step_multistep(edus_ctx *c, int32_t nsteps, const double *fields, const uint8_t *grad_on, bool taylor) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->stepper != (taylor ? EDUS_STEPPER_TAYLOR : EDUS_STEPPER_AB)) return set_err(EDUS_ERR_INVALID, "context was created for another stepper");
    if (nsteps < 0 || (nsteps > 0 && (!fields || !grad_on))) return set_err(EDUS_ERR_INVALID, "bad step arguments");
    const int rs = 7 + c->n_nbr;
    std::vector<double> rows((size_t)nsteps * rs);
    for (int s = 0; s < nsteps; s++) table_row(c, fields + (size_t)s * 6, grad_on[s], &rows[(size_t)s * rs]);
    if ((rc = upload_table(c, rows))) return rc;
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; s++)
        if ((rc = multistep_one(c, c->tab + (size_t)s * rs, taylor, c->dt, c->tinv.data()))) return rc;
    return finish_step(c, nsteps);
}

int edus_step_ab(edus_ctx *c, int32_t nsteps, const double *fields, const uint8_t *grad_on) {
    return step_multistep(c, nsteps, fields, grad_on, false);
}

int edus_step_taylor(edus_ctx *c, int32_t nsteps, const double *fields, const uint8_t *grad_on) {
    return step_multistep(c, nsteps, fields, grad_on, true);
}

static const double *current_state(const edus_ctx *c) {
    if (c->stepper == EDUS_STEPPER_RK4) return c->P0;
    return c->cur ? c->YA : c->P0;
}

int edus_step_dyn(edus_ctx *c, double dt, const double *fields, uint8_t grad_on, const double *taylor_tau_inv, double cond,
                  double *max_dp) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!(dt > 0) || !fields || !max_dp) return set_err(EDUS_ERR_INVALID, "bad adaptive step arguments");
    const bool taylor = c->stepper == EDUS_STEPPER_TAYLOR;
    if (!taylor && c->stepper != EDUS_STEPPER_RK4)
        return set_err(EDUS_ERR_UNSUPPORTED, "adaptive steps run RK4 or Taylor contexts");
    if (taylor && !taylor_tau_inv) return set_err(EDUS_ERR_INVALID, "Taylor adaptive step needs tau^-1");
    const int nblk = (c->nk + 255) / 256;
    if (!c->DG) {
        if ((rc = dalloc(c, &c->DG, (size_t)c->nb * c->nk)) || (rc = dalloc(c, &c->dpart, (size_t)nblk))) return rc;
    }
    const int rs = 7 + c->n_nbr;
    std::vector<double> rows;
    if (taylor) {
        rows.resize(rs);
        table_row(c, fields, grad_on, rows.data());
    } else {
        rk4_rows(c, 1, fields, &grad_on, rows);
    }
    if ((rc = upload_table(c, rows))) return rc;
    const size_t sk = c->soa ? 1 : (size_t)c->np, se = c->soa ? (size_t)c->nkp : 1;
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_diag_save, dim3(nblk), dim3(256), 0, c->stream, current_state(c), c->DG, c->nk, c->nb, sk, se);
    HIP_TRY(hipGetLastError());
    if (taylor) rc = multistep_one(c, c->tab, true, dt, taylor_tau_inv);
    else {
        rc = rk4_step(c, c->tab, rs, dt);
        c->nstep++;
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_diag_dp, dim3(nblk), dim3(256), 0, c->stream, current_state(c), (const double *)c->DG, c->nk, c->nb, sk, se,
                       cond, c->dpart);
    HIP_TRY(hipGetLastError());
    std::vector<double> part(nblk);
    HIP_TRY(hipMemcpyAsync(part.data(), c->dpart, nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if ((rc = finish_step(c, taylor ? 1 : 4))) return rc;
    double m = 0.0;
    for (double v : part) m = std::max(m, v);
    *max_dp = m;
    return 0;
}

int edus_eigen_element(edus_ctx *c, int32_t ic, int32_t jc, double *out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!out || ic < 0 || jc < 0 || ic >= c->nb || jc >= c->nb) return set_err(EDUS_ERR_INVALID, "bad snapshot element");
    if (!c->snap && (rc = dalloc(c, &c->snap, (size_t)2 * c->nk))) return rc;
    const size_t sk = c->soa ? 1 : (size_t)c->np, se = c->soa ? (size_t)c->nkp : 1;
    hipLaunchKernelGGL(k_eigen_element, dim3((c->nk + 255) / 256), dim3(256), 0, c->stream, current_state(c), (const double *)c->U, c->nk,
                       c->nb, sk, se, ic, jc, c->snap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, c->snap, (size_t)2 * c->nk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int edus_renorm_energies(edus_ctx *c, double *out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!out) return set_err(EDUS_ERR_INVALID, "null output");
    if (!c->coul) return set_err(EDUS_ERR_INVALID, "no Coulomb term in this context");
    if (!c->cEk_valid) return set_err(EDUS_ERR_INVALID, "call edus_observables first");
    HIP_TRY(hipMemcpyAsync(out, c->cEk, (size_t)c->nk * c->nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int edus_observables(edus_ctx *c, double *pop, double *j1, double *j2, double *tabs, double *pminmax) {
    int rc = check_ctx(c);
    if (rc) return rc;
    ObsArgs o;
    o.P = current_state(c); o.U = c->U; o.GE = c->GE; o.Q = c->Q; o.DT = c->DT; o.part = c->part;
    if (c->coul) {
        const int blocks = (c->nk + 255) / 256;
        switch (c->nb) {
            case 1: hipLaunchKernelGGL(k_coul_obs<1>, dim3(blocks), dim3(256), 0, c->stream, (const double *)c->cHr, (const double *)c->D, (const int *)c->nbr, (const double *)c->dwb, c->n_nbr, c->nk, c->nkp, c->dherm, c->cUexc, c->cGE, c->cQ, c->cEk); break;
            case 2: hipLaunchKernelGGL(k_coul_obs<2>, dim3(blocks), dim3(256), 0, c->stream, (const double *)c->cHr, (const double *)c->D, (const int *)c->nbr, (const double *)c->dwb, c->n_nbr, c->nk, c->nkp, c->dherm, c->cUexc, c->cGE, c->cQ, c->cEk); break;
            case 3: hipLaunchKernelGGL(k_coul_obs<3>, dim3(blocks), dim3(256), 0, c->stream, (const double *)c->cHr, (const double *)c->D, (const int *)c->nbr, (const double *)c->dwb, c->n_nbr, c->nk, c->nkp, c->dherm, c->cUexc, c->cGE, c->cQ, c->cEk); break;
            default: hipLaunchKernelGGL(k_coul_obs<4>, dim3(blocks), dim3(256), 0, c->stream, (const double *)c->cHr, (const double *)c->D, (const int *)c->nbr, (const double *)c->dwb, c->n_nbr, c->nk, c->nkp, c->dherm, c->cUexc, c->cGE, c->cQ, c->cEk); break;
        }
        HIP_TRY(hipGetLastError());
        c->cEk_valid = true;
        o.U = c->cUexc;
        o.Q = c->cQ;
        if (c->GE) o.GE = c->cGE;
    }
    o.nk = c->nk; o.nb = c->nb; o.nch = c->nch; o.ncc = c->ncc;
    o.sk = c->soa ? 1 : c->np;
    o.se = c->soa ? c->nkp : 1;
    const size_t obs_lds = (size_t)c->nb * (2 * c->nb + 1) * sizeof(cd) + (4 * 17 + 32) * sizeof(double);
    if (c->nb > 8)
        hipLaunchKernelGGL(k_observables<256>, dim3(c->obs_blocks), dim3(256), obs_lds, c->stream, o);
    else
        hipLaunchKernelGGL(k_observables<64>, dim3(c->obs_blocks), dim3(64), obs_lds, c->stream, o);
    HIP_TRY(hipGetLastError());
    const int nv = c->nb + 15;
    hipLaunchKernelGGL(k_obs_final, dim3(nv + 2), dim3(256), 0, c->stream, (const double *)c->part, c->obs_blocks, nv, c->obs);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(nv + 2);
    HIP_TRY(hipMemcpyAsync(h.data(), c->obs, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int nb = c->nb;
    const double inv = 1. / c->nktot;
    if (pop)
        for (int i = 0; i < nb; i++) pop[i] = h[i] * inv;
    if (j1)
        for (int x = 0; x < 3; x++) { j1[2 * x] = h[nb + x] * c->j_norm; j1[2 * x + 1] = 0.0; }
    if (j2)
        for (int x = 0; x < 3; x++) { j2[2 * x] = h[nb + 3 + 2 * x] * c->j_norm; j2[2 * x + 1] = h[nb + 4 + 2 * x] * c->j_norm; }
    if (tabs)
        for (int x = 0; x < 3; x++) { tabs[2 * x] = h[nb + 9 + 2 * x] * c->tabs_norm; tabs[2 * x + 1] = h[nb + 10 + 2 * x] * c->tabs_norm; }
    if (pminmax) { pminmax[0] = h[nv]; pminmax[1] = h[nv + 1]; }
    if (c->coul) {  // n_cond: largest printed population (Observables_MPI.h:413-426)
        double nc = 0;
        for (int i = 0; i < nb; i++) {
            const double v = (i < c->ncc) ? 1.0 - h[i] * inv : h[i] * inv;
            if (v > nc) nc = v;
        }
        c->n_cond = nc;
    }
    return 0;
}

int edus_rhs(edus_ctx *c, const double *field6, uint8_t grad_on, double *pv) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!field6 || !pv) return set_err(EDUS_ERR_INVALID, "null argument");
    std::vector<double> row(7 + c->n_nbr);
    table_row(c, field6, grad_on, row.data());
    if ((rc = upload_table(c, row))) return rc;
    StageArgs a;
    std::memset(&a, 0, sizeof a);
    a.tab = c->tab;
    a.Y = current_state(c);
    a.OUT = c->YB;
    if ((rc = launch_stage<RHSO>(c, a))) return rc;
    std::vector<double> st(state_elems(c));
    HIP_TRY(hipMemcpyAsync(st.data(), c->YB, st.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    unpack_state(c, st, pv);
    return 0;
}

int edus_get_state(edus_ctx *c, double *rho) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!rho) return set_err(EDUS_ERR_INVALID, "null output");
    std::vector<double> st(state_elems(c));
    HIP_TRY(hipMemcpyAsync(st.data(), current_state(c), st.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    unpack_state(c, st, rho);
    return 0;
}

int edus_set_state(edus_ctx *c, const double *rho) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!rho) return set_err(EDUS_ERR_INVALID, "null input");
    std::vector<double> st;
    pack_state(c, rho, st);
    HIP_TRY(hipMemcpyAsync(const_cast<double *>(current_state(c)), st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Schedule of the stage kernels picked by timing the context's own right-hand
// side: every candidate gives the same sums for every k (only the order of the
// workgroups or their size changes), so the choice cannot change a result.
int edus_autotune(edus_ctx *c, const double *field6, uint8_t grad_on, int32_t reps, int32_t *chosen) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!field6 || reps < 1 || reps > 1000) return set_err(EDUS_ERR_INVALID, "bad autotune arguments");
    std::vector<double> row(7 + c->n_nbr);
    table_row(c, field6, grad_on, row.data());
    if ((rc = upload_table(c, row))) return rc;
    StageArgs a;
    std::memset(&a, 0, sizeof a);
    a.tab = c->tab;
    a.Y = current_state(c);
    a.OUT = c->YB;  // scratch between steps: the state is not touched
    std::vector<int> cand;
    if (c->soa) {
        cand = {64, 128, 256};
    } else {
        cand.push_back(0);
        if (c->scol > 0)
            for (int w : {16, 32, 64, 128, 256})
                if (w < c->scol) cand.push_back(w);
    }
    int &knob = c->soa ? c->tpb : c->sw;
    const int before = knob;
    int best = cand[0];
    float best_ms = 0.f;
    for (size_t i = 0; i < cand.size(); i++) {
        knob = cand[i];
        if ((rc = launch_stage<RHSO>(c, a))) { knob = before; return rc; }  // untimed
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < reps; r++)
            if ((rc = launch_stage<RHSO>(c, a))) { knob = before; return rc; }
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
        HIP_TRY(hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (i == 0 || ms < best_ms) { best_ms = ms; best = cand[i]; }
    }
    knob = best;
    if (best != before) drop_graph(c);  // a captured step holds the old launch shape
    if (chosen) *chosen = best;
    return 0;
}

int edus_last_step_time(edus_ctx *c, double *ms, int32_t *nlaunch) {
    if (!c) return set_err(EDUS_ERR_INVALID, "null context");
    if (ms) *ms = c->last_ms;
    if (nlaunch) *nlaunch = c->last_launches;
    return 0;
}

int edus_launch_schedule(edus_ctx *c, int32_t *out) {
    if (!c || !out) return set_err(EDUS_ERR_INVALID, "null argument");
    out[0] = c->soa ? 0 : (c->nb <= 8 ? 8 : c->nb <= 16 ? 16 : 32);
    out[1] = c->sw;
    out[2] = c->pair;
    out[3] = c->pair ? c->npnew : 0;
    out[4] = c->soa ? (c->tpb ? c->tpb : (c->nk < 256 * 256 ? 64 : 256)) : 0;
    return 0;
}

// ---------------------------------------------------------------- batched instances
struct edus_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<edus_ctx *> inst;
    struct Cls {
        bool soa = true;
        int nb_small = 0, nbt = 0, precision = EDUS_PRECISION_FP64, diss = 0, pair = 0;
        std::vector<int> members;
        int2 *blk = nullptr;
        int nblocks = 0;
    };
    std::vector<Cls> cls;
    BatchEntry *entries = nullptr;
    int *masks = nullptr;   // [n] dipole components the call's fields need, per member
    double *tab = nullptr;  // [step][stage][R_tot]; instance i's row at roff[i]
    size_t tab_steps = 0;
    int R_tot = 0;
    std::vector<int> roff;
    std::vector<void *> allocs;
    double last_ms = 0.0;
    int last_launches = 0;
};

extern "C++" {
namespace {

int batch_write_entries(edus_batch *b) {
    std::vector<BatchEntry> e(b->inst.size());
    for (size_t i = 0; i < b->inst.size(); i++) {
        const edus_ctx *c = b->inst[i];
        e[i].c = const_args(c);
        for (int st = 0; st < 4; st++) {
            e[i].a[st] = rk4_stage_args(c, st, b->tab + (size_t)st * b->R_tot + b->roff[i], c->dt);
            e[i].a[st].dmask = 7;  // replaced per call by the member's mask (b->masks)
        }
    }
    HIP_TRY(hipMemcpyAsync(b->entries, e.data(), e.size() * sizeof(BatchEntry), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int batch_table_capacity(edus_batch *b, size_t nsteps) {
    if (nsteps <= b->tab_steps) return 0;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->tab) HIP_TRY(hipFree(b->tab));
    b->tab = nullptr;
    const size_t cap = std::max<size_t>(nsteps, 16);
    HIP_TRY(hipMalloc(&b->tab, cap * 4 * (size_t)b->R_tot * sizeof(double)));
    b->tab_steps = cap;
    return batch_write_entries(b);
}

template <int MODE>
int batch_launch(edus_batch *b, const edus_batch::Cls &k, size_t tab_step) {
    const dim3 grid(k.nblocks);
    if (k.soa) {
        auto go = [&](auto dtag) {
            constexpr bool DISS = decltype(dtag)::value;
            switch (k.nb_small) {
                case 1: hipLaunchKernelGGL((k_stage_small_batch<1, MODE, DISS>), grid, dim3(kSmallBatchThreads), 0, b->stream, b->entries, k.blk, b->masks, tab_step); break;
                case 2: hipLaunchKernelGGL((k_stage_small_batch<2, MODE, DISS>), grid, dim3(kSmallBatchThreads), 0, b->stream, b->entries, k.blk, b->masks, tab_step); break;
                case 3: hipLaunchKernelGGL((k_stage_small_batch<3, MODE, DISS>), grid, dim3(kSmallBatchThreads), 0, b->stream, b->entries, k.blk, b->masks, tab_step); break;
                default: hipLaunchKernelGGL((k_stage_small_batch<4, MODE, DISS>), grid, dim3(kSmallBatchThreads), 0, b->stream, b->entries, k.blk, b->masks, tab_step); break;
            }
        };
        if (k.diss) go(std::true_type{});
        else go(std::false_type{});
    } else {
        auto go = [&](auto tag, auto ptag) {
            constexpr int NBT = decltype(tag)::value, KPB = WaveGeom<NBT>::KPB;
            constexpr bool PAIR = decltype(ptag)::value;
            const dim3 block(WaveGeom<NBT>::TPK * KPB);
            const size_t lds_f = (size_t)KPB * wave_lds_doubles<NBT, float>(k.diss) * 8;
            const size_t lds_d = (size_t)KPB * wave_lds_doubles<NBT, double>(k.diss) * 8;
            if (k.precision == EDUS_PRECISION_MIXED)
                hipLaunchKernelGGL((k_stage_wave_batch<NBT, MODE, float, PAIR>), grid, block, lds_f, b->stream, b->entries, k.blk, b->masks, tab_step);
            else
                hipLaunchKernelGGL((k_stage_wave_batch<NBT, MODE, double, PAIR>), grid, block, lds_d, b->stream, b->entries, k.blk, b->masks, tab_step);
        };
        if (k.nbt == 8) go(std::integral_constant<int, 8>{}, std::false_type{});
        else if (k.nbt == 16 && k.pair) go(std::integral_constant<int, 16>{}, std::true_type{});
        else if (k.nbt == 16) go(std::integral_constant<int, 16>{}, std::false_type{});
        else if (k.pair) go(std::integral_constant<int, 32>{}, std::true_type{});
        else go(std::integral_constant<int, 32>{}, std::false_type{});
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace
}  // extern "C++"

int edus_batch_destroy(edus_batch *b) {
    if (!b) return 0;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (edus_ctx *c : b->inst) edus_destroy(c);
    for (auto &k : b->cls)
        if (k.blk) (void)hipFree(k.blk);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->tab) (void)hipFree(b->tab);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int edus_batch_create(const edus_problem *const *problems, int32_t n, edus_batch **out) {
    if (!problems || !out || n < 1) return set_err(EDUS_ERR_INVALID, "null argument or empty batch");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!problems[i]) return set_err(EDUS_ERR_INVALID, "null problem in batch");
        if (problems[i]->stepper != EDUS_STEPPER_RK4) return set_err(EDUS_ERR_UNSUPPORTED, "batches run RK4 problems");
        if (problems[i]->coulomb) return set_err(EDUS_ERR_UNSUPPORTED, "batches do not carry the Coulomb term");
        if (problems[i]->precision == EDUS_PRECISION_MIXED32 || problems[i]->precision == EDUS_PRECISION_BF16X3)
            return set_err(EDUS_ERR_UNSUPPORTED, "batches run FP64 or EDUS_PRECISION_MIXED members");
    }
    auto *b = new edus_batch();
    auto fail = [&](int rc) { edus_batch_destroy(b); return rc; };
    if (hipGetDevice(&b->dev) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "stream create failed"));
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) return fail(set_err(EDUS_ERR_HIP, "event create failed"));
    for (int i = 0; i < n; i++) {
        edus_ctx *c = nullptr;
        const int rc = edus_create(problems[i], &c);
        if (rc) return fail(rc);
        // every instance works on the batch's stream
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->stream = b->stream;
        c->own_stream = false;
        b->inst.push_back(c);
        b->roff.push_back(b->R_tot);
        b->R_tot += 7 + c->n_nbr;
    }
    // kernel classes: small kernels by band count, wave kernels by tile and
    // precision, each split by the presence of dephasing and by stencil pairing
    for (int i = 0; i < n; i++) {
        const edus_ctx *c = b->inst[i];
        const int nbt = c->soa ? 0 : (c->nb <= 8 ? 8 : c->nb <= 16 ? 16 : 32);
        edus_batch::Cls *hit = nullptr;
        for (auto &k : b->cls)
            if (k.soa == c->soa && k.nb_small == (c->soa ? c->nb : 0) && k.nbt == nbt && k.precision == c->precision &&
                k.diss == c->diss && k.pair == c->pair)
                hit = &k;
        if (!hit) {
            b->cls.emplace_back();
            hit = &b->cls.back();
            hit->soa = c->soa; hit->nb_small = c->soa ? c->nb : 0; hit->nbt = nbt; hit->precision = c->precision;
            hit->diss = c->diss; hit->pair = c->pair;
        }
        hit->members.push_back(i);
    }
    for (auto &k : b->cls) {
        std::vector<int2> map;
        for (int i : k.members) {
            const edus_ctx *c = b->inst[i];
            const int kpb = c->nb <= 8 ? WaveGeom<8>::KPB : c->nb <= 16 ? WaveGeom<16>::KPB : WaveGeom<32>::KPB;
            const int nblk = c->soa    ? (c->nk + kSmallBatchThreads - 1) / kSmallBatchThreads
                             : c->pair ? c->nk / 2
                                       : (c->nk + kpb - 1) / kpb;
            for (int q = 0; q < nblk; q++) map.push_back(make_int2(i, q));
        }
        k.nblocks = (int)map.size();
        if (hipMalloc(&k.blk, map.size() * sizeof(int2)) != hipSuccess ||
            hipMemcpy(k.blk, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(EDUS_ERR_NOMEM, "batch block table"));
    }
    void *ent = nullptr;
    if (hipMalloc(&ent, (size_t)n * sizeof(BatchEntry)) != hipSuccess) return fail(set_err(EDUS_ERR_NOMEM, "batch entries"));
    b->allocs.push_back(ent);
    b->entries = static_cast<BatchEntry *>(ent);
    void *msk = nullptr;
    if (hipMalloc(&msk, (size_t)n * sizeof(int)) != hipSuccess) return fail(set_err(EDUS_ERR_NOMEM, "batch masks"));
    b->allocs.push_back(msk);
    b->masks = static_cast<int *>(msk);
    int rc = batch_table_capacity(b, 16);
    if (rc) return fail(rc);
    *out = b;
    g_err.clear();
    return 0;
}

// ---------------------------------------------------------------- setup: batched eigensolve
// Rows a of U are conj(v_a), the eigenvectors of H in ascending eigenvalue order
// (Wannier.h:496-542).  rocSOLVER's batched Jacobi solver reads our row-major
// H as its column-major transpose conj(H) (its lower triangle is our upper
// one, the triangle the engine packs), whose eigenvectors conj(v_a) it
// returns as columns, i.e. as our rows: the memory is U up to the phases,
// which k_eig_phase then fixes (largest-modulus component real positive, the
// first of equal moduli within 1e-12 -- the host Jacobi's convention).
__global__ void k_eig_phase(cd *U, int nmat, int nb) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long long)nmat * nb) return;
    cd *row = U + r * nb;
    int best = 0;
    double bm = -1.0;
    for (int i = 0; i < nb; i++) {
        const double m = sqrt(row[i].x * row[i].x + row[i].y * row[i].y);
        if (m > bm + 1e-12) { bm = m; best = i; }
    }
    const cd ub = row[best];
    const double mag = sqrt(ub.x * ub.x + ub.y * ub.y);
    if (mag == 0.0) return;
    const cd ph = mk(ub.x / mag, -ub.y / mag);  // conj(u_best) / |u_best|
    for (int i = 0; i < nb; i++) row[i] = cmul(row[i], ph);
}

int edus_eigh_batched(int32_t nmat, int32_t nb, const double *H, double *w, double *U) {
    if (nmat < 0 || nb < 1 || nb > EDUS_MAX_NB || (nmat > 0 && (!H || !w || !U))) return set_err(EDUS_ERR_INVALID, "bad eigensolve arguments");
    if (nmat == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(EDUS_ERR_NODEVICE, "no HIP device visible");
    const size_t nn = (size_t)nb * nb, chunk = std::min<size_t>((size_t)nmat, 32768);
    rocblas_handle hd = nullptr;
    if (rocblas_create_handle(&hd) != rocblas_status_success) return set_err(EDUS_ERR_HIP, "rocblas_create_handle failed");
    hipStream_t st = nullptr;
    void *dA = nullptr, *dW = nullptr, *dres = nullptr, *dinfo = nullptr, *dsw = nullptr;
    int rc = 0;
    auto done = [&](int r) {
        if (st) (void)hipStreamSynchronize(st);
        (void)hipFree(dA); (void)hipFree(dW); (void)hipFree(dres); (void)hipFree(dinfo); (void)hipFree(dsw);
        if (st) (void)hipStreamDestroy(st);
        rocblas_destroy_handle(hd);
        return r;
    };
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || rocblas_set_stream(hd, st) != rocblas_status_success)
        return done(set_err(EDUS_ERR_HIP, "eigensolve stream"));
    if (hipMalloc(&dA, chunk * nn * 16) != hipSuccess || hipMalloc(&dW, chunk * nb * 8) != hipSuccess ||
        hipMalloc(&dres, chunk * 8) != hipSuccess || hipMalloc(&dinfo, chunk * 4) != hipSuccess || hipMalloc(&dsw, chunk * 4) != hipSuccess)
        return done(set_err(EDUS_ERR_NOMEM, "eigensolve buffers"));
    std::vector<int> info(chunk);
    for (size_t m0 = 0; m0 < (size_t)nmat; m0 += chunk) {
        const size_t m = std::min(chunk, (size_t)nmat - m0);
        if (hipMemcpyAsync(dA, H + m0 * nn * 2, m * nn * 16, hipMemcpyHostToDevice, st) != hipSuccess)
            return done(set_err(EDUS_ERR_HIP, "eigensolve upload"));
        const rocblas_status s1 = rocsolver_zheevj_strided_batched(
            hd, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_lower, nb, static_cast<rocblas_double_complex *>(dA), nb,
            (rocblas_stride)nn, 0.0, static_cast<double *>(dres), 100, static_cast<rocblas_int *>(dsw), static_cast<double *>(dW), nb,
            static_cast<rocblas_int *>(dinfo), (rocblas_int)m);
        if (s1 != rocblas_status_success) return done(set_err(EDUS_ERR_HIP, "rocsolver_zheevj_strided_batched failed"));
        const long long rows = (long long)m * nb;
        hipLaunchKernelGGL(k_eig_phase, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, static_cast<cd *>(dA), (int)m, nb);
        if (hipGetLastError() != hipSuccess) return done(set_err(EDUS_ERR_HIP, "k_eig_phase launch"));
        if (hipMemcpyAsync(U + m0 * nn * 2, dA, m * nn * 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(w + m0 * nb, dW, m * nb * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(info.data(), dinfo, m * 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            return done(set_err(EDUS_ERR_HIP, "eigensolve download"));
        for (size_t i = 0; i < m; i++)
            if (info[i] != 0) return done(set_err(EDUS_ERR_HIP, "Jacobi eigensolver did not converge for matrix " + std::to_string(m0 + i)));
    }
    g_err.clear();
    return done(rc);
}

int edus_batch_size(const edus_batch *b) { return b ? (int)b->inst.size() : 0; }

int edus_batch_step_rk4(edus_batch *b, int32_t nsteps, const double *const *fields, const uint8_t *const *grad_on) {
    if (!b) return set_err(EDUS_ERR_INVALID, "null batch");
    HIP_TRY(hipSetDevice(b->dev));
    const int n = (int)b->inst.size();
    if (nsteps < 0 || (nsteps > 0 && (!fields || !grad_on))) return set_err(EDUS_ERR_INVALID, "bad step arguments");
    if (nsteps == 0) return 0;
    for (int i = 0; i < n; i++)
        if (!fields[i] || !grad_on[i]) return set_err(EDUS_ERR_INVALID, "null fields of a batch member");
    int rc = batch_table_capacity(b, (size_t)nsteps);
    if (rc) return rc;
    std::vector<double> rows((size_t)nsteps * 4 * b->R_tot), one;
    std::vector<int> masks(n, 0);
    for (int i = 0; i < n; i++) {
        const edus_ctx *c = b->inst[i];
        const int rs = 7 + c->n_nbr;
        rk4_rows(c, nsteps, fields[i], grad_on[i], one);
        for (size_t r = 0; r < (size_t)nsteps * 4; r++) {
            std::copy(one.begin() + r * rs, one.begin() + (r + 1) * rs, rows.begin() + r * b->R_tot + b->roff[i]);
            for (int x = 0; x < 3; x++)
                if (one[r * rs + x] != 0.0 || one[r * rs + 3 + x] != 0.0) masks[i] |= 1 << x;
        }
    }
    HIP_TRY(hipMemcpyAsync(b->tab, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->masks, masks.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    int launches = 0;
    for (int s = 0; s < nsteps; s++) {
        const size_t ts = (size_t)s * 4 * b->R_tot;
        for (const auto &k : b->cls) if ((rc = batch_launch<RK1>(b, k, ts))) return rc;
        for (const auto &k : b->cls) if ((rc = batch_launch<RK2>(b, k, ts))) return rc;
        for (const auto &k : b->cls) if ((rc = batch_launch<RK3>(b, k, ts))) return rc;
        for (const auto &k : b->cls) if ((rc = batch_launch<RK4S>(b, k, ts))) return rc;
        launches += 4 * (int)b->cls.size();
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipEventSynchronize(b->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_launches = launches;
    for (edus_ctx *c : b->inst) c->nstep += nsteps;
    return 0;
}

int edus_batch_last_step_time(const edus_batch *b, double *ms, int32_t *nlaunch) {
    if (!b) return set_err(EDUS_ERR_INVALID, "null batch");
    if (ms) *ms = b->last_ms;
    if (nlaunch) *nlaunch = b->last_launches;
    return 0;
}

edus_ctx *edus_batch_member(edus_batch *b, int32_t i) {
    if (!b || i < 0 || i >= (int)b->inst.size()) {
        set_err(EDUS_ERR_INVALID, "batch member index out of range");
        return nullptr;
    }
    return b->inst[i];
}

int edus_coulomb_coefficients(const edus_coulomb_gen *g, double *Ao, double *Bo, double *Co, double *Do, double *screen, double *vh) {
    if (!g || !Ao || !Bo || !Co || !Do || !screen || !vh || !g->g || !g->disp) return set_err(EDUS_ERR_INVALID, "null argument");
    const int nb = g->nb, nc = g->ncut, N = g->n_sum, Np = g->n_peak;
    if (nb < 1 || nb > 4) return set_err(EDUS_ERR_UNSUPPORTED, "the coefficient generator supports 1..4 bands");
    if (nc < 1 || N < 3 || Np < 3 || g->n_g < 1) return set_err(EDUS_ERR_INVALID, "ncut, n_sum, n_peak, n_g out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(EDUS_ERR_NODEVICE, "no HIP device visible");
    const int nb2 = nb * nb, npair = nb * (nb - 1) / 2, nw = 1 + 2 * npair, nmm = nc * nc, ntrig = 2 * nc;
    const double twopi = 2 * M_PI;
    auto vq = [&](double q2, double qtf) {
        return (g->rytova_keldysh ? g->eps_area * std::sqrt(q2 + qtf * qtf) * (std::sqrt(q2) * g->r0 + 1.0)
                                  : g->eps_area * std::sqrt(q2 + qtf * qtf));
    };
    // band pairs ic < jc and their beta = b.(d_ic - d_jc)
    std::vector<int> pic, pjc;
    CGenArgs a;
    std::memset(&a, 0, sizeof a);
    for (int ic = 0; ic < nb; ic++)
        for (int jc = ic + 1; jc < nb; jc++) {
            const int p = (int)pic.size();
            pic.push_back(ic);
            pjc.push_back(jc);
            double b1d = 0, b2d = 0;
            for (int x = 0; x < 3; x++) {
                const double dd = g->disp[ic * 3 + x] - g->disp[jc * 3 + x];
                b1d += g->b1[x] * dd;
                b2d += g->b2[x] * dd;
            }
            a.beta1[p] = b1d;
            a.beta2[p] = b2d;
        }
    // Hartree term at q = 0 on the host (Fourier_Coulomb.cpp:173-193): V = 0 where
    // the G offsets sum to <= 0, as written there.
    {
        double w00 = 0;
        std::vector<std::complex<double>> wp(npair, 0.0);
        for (int ig = 0; ig < g->n_g; ig++) {
            const double gx = g->g[2 * ig], gy = g->g[2 * ig + 1];
            double cart[3];
            for (int x = 0; x < 3; x++) cart[x] = g->b1[x] * gx + g->b2[x] * gy;
            const double q2 = cart[0] * cart[0] + cart[1] * cart[1] + cart[2] * cart[2];
            const double V = ((gx + gy) < 1e-16) ? 0.0 : twopi / vq(q2, g->qtf);
            w00 += V;
            for (int p = 0; p < npair; p++) {
                double phi = 0;
                for (int x = 0; x < 3; x++) phi += cart[x] * (g->disp[pic[p] * 3 + x] - g->disp[pjc[p] * 3 + x]);
                wp[p] += V * std::exp(std::complex<double>(0.0, phi));
            }
        }
        for (int i = 0; i < nb2 * 2; i++) vh[i] = 0.0;
        for (int ic = 0; ic < nb; ic++) vh[(ic * nb + ic) * 2] = w00;
        for (int p = 0; p < npair; p++) {
            const double r = wp[p].real();
            vh[(pic[p] * nb + pjc[p]) * 2] = r;
            vh[(pjc[p] * nb + pic[p]) * 2] = r;
        }
    }
    // per-G data: offsets and e^{i G.beta_p}
    std::vector<double> gd((size_t)g->n_g * 2), gph((size_t)g->n_g * std::max(npair, 1) * 2, 0.0);
    for (int ig = 0; ig < g->n_g; ig++) {
        const double gx = g->g[2 * ig], gy = g->g[2 * ig + 1];
        gd[2 * ig] = gx;
        gd[2 * ig + 1] = gy;
        for (int p = 0; p < npair; p++) {
            const double ph = gx * a.beta1[p] + gy * a.beta2[p];
            gph[((size_t)ig * npair + p) * 2] = std::cos(ph);
            gph[((size_t)ig * npair + p) * 2 + 1] = std::sin(ph);
        }
    }
    // Simpson weights and trig tables (integrationWeight.h:5-18)
    auto simpson = [](int n, std::vector<double> &q, std::vector<double> &w) {
        q.resize(n);
        w.resize(n);
        const double h = 2 * 0.5 / (n - 1), hs = 3 * (n - 1);
        for (int i = 0; i < n; i++) q[i] = -0.5 + i * h;
        w[0] = 1 / hs;
        for (int i = 1; i < n; i += 2) w[i] = 4.0 / hs;
        for (int i = 2; i < n; i += 2) w[i] = 2.0 / hs;
        w[n - 1] = 1 / hs;
    };
    std::vector<double> q, w;
    simpson(N, q, w);
    std::vector<double> wtrig((size_t)ntrig * N);
    for (int t = 0; t < ntrig; t++)
        for (int i = 0; i < N; i++) {
            const int m = t < nc ? t : t - nc;
            const double arg = 2 * M_PI * m * q[i];
            wtrig[(size_t)t * N + i] = w[i] * (t < nc ? std::cos(arg) : std::sin(arg));
        }
    const size_t nq = (size_t)N * N;
    double *dG = nullptr, *dPh = nullptr, *dW = nullptr, *dTr = nullptr, *dT = nullptr, *dS = nullptr, *dCoef = nullptr, *dW0 = nullptr,
           *dPart = nullptr;
    int *dQl = nullptr;
    auto cleanup = [&]() {
        for (void *p : {(void *)dG, (void *)dPh, (void *)dW, (void *)dTr, (void *)dT, (void *)dS, (void *)dCoef, (void *)dW0, (void *)dPart,
                        (void *)dQl})
            (void)hipFree(p);
    };
    auto fail = [&](int code, const std::string &m) {
        cleanup();
        return set_err(code, m);
    };
    if (hipMalloc(&dG, gd.size() * 8) != hipSuccess || hipMalloc(&dPh, gph.size() * 8) != hipSuccess ||
        hipMalloc(&dW, nq * nw * 8) != hipSuccess || hipMalloc(&dTr, wtrig.size() * 8) != hipSuccess ||
        hipMalloc(&dT, (size_t)ntrig * N * nw * 8) != hipSuccess || hipMalloc(&dS, (size_t)ntrig * ntrig * nw * 8) != hipSuccess ||
        hipMalloc(&dCoef, (size_t)4 * nmm * nb2 * 2 * 8) != hipSuccess)
        return fail(EDUS_ERR_NOMEM, "coefficient generator buffers");
    if (hipMemcpy(dG, gd.data(), gd.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dPh, gph.data(), gph.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dTr, wtrig.data(), wtrig.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
        return fail(EDUS_ERR_HIP, "coefficient generator upload");
    a.npair = npair; a.n_g = g->n_g; a.N = N; a.mode = 0; a.h = 2 * 0.5 / (N - 1);
    a.qtf2 = g->qtf * g->qtf; a.eps = g->eps_area; a.r0 = g->r0; a.rk = g->rytova_keldysh ? 1.0 : 0.0;
    for (int x = 0; x < 3; x++) { a.b1[x] = g->b1[x]; a.b2[x] = g->b2[x]; }
    a.g = dG; a.gph = dPh; a.qlist = nullptr;
    // W on the Simpson grid, in chunks with a progress line each
    const long long chunk = 1LL << 21;
    for (long long q0 = 0; q0 < (long long)nq; q0 += chunk) {
        const int n = (int)std::min<long long>(chunk, (long long)nq - q0);
        hipLaunchKernelGGL(k_cgen_w, dim3((n + 255) / 256), dim3(256), 0, 0, a, q0, n, dW + (size_t)q0 * nw);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return fail(EDUS_ERR_HIP, "k_cgen_w");
        if (g->verbose) std::fprintf(stderr, "Coulomb coefficients: %lld / %zu q points\n", q0 + n, nq);
    }
    {
        const long long n1 = (long long)ntrig * nw * N;
        hipLaunchKernelGGL(k_cgen_pass1, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, 0, (const double *)dW, (const double *)dTr, N, ntrig, nw, dT);
        const int n2 = ntrig * ntrig * nw;
        hipLaunchKernelGGL(k_cgen_pass2, dim3((n2 + 255) / 256), dim3(256), 0, 0, (const double *)dT, (const double *)dTr, N, ntrig, nw, dS);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return fail(EDUS_ERR_HIP, "k_cgen_pass");
    }
    std::vector<double> S((size_t)ntrig * ntrig * nw);
    if (hipMemcpy(S.data(), dS, S.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return fail(EDUS_ERR_HIP, "download S");
    auto sv = [&](int t1, int t2, int c) { return S[((size_t)t1 * ntrig + t2) * nw + c]; };
    for (size_t i = 0; i < (size_t)nmm * nb2 * 2; i++) Ao[i] = Bo[i] = Co[i] = Do[i] = 0.0;
    for (int mm = 0; mm < nmm; mm++) {
        const int m = mm / nc, n = mm % nc;
        const double sc = 4.0 * ((m == 0 && n == 0) ? 0.25 : ((m == 0 || n == 0) ? 0.5 : 1.0));
        Ao[(size_t)mm * nb2 * 2] = sc * sv(m, n, 0);
        Do[(size_t)mm * nb2 * 2] = sc * sv(nc + m, nc + n, 0);
        for (int p = 0; p < npair; p++) {
            const size_t o = ((size_t)mm * nb2 + pic[p] * nb + pjc[p]) * 2;
            Ao[o] = sc * sv(m, n, 1 + 2 * p);
            Bo[o + 1] = sc * sv(m, nc + n, 2 + 2 * p);
            Co[o + 1] = sc * sv(nc + m, n, 2 + 2 * p);
            Do[o] = sc * sv(nc + m, nc + n, 1 + 2 * p);
        }
    }
    // correction around q = 0: points of the n_peak grid with 1e-17 < |q|^2 < q_peak^2
    for (int i = 0; i < nb2 * 2; i++) screen[i] = 0.0;
    const double hp = 2 * 0.5 / (Np - 1);
    std::vector<int> ql;
    {
        const double g11 = g->b1[0] * g->b1[0] + g->b1[1] * g->b1[1] + g->b1[2] * g->b1[2];
        const double g22 = g->b2[0] * g->b2[0] + g->b2[1] * g->b2[1] + g->b2[2] * g->b2[2];
        const double g12 = g->b1[0] * g->b2[0] + g->b1[1] * g->b2[1] + g->b1[2] * g->b2[2];
        const double lmin = 0.5 * (g11 + g22) - std::sqrt(0.25 * (g11 - g22) * (g11 - g22) + g12 * g12);
        const double rc = (lmin > 0) ? g->q_peak / std::sqrt(lmin) : 1.0;
        const int lo = std::max(0, (int)std::floor((0.5 - rc) / hp) - 2), hi = std::min(Np - 1, (int)std::ceil((0.5 + rc) / hp) + 2);
        const double qp2 = g->q_peak * g->q_peak;
        for (int ix = lo; ix <= hi; ix++)
            for (int iy = lo; iy <= hi; iy++) {
                const double qx = -0.5 + ix * hp, qy = -0.5 + iy * hp;
                double cart[3];
                for (int x = 0; x < 3; x++) cart[x] = g->b1[x] * qx + g->b2[x] * qy;
                const double q2 = cart[0] * cart[0] + cart[1] * cart[1] + cart[2] * cart[2];
                if (q2 < qp2 && q2 > std::pow(10, -17)) {
                    ql.push_back(ix);
                    ql.push_back(iy);
                }
            }
    }
    const int npk = (int)(ql.size() / 2);
    if (npk > 0) {
        const int nblk = (npk + 255) / 256;
        std::vector<double> coef((size_t)4 * nmm * nb2 * 2);
        std::copy(Ao, Ao + (size_t)nmm * nb2 * 2, coef.begin());
        std::copy(Bo, Bo + (size_t)nmm * nb2 * 2, coef.begin() + (size_t)nmm * nb2 * 2);
        std::copy(Co, Co + (size_t)nmm * nb2 * 2, coef.begin() + (size_t)2 * nmm * nb2 * 2);
        std::copy(Do, Do + (size_t)nmm * nb2 * 2, coef.begin() + (size_t)3 * nmm * nb2 * 2);
        if (hipMalloc(&dQl, ql.size() * sizeof(int)) != hipSuccess || hipMalloc(&dW0, (size_t)npk * nw * 8) != hipSuccess ||
            hipMalloc(&dPart, (size_t)nblk * nb2 * 2 * 8) != hipSuccess)
            return fail(EDUS_ERR_NOMEM, "peak buffers");
        if (hipMemcpy(dQl, ql.data(), ql.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dCoef, coef.data(), coef.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(dPart, 0, (size_t)nblk * nb2 * 2 * 8) != hipSuccess)
            return fail(EDUS_ERR_HIP, "peak upload");
        CGenArgs b = a;
        b.mode = 1; b.N = Np; b.h = hp; b.qtf2 = 0.0; b.qlist = dQl;
        hipLaunchKernelGGL(k_cgen_w, dim3(nblk), dim3(256), 0, 0, b, 0LL, npk, dW0);
        const size_t cs = (size_t)nmm * nb2 * 2;
        hipLaunchKernelGGL(k_cgen_peak, dim3(nblk), dim3(256), 0, 0, (const double *)dW0, (const int *)dQl, npk, nb, nc, hp,
                           1. / ((double)Np * Np), (const double *)dCoef, (const double *)(dCoef + cs), (const double *)(dCoef + 2 * cs),
                           (const double *)(dCoef + 3 * cs), dPart);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return fail(EDUS_ERR_HIP, "peak kernels");
        std::vector<double> part((size_t)nblk * nb2 * 2);
        if (hipMemcpy(part.data(), dPart, part.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return fail(EDUS_ERR_HIP, "download peak");
        for (int b2i = 0; b2i < nblk; b2i++)
            for (int i = 0; i < nb2 * 2; i++) screen[i] += part[(size_t)b2i * nb2 * 2 + i];
        if (g->verbose) std::fprintf(stderr, "Coulomb coefficients: %d points in the q ~ 0 correction\n", npk);
    }
    cleanup();
    g_err.clear();
    return 0;
}

// Twiddles e^{2 pi i m / N_l} of the three grid directions, evaluated on the
// host in long double, and the offset of each direction's run.
static void wannier_twiddles(const int32_t nk3[3], std::vector<double> &tw, int offs[3]) {
    tw.clear();
    for (int l = 0; l < 3; l++) {
        offs[l] = (int)(tw.size() / 2);
        for (int m = 0; m < nk3[l]; m++) {
            long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)nk3[l];
            tw.push_back((double)cosl(ang));
            tw.push_back((double)sinl(ang));
        }
    }
}

// form: 0 FP64 matrix cores, 1 FP32, 2 BF16x3 split products.
static int wannier_hk(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R, const double *HR, const double *rR, double *H_out,
                     double *D_out, int form) {
    if (!nk3 || !R || !HR || !rR || !H_out || nR < 1 || nw < 1) return set_err(EDUS_ERR_INVALID, "bad Wannier arguments");
    for (int l = 0; l < 3; l++)
        if (nk3[l] < 1) return set_err(EDUS_ERR_INVALID, "bad grid");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(EDUS_ERR_NODEVICE, "no HIP device visible");
    const long long nk = (long long)nk3[0] * nk3[1] * nk3[2];
    const int nw2 = nw * nw, nm = 4 * nw2;
    // Val[R][c][i][j]: c = 0 H, 1..3 r_x, r_y, r_z
    std::vector<double> val((size_t)nR * nm * 2);
    for (int r = 0; r < nR; r++)
        for (int t = 0; t < nw2; t++) {
            val[((size_t)r * nm + t) * 2] = HR[((size_t)r * nw2 + t) * 2];
            val[((size_t)r * nm + t) * 2 + 1] = HR[((size_t)r * nw2 + t) * 2 + 1];
            for (int x = 0; x < 3; x++) {
                val[((size_t)r * nm + (x + 1) * nw2 + t) * 2] = rR[((size_t)r * nw2 + t) * 6 + 2 * x];
                val[((size_t)r * nm + (x + 1) * nw2 + t) * 2 + 1] = rR[((size_t)r * nw2 + t) * 6 + 2 * x + 1];
            }
        }
    std::vector<double> tw;
    int offs[3];
    wannier_twiddles(nk3, tw, offs);
    int *dR = nullptr;
    double *dVal = nullptr, *dTw = nullptr, *dOut = nullptr, *dPT = nullptr, *dH = nullptr, *dD = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dR); (void)hipFree(dVal); (void)hipFree(dTw); (void)hipFree(dOut); (void)hipFree(dPT);
        (void)hipFree(dH); (void)hipFree(dD);
    };
    const size_t npt = (size_t)nR * (nk3[0] + nk3[1] + nk3[2]);
    const size_t chunk = std::min<long long>(nk, std::max<long long>(64, (1LL << 28) / ((long long)nm * 16)));
    if (hipMalloc(&dR, (size_t)nR * 3 * sizeof(int)) != hipSuccess || hipMalloc(&dVal, val.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&dTw, tw.size() * sizeof(double)) != hipSuccess || hipMalloc(&dOut, chunk * nm * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc(&dPT, npt * 2 * sizeof(double)) != hipSuccess || hipMalloc(&dH, chunk * nw2 * 2 * sizeof(double)) != hipSuccess ||
        (D_out && hipMalloc(&dD, chunk * nw2 * 6 * sizeof(double)) != hipSuccess)) {
        cleanup();
        return set_err(EDUS_ERR_NOMEM, "Wannier buffers");
    }
    if (hipMemcpy(dR, R, (size_t)nR * 3 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dVal, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dTw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return set_err(EDUS_ERR_HIP, "Wannier upload");
    }
    hipLaunchKernelGGL(k_wannier_ptab, dim3((unsigned)std::min<size_t>((npt + 255) / 256, 65536)), dim3(256), 0, 0, (const int *)dR,
                       (const cd *)dTw, nR, nk3[0], nk3[1], nk3[2], offs[1], offs[2], (cd *)dPT);
    for (long long k0 = 0; k0 < nk; k0 += chunk) {
        const int kc = (int)std::min<long long>(chunk, nk - k0);
        const dim3 grid((nm + WMN - 1) / WMN, (unsigned)((kc + WMK - 1) / WMK));
        if (form == 2)
            hipLaunchKernelGGL(k_wannier_mfma_bf16x3, grid, dim3(256), 0, 0, (const cd *)dPT, (const cd *)dVal, nk3[0], nk3[1], nk3[2], nR,
                               nm, kc, (int)k0, (cd *)dOut);
        else if (form == 1)
            hipLaunchKernelGGL(k_wannier_mfma32, grid, dim3(256), 0, 0, (const cd *)dPT, (const cd *)dVal, nk3[0], nk3[1], nk3[2], nR, nm,
                               kc, (int)k0, (cd *)dOut);
        else
            hipLaunchKernelGGL(k_wannier_mfma, grid, dim3(256), 0, 0, (const cd *)dPT, (const cd *)dVal, nk3[0], nk3[1], nk3[2], nR, nm,
                               kc, (int)k0, (cd *)dOut);
        const long long nt = (long long)kc * nw2;
        hipLaunchKernelGGL(k_wannier_finish, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, 0, (const cd *)dOut, kc, nw, dH, dD, 0);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(H_out + (size_t)k0 * nw2 * 2, dH, (size_t)nt * 16, hipMemcpyDeviceToHost);
        if (e == hipSuccess && D_out) e = hipMemcpy(D_out + (size_t)k0 * nw2 * 6, dD, (size_t)nt * 48, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            cleanup();
            return set_err(EDUS_ERR_HIP, std::string("k_wannier: ") + hipGetErrorString(e));
        }
    }
    cleanup();
    g_err.clear();
    return 0;
}

int edus_wannier_hk(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R, const double *HR, const double *rR, double *H_out,
                    double *D_out) {
    return wannier_hk(nk3, nR, nw, R, HR, rR, H_out, D_out, 0);
}

int edus_wannier_hk_fp32(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R, const double *HR, const double *rR,
                         double *H_out, double *D_out) {
    return wannier_hk(nk3, nR, nw, R, HR, rR, H_out, D_out, 1);
}

int edus_wannier_hk_bf16x3(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R, const double *HR, const double *rR,
                           double *H_out, double *D_out) {
    return wannier_hk(nk3, nR, nw, R, HR, rR, H_out, D_out, 2);
}

// The contraction of edus_wannier_hk with every device buffer carved from the
// fixed scratch array: lattice vectors, twiddles, phase tables, H(R) and r(R)
// (copied in with 2D copies, the r part of each row in the caller's [i][j][x]
// order, so no host repacking), then the grid in as many k chunks as the rest
// of the array holds.  Nothing in the array needs clearing: every buffer is
// written before it is read.
int edus_wannier_hk_noalloc(const int32_t nk3[3], int32_t nR, int32_t nw, const int32_t *R, const double *HR, const double *rR,
                            double *H_out, double *D_out) {
    if (!nk3 || !R || !HR || !rR || !H_out || nR < 1 || nw < 1) return set_err(EDUS_ERR_INVALID, "bad Wannier arguments");
    for (int l = 0; l < 3; l++)
        if (nk3[l] < 1) return set_err(EDUS_ERR_INVALID, "bad grid");
    const long long nk = (long long)nk3[0] * nk3[1] * nk3[2];
    const long long nw2 = (long long)nw * nw, nm = 4 * nw2, nsum = (long long)nk3[0] + nk3[1] + nk3[2];
    auto al = [](long long b) { return (b + 255) / 256 * 256; };
    const long long oR = 0, oTw = al(oR + 12LL * nR), oPT = al(oTw + 16 * nsum), oVal = al(oPT + 16LL * nR * nsum),
                    oOut = al(oVal + 16LL * nR * nm);
    const long long per_k = 16 * nm + 16 * nw2 + (D_out ? 48 * nw2 : 0), slack = 512;
    const long long cap = EDUS_NOALLOC_SCRATCH_BYTES;
    if (oOut + per_k + slack > cap)
        return set_err(EDUS_ERR_UNSUPPORTED, "edus_wannier_hk_noalloc: the model needs " + std::to_string(oOut + per_k + slack) +
                                                 " bytes of scratch, above the " + std::to_string(cap) +
                                                 " of EDUS_NOALLOC_SCRATCH_BYTES (use edus_wannier_hk)");
    const long long chunk = std::min<long long>(nk, (cap - oOut - slack) / per_k);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(EDUS_ERR_NODEVICE, "no HIP device visible");
    void *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    char *b = static_cast<char *>(base);
    int *dR = reinterpret_cast<int *>(b + oR);
    double *dTw = reinterpret_cast<double *>(b + oTw), *dPT = reinterpret_cast<double *>(b + oPT);
    cd *dVal = reinterpret_cast<cd *>(b + oVal), *dOut = reinterpret_cast<cd *>(b + oOut);
    const long long oH = al(oOut + 16 * chunk * nm), oD = al(oH + 16 * chunk * nw2);
    double *dH = reinterpret_cast<double *>(b + oH), *dD = D_out ? reinterpret_cast<double *>(b + oD) : nullptr;
    std::vector<double> tw;
    int offs[3];
    wannier_twiddles(nk3, tw, offs);
    HIP_TRY(hipMemcpy(dR, R, (size_t)nR * 3 * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dTw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
    // Val[R] = [H(R) (nw^2) | r(R) as [i][j][x] (3 nw^2)]
    HIP_TRY(hipMemcpy2D(dVal, (size_t)nm * 16, HR, (size_t)nw2 * 16, (size_t)nw2 * 16, (size_t)nR, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy2D(dVal + nw2, (size_t)nm * 16, rR, (size_t)3 * nw2 * 16, (size_t)3 * nw2 * 16, (size_t)nR,
                        hipMemcpyHostToDevice));
    const size_t npt = (size_t)nR * nsum;
    hipLaunchKernelGGL(k_wannier_ptab, dim3((unsigned)std::min<size_t>((npt + 255) / 256, 65536)), dim3(256), 0, 0, (const int *)dR,
                       (const cd *)dTw, nR, nk3[0], nk3[1], nk3[2], offs[1], offs[2], (cd *)dPT);
    HIP_TRY(hipGetLastError());
    for (long long k0 = 0; k0 < nk; k0 += chunk) {
        const int kc = (int)std::min<long long>(chunk, nk - k0);
        const dim3 grid((unsigned)((nm + WMN - 1) / WMN), (unsigned)((kc + WMK - 1) / WMK));
        hipLaunchKernelGGL(k_wannier_mfma, grid, dim3(256), 0, 0, (const cd *)dPT, (const cd *)dVal, nk3[0], nk3[1], nk3[2], nR,
                           (int)nm, kc, (int)k0, dOut);
        const long long nt = (long long)kc * nw2;
        hipLaunchKernelGGL(k_wannier_finish, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, 0, (const cd *)dOut, kc, nw, dH, dD, 1);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpy(H_out + (size_t)k0 * nw2 * 2, dH, (size_t)nt * 16, hipMemcpyDeviceToHost));
        if (D_out) HIP_TRY(hipMemcpy(D_out + (size_t)k0 * nw2 * 6, dD, (size_t)nt * 48, hipMemcpyDeviceToHost));
    }
    g_err.clear();
    return 0;
}

}  // extern "C"
