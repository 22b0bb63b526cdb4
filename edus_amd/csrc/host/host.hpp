// Internal declarations of the EDUS host library (C++17).
// The host side mirrors the reference's setup path: input grammar
// (headers/ReadInput.h), lattice and k-grid (InitParameters.h, Coordinate.h),
// gradient shells (CalculateWeigths.h), models (TightBinding.h, Wannier.h),
// initial state (InitialPopulation.h) and the observable writers
// (Observables_MPI.h).  Only the arithmetic that fixes results is kept alike;
// the code structure is this project's own.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace edus {

using cplx = std::complex<double>;

// Allocator whose value-initialisation does nothing: the per-k tables (up to
// 10^10 bytes) are sized without a serial zero fill and then filled by parallel
// loops (first touch) or copies.
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U> &) noexcept {}
    template <class U>
    void construct(U *) noexcept {}
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
using ctable = std::vector<cplx, UninitAlloc<cplx>>;

// Sizes a table to n elements and zeroes it with all threads.
void zero_table(ctable &v, size_t n);

// Unit conversions with the digits of headers/Constants.h:14-37 (they fix
// every time, energy and field value the reference writes).
struct Units {
    static double pi() { return 4.0 * std::atan(1.0); }
    static constexpr double c_light = 137.036;
    static constexpr double time_au_fs = 0.02418884326505;
    static double time_fs_au() { return 1.0 / 0.02418884326505; }
    static constexpr double energy_au_eV = 27.211396;
    static double energy_eV_au() { return 1.0 / 27.211396; }
    static constexpr double space_au_A = 0.52917721092;
    static double space_A_au() { return 1.0 / 0.52917721092; }
    static constexpr double intensity_au_Wcm2 = 3.50944758e+16;
    static constexpr double Efield_Vm_au = 5.14221e11;
};

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------- laser
// A pulse with the parametrisation and setter side effects of headers/Laser.h:1-51.
struct Pulse {
    double lambdanm = 0, wl = 0, Period = 0, Intensity = 0, A0 = 0, E0 = 0;
    double t0 = 0, tf = 0, w1 = 0, w2 = 0, ncycle = 0, sigma = 0, phase = 0;
    bool envelope = false, gaussian = false, PW = false, sin2 = false;
    std::array<double, 3> pol1{0, 0, 0}, pol2{0, 0, 0};

    void set_freq(double a);
    void set_lambdanm(double a);
    void set_intensity_wcm2(double a);
    void set_ncycle_sigma(double a);
    void set_t0(double a);
    void set_pol(double a0, double a1, double a2, double b0, double b1, double b2);
    // Electric field in cartesian a.u. (Laser.h:128-144).
    void field(double t, double out[3]) const;
};

// ---------------------------------------------------------------- input
struct Input {
    std::string ref_coord;                 // first token of input.txt (Wannier seed)
    std::string base_dir;                  // directory of input.txt
    double a[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double b[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    int Nk[3] = {0, 0, 0};
    bool kptread = false;
    std::vector<double> kpt_list;          // nkPT read: [Nk0 Nk1 Nk2][3] crystal k points of the input
    int Nb[3] = {0, 0, 0};
    double change_gap_constant = 0;
    std::vector<Pulse> pumps;
    int num_pump_pulses = 1;
    Pulse probe;
    double DELAY = 0;
    std::string iMode = "none";
    std::string TBtype = "CoreGraphene";
    std::vector<double> hopping;
    double dt = 0.1, t_fin = 0.0, FermiE = 0.0;
    bool const_dt = true, dyn_dt = false, taylor = false, adams_bashforth = false;
    double epsStepAbs = 1e-2, gap_correction = 0;
    int solver_order = 5;
    // PrintPopulation [t0 t1 dt] (a.u.; defaults of variables.cpp:84-86) and the
    // Coulomb section's Print_new_band_dispersion: k-resolved snapshots.
    bool print_population = false, print_band_dispersion = false;
    double print_start = 0.0, print_end = 100.0 * 0.02418884326505, print_step = 20.0 * 0.02418884326505;
    int it_resolution = 10;
    bool iCurrent = false, iTAbs = true, iTAbsK = false, iWFDs = false;
    bool vectorization = true;
    // Coulomb section.  The mean-field term runs with coefficients read from
    // files (coulomb_read) in the Wannier basis; the other options are parsed
    // for their error messages only.
    bool coulomb = false, coulomb_diag_basis = false, rytova_keldysh = false, coulomb_read = false;
    double qTF = 0, epsilon_static = 1.0, r0 = 18.89726, G_distance = 50;
    int Ncut = 30;
    std::string coulomb_label = "_";
    double T1 = 0, T2 = 0, Tch = 0;
};

// Parses the reference grammar (ReadInput.h:332-1234); throws Error with the
// reference's message on malformed input.
Input parse_input(const std::string &path);

// ---------------------------------------------------------------- lattice
// Crystal <-> cartesian transform of Coordinate.h:68-149 for reciprocal space.
struct Lattice {
    double M[3][3];   // cart = M * crys, M[i][j] = b[j][i]
    double MI[3][3];  // inverse
    double jacobian;  // |det M|
    void set_from_b(const double b[3][3]);
    void to_cart(const double crys[3], double cart[3]) const;
    void to_crys(const double cart[3], double crys[3]) const;
};

void reciprocal_from_direct(const double a[3][3], double b[3][3]);

struct Shells {
    std::vector<std::vector<std::array<double, 3>>> bvec;  // [shell][b] cartesian
    std::vector<std::array<int, 3>> offsets;               // flattened [n_nbr] grid offsets
    std::vector<double> weights;                           // [shell]
    double residual = 0;
};

// k-grid of BZ_parameters (InitParameters.h:13-40): ik -> (i0,i1,i2), i2 fastest.
std::vector<double> make_kgrid(const int Nk[3]);
// grid: kpt is the uniform grid of Nk (shell vectors then carry integer offsets);
// false for an explicit k list (nkPT read).
Shells find_shells(const std::vector<double> &kpt, const int Nk[3], const Lattice &L, bool grid = true);
std::vector<int32_t> neighbours_closed_form(const int Nk[3], const Shells &sh);
// The reference's search of k + b in an explicit k list, by hashing (first match in list order).
std::vector<int32_t> neighbours_search(const std::vector<double> &kpt, const Shells &sh, const Lattice &L);
std::vector<int32_t> neighbours_bruteforce(const std::vector<double> &kpt, const Shells &sh, const Lattice &L);

// ---------------------------------------------------------------- small linear algebra
void inv3(const double A[3][3], double Ainv[3][3], double *det);
std::vector<double> lstsq(const std::vector<double> &A, int m, int n, const std::vector<double> &q);
std::vector<double> inverse_gauss_jordan(const std::vector<double> &A, int n);
// Hermitian eigen-decomposition (cyclic Jacobi): ascending eigenvalues, columns of V.
void eigh_jacobi(int n, const cplx *H, double *w, cplx *V);

// ---------------------------------------------------------------- models
struct TBModel {
    std::string name;
    int Ncv = 0, Nch = 0;
    double a = 0, b = 0, c = 0, d = 0, f = 0, g = 0, t2 = 0, phi0 = 0, core = 0, rchv = 0, rchc = 0;
    double vampa[2][3][6] = {};  // Vampa2015 band coefficients [band][direction][harmonic]
    double kcart[3][3] = {};     // crystal -> cartesian k (columns b_1..b_3), for models written in cartesian k
    void init(const std::string &nm, int nch, int ncv, const std::vector<double> &hopping, const double b_rows[3][3] = nullptr);
    bool supported() const;
    cplx energy(int ic, int jc, const double k[3]) const;
    cplx unitary(int ic, int jc, const double k[3]) const;
    cplx dipole(int ic, int jc, int x, const double k[3]) const;
};

struct WannierModel {
    int nw = 0, nR = 0;
    std::vector<int32_t> R;        // [nR][3]
    std::vector<double> ndegen;    // [nR]
    std::vector<cplx> HR;          // [nR][nw][nw]  (a.u., / ndegen)
    std::vector<cplx> rR;          // [nR][nw][nw][3]
    void read_tb(const std::string &path);
};

// Crystal interface (Crystal_interface.h): tabulated H and Berry connections,
// interpolated by periodic bicubic convolution (Interpolation.h, 2D case).
struct CrystalModel {
    int norb = 0, nkf = 0;
    std::vector<cplx> h;   // [norb][norb][nkf][nkf], entries jc <= ic
    std::vector<cplx> bc;  // [norb][norb][2][nkf][nkf], entries jc <= ic
    void read(const std::string &dir, double a[3][3], double b[3][3]);
    cplx interp(const cplx *grid, const double k[3]) const;
    cplx energy(int ic, int jc, const double k[3]) const;
    cplx dipole(int ic, int jc, int x, const double k[3]) const;
};

// ---------------------------------------------------------------- setup
struct Setup {
    Input in;
    CrystalModel cy;   // iMode "CY"
    Lattice L;
    int nk = 0, nb = 0, nch = 0, ncc = 0;
    std::vector<double> kpt, kcart;
    Shells sh;
    std::vector<int32_t> nbr;      // [nk][n_nbr]
    std::vector<double> nbr_wb;    // [n_nbr][3]
    std::vector<double> bvec_flat; // [n_nbr][3]
    ctable H, U, rho0;             // [nk][nb][nb]
    ctable D, gradH;               // [nk][nb][nb][3]
    std::vector<double> T;         // [nb][nb]
    std::vector<double> tau_inv;   // Taylor
    int n_diss = 0;
    double det_a = 0, det_b = 0, t_start = 0;
    int64_t itfi = 0;
    int stepper = 0, order = 1;
    std::vector<double> times;     // accumulated time_loop per step, [itfi+2]
    // Coulomb mean field (read from the coefficient files)
    int ncut = 0;
    std::vector<cplx> cA, cB, cC, cD, cscreen, cvh;
    std::string coulomb_label_written;  // label of generated coefficient files
    void read_coulomb();
    std::vector<double> disp_at_reference_point() const;
    void generate_coulomb(const std::vector<double> &disp);

    void build(const std::string &input_path, int flags);
    void fields_at(double t, double ef[6]) const;   // pump sum + 0.5 probe
    void step_fields(int64_t it, double *f, uint8_t *grad_on) const;
    int fields_per_step() const { return stepper == 0 ? 3 : 1; }
};

// tau^{-1} of Taylor_DE_solver_deriv_Polynom.cpp:95-123 for the dt history
// dt_prev [order] (most recent first): t_0 = dt_prev[0], t_i = t_{i-1} + dt_prev[i-1]
// as written there, tau_ip = 1e7 t_i^(p+1), inverted, then scaled by 1e7.
std::vector<double> taylor_tau_inverse(const std::vector<double> &dt_prev);

void write_synthetic_tb(const std::string &path, int nw, int rmax, double a_ang, double sigma_h_ev,
                        double sigma_r_ang, double decay_ang, uint64_t seed);

}  // namespace edus
