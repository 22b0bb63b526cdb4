// Builds every table the time-stepping path reads, from an input.txt.
//
// Order and arithmetic follow the reference's setup: Source_Main/input.cpp
// (parse, BZ_parameters, model init, det_a, CalculateWeigths), save_HUD.cpp
// (H/U/D per k, scissor, Wannier dipole symmetrisation, GradientEnergy),
// allocate_arr.cpp + InitParameters.h:203-224 (T), main_MPI.cpp:66-79
// (nonzero-T list), runge_kutta_init.cpp + InitialPopulation.h:377-414 (rho0),
// const_dt_evolution.cpp:164-169 (start time, last step).
#include "host.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <sys/stat.h>
#include <fstream>
#include <omp.h>

#include "../../../include/edus.h"

namespace edus {

static inline size_t mi(int nb, int k, int i, int j) { return ((size_t)k * nb + i) * nb + j; }

void zero_table(ctable &v, size_t n) {
    v.clear();
    v.resize(n);
    cplx *p = v.data();
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; i++) p[i] = cplx(0., 0.);
}

void Setup::build(const std::string &input_path, int flags) {
    // EDUS_SETUP_TIMING=1 prints the wall time of each setup phase to stderr.
    const bool timing = std::getenv("EDUS_SETUP_TIMING") != nullptr;
    double t_last = omp_get_wtime();
    auto mark = [&](const char *phase) {
        const double t = omp_get_wtime();
        if (timing) std::fprintf(stderr, "[edus setup] %-26s %8.3f s\n", phase, t - t_last);
        t_last = t;
    };
    in = parse_input(input_path);
    if (in.iMode != "TB" && in.iMode != "W" && in.iMode != "CY") throw Error("iMode = " + in.iMode + " is incorrect");
    // The crystal interface replaces the lattice with the one of its k-mesh file
    // (input.cpp:57, Crystal_interface.h:58-96).
    if (in.iMode == "CY") cy.read(in.base_dir, in.a, in.b);

    L.set_from_b(in.b);
    // BZ_parameters re-applies the polarisation of every pulse (InitParameters.h:4-11).
    for (auto &p : in.pumps) p.set_pol(p.pol1[0], p.pol1[1], p.pol1[2], p.pol2[0], p.pol2[1], p.pol2[2]);
    in.probe.set_pol(in.probe.pol1[0], in.probe.pol1[1], in.probe.pol1[2], in.probe.pol2[0], in.probe.pol2[1], in.probe.pol2[2]);

    for (int c = 0; c < 3; c++)
        if (in.Nk[c] < 1) throw Error("nkPT section missing or invalid");
    kpt = in.kptread ? in.kpt_list : make_kgrid(in.Nk);
    nk = in.Nk[0] * in.Nk[1] * in.Nk[2];
    nb = in.Nb[0] + in.Nb[1] + in.Nb[2];
    nch = in.Nb[0];
    ncc = in.Nb[0] + in.Nb[1];
    if (nb < 1 || nb > EDUS_MAX_NB) throw Error("band count out of range (1.." + std::to_string(EDUS_MAX_NB) + ")");
    kcart.resize((size_t)nk * 3);
    for (int ik = 0; ik < nk; ik++) L.to_cart(&kpt[(size_t)ik * 3], &kcart[(size_t)ik * 3]);

    // In-plane cell areas (input.cpp:65-73).
    if (in.Nk[2] < in.Nk[1]) {  // sample in xy
        det_b = std::fabs(in.b[0][0] * in.b[1][1] - in.b[0][1] * in.b[1][0]);
        det_a = std::fabs(in.a[0][0] * in.a[1][1] - in.a[0][1] * in.a[1][0]);
    } else {                     // sample in yz (default "011")
        det_b = std::fabs(in.b[1][1] * in.b[2][2] - in.b[1][2] * in.b[2][1]);
        det_a = std::fabs(in.a[1][1] * in.a[2][2] - in.a[1][2] * in.a[2][1]);
    }

    if (in.coulomb) {
        if (in.Nk[0] != 1 || in.Nk[2] < in.Nk[1]) throw Error("the Coulomb term needs a 1 x N1 x N2 grid in the yz plane");
        if (nb > 4) throw Error("the Coulomb term is implemented for up to 4 bands");
        if (in.coulomb_read) read_coulomb();
    }
    if (in.kptread && in.coulomb) throw Error("the Coulomb term needs the uniform k-grid (not nkPT read)");
    sh = find_shells(kpt, in.Nk, L, !in.kptread);
    nbr = in.kptread ? neighbours_search(kpt, sh, L) : neighbours_closed_form(in.Nk, sh);
    int nn = 0;
    for (auto &v : sh.bvec) nn += (int)v.size();
    if (flags & 1) {
        std::vector<int32_t> ref = neighbours_bruteforce(kpt, sh, L);
        if (ref != nbr) throw Error("closed-form neighbour table differs from the brute-force search");
    }
    nbr_wb.assign((size_t)nn * 3, 0.);
    bvec_flat.assign((size_t)nn * 3, 0.);
    {
        int n = 0;
        for (size_t s = 0; s < sh.bvec.size(); s++)
            for (auto &v : sh.bvec[s]) {
                for (int c = 0; c < 3; c++) {
                    nbr_wb[(size_t)n * 3 + c] = sh.weights[s] * v[c];
                    bvec_flat[(size_t)n * 3 + c] = v[c];
                }
                n++;
            }
    }

    mark("input, k-grid, neighbours");
    // ---- model tables (save_HUD.cpp:142-210)
    zero_table(H, (size_t)nk * nb * nb);
    zero_table(U, (size_t)nk * nb * nb);
    zero_table(D, (size_t)nk * nb * nb * 3);
    mark("table allocation");
    if (in.iMode == "TB") {
        TBModel tb;
        tb.init(in.TBtype, nch, nb, in.hopping, in.b);
#pragma omp parallel for schedule(static)
        for (int ik = 0; ik < nk; ik++) {
            const double *k = &kpt[(size_t)ik * 3];
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) {
                    H[mi(nb, ik, i, j)] = tb.energy(i, j, k);
                    U[mi(nb, ik, i, j)] = tb.unitary(i, j, k);
                    for (int x = 0; x < 3; x++) D[mi(nb, ik, i, j) * 3 + x] = tb.dipole(i, j, x, k);
                }
        }
    } else if (in.iMode == "CY") {
        if (cy.norb < nb) throw Error("the crystal interface files hold fewer orbitals than bands in the input");
#pragma omp parallel for schedule(dynamic, 64)
        for (int ik = 0; ik < nk; ik++) {
            const double *k = &kpt[(size_t)ik * 3];
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) {
                    H[mi(nb, ik, i, j)] = cy.energy(i, j, k);
                    for (int x = 0; x < 3; x++) D[mi(nb, ik, i, j) * 3 + x] = cy.dipole(i, j, x, k);
                }
            std::vector<double> ev(nb);
            std::vector<cplx> V((size_t)nb * nb);
            eigh_jacobi(nb, &H[mi(nb, ik, 0, 0)], ev.data(), V.data());
            for (int a = 0; a < nb; a++)
                for (int j = 0; j < nb; j++) U[mi(nb, ik, a, j)] = std::conj(V[(size_t)j * nb + a]);
        }
    } else {
        WannierModel w;
        std::string seed = in.ref_coord;
        std::string path = (seed.size() && seed[0] == '/') ? seed + "_tb.dat" : in.base_dir + "/" + seed + "_tb.dat";
        w.read_tb(path);
        mark("_tb.dat read");
        if (w.nw != nb) throw Error("bands in input (" + std::to_string(nb) + ") differ from wannier90 nwannier (" + std::to_string(w.nw) + ")");
        std::vector<double> HR((size_t)w.nR * nb * nb * 2), rR((size_t)w.nR * nb * nb * 6);
        std::memcpy(HR.data(), w.HR.data(), HR.size() * sizeof(double));
        std::memcpy(rR.data(), w.rR.data(), rR.size() * sizeof(double));
        int rc = 0;
        if (in.kptread) {
            // An explicit k list has no per-direction phase tables: the sums of
            // Wannier.h:278-304, 336-471 term by term (e^{2 pi i k.R} per k and R),
            // H with a real diagonal, D = (r + r^T)/2 with a real diagonal.
            const double tp = 2.0 * Units::pi();
#pragma omp parallel for schedule(dynamic, 16)
            for (int ik = 0; ik < nk; ik++) {
                const double *k = &kpt[(size_t)ik * 3];
                std::vector<cplx> h((size_t)nb * nb, 0.), r((size_t)nb * nb * 3, 0.);
                for (int q = 0; q < w.nR; q++) {
                    const int32_t *R = &w.R[(size_t)q * 3];
                    const cplx ph = std::exp(cplx(0.0, tp * (k[0] * R[0] + k[1] * R[1] + k[2] * R[2])));
                    for (int t = 0; t < nb * nb; t++) {
                        h[t] += ph * w.HR[(size_t)q * nb * nb + t];
                        for (int x = 0; x < 3; x++) r[(size_t)t * 3 + x] += ph * w.rR[((size_t)q * nb * nb + t) * 3 + x];
                    }
                }
                for (int i = 0; i < nb; i++)
                    for (int j = 0; j < nb; j++) {
                        cplx v = h[(size_t)i * nb + j];
                        H[mi(nb, ik, i, j)] = (i == j) ? cplx(v.real(), 0.0) : v;
                        for (int x = 0; x < 3; x++) {
                            cplx a = r[((size_t)i * nb + j) * 3 + x], b = r[((size_t)j * nb + i) * 3 + x];
                            if (i == j) { a = cplx(a.real(), 0.0); b = cplx(b.real(), 0.0); }
                            D[mi(nb, ik, i, j) * 3 + x] = 0.5 * (a + b);
                        }
                    }
            }
        } else {
            auto *contract = (flags & 2) ? edus_wannier_hk_fp32 : edus_wannier_hk;
            rc = contract(in.Nk, w.nR, nb, w.R.data(), HR.data(), rR.data(), reinterpret_cast<double *>(H.data()),
                          reinterpret_cast<double *>(D.data()));
            if (rc != 0) throw Error(std::string("Wannier contraction on the GPU failed: ") + edus_last_error());
        }
        mark("Wannier contraction");
        // U from the eigenvectors of H(k) (Wannier.h:496-542): rows are conj(v_a),
        // batched on the GPU like the contraction.
        std::vector<double> ev((size_t)nk * nb);
        rc = edus_eigh_batched(nk, nb, reinterpret_cast<const double *>(H.data()), ev.data(), reinterpret_cast<double *>(U.data()));
        if (rc != 0) throw Error(std::string("eigensolve on the GPU failed: ") + edus_last_error());
    }
    mark("model tables");
    // Scissor: conduction_shift on bands ic >= Nb[1] in the eigenbasis (save_HUD.cpp:147-181).
    if (in.change_gap_constant > 1e-16) {
#pragma omp parallel for schedule(static)
        for (int ik = 0; ik < nk; ik++) {
            std::vector<cplx> add((size_t)nb * nb, 0.);
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) {
                    cplx v = 0.;
                    for (int m = 0; m < nb; m++) {
                        double s = (m >= in.Nb[1]) ? in.change_gap_constant : 0.0;
                        v += std::conj(U[mi(nb, ik, m, i)]) * s * U[mi(nb, ik, m, j)];
                    }
                    add[(size_t)i * nb + j] = v;
                }
            for (int i = 0; i < nb * nb; i++) H[(size_t)ik * nb * nb + i] += add[i];
        }
    }
    if (in.coulomb && !in.coulomb_read) generate_coulomb(disp_at_reference_point());
    // GradientEnergy for Current1 (save_HUD.cpp:236-254): sum_shell,b w*b_x*(H(k+b)-H(k)).
    if (in.iCurrent) {
        zero_table(gradH, (size_t)nk * nb * nb * 3);
#pragma omp parallel for schedule(static)
        for (int ik = 0; ik < nk; ik++)
            for (int x = 0; x < 3; x++)
                for (int i = 0; i < nb; i++)
                    for (int j = 0; j < nb; j++) {
                        cplx acc = 0.;
                        int n = 0;
                        for (size_t s = 0; s < sh.bvec.size(); s++)
                            for (auto &bv : sh.bvec[s]) {
                                int jk = nbr[(size_t)ik * nn + n];
                                acc += sh.weights[s] * bv[x] * (H[mi(nb, jk, i, j)] - H[mi(nb, ik, i, j)]);
                                n++;
                            }
                        gradH[mi(nb, ik, i, j) * 3 + x] = acc;
                    }
    }

    mark("scissor, Coulomb, grad H");
    // ---- decoherence matrix (InitParameters.h:203-224) and its nonzero count
    T.assign((size_t)nb * nb, 0.);
    for (int i = nch; i < nb; i++)
        for (int j = nch; j < nb; j++) T[(size_t)i * nb + j] = (i == j) ? in.T1 : in.T2;
    for (int i = 0; i < nch; i++) {
        for (int j = 0; j < nch; j++) T[(size_t)i * nb + j] = in.Tch;
        for (int j = nch; j < nb; j++) T[(size_t)i * nb + j] = in.Tch / 2.;
    }
    for (int i = 0; i < nb; i++)
        for (int j = i + 1; j < nb; j++) T[(size_t)j * nb + i] = T[(size_t)i * nb + j];
    n_diss = 0;
    for (double t : T)
        if (std::fabs(t) > 1e-24) n_diss++;

    // ---- initial density matrix, stored transposed: P0[a][b] = sum_{i<ncc} U[i][a] conj(U[i][b])
    zero_table(rho0, (size_t)nk * nb * nb);
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < nk; ik++)
        for (int a = 0; a < nb; a++)
            for (int b = 0; b < nb; b++) {
                cplx acc = 0.;
                for (int i = 0; i < ncc; i++) acc += U[mi(nb, ik, i, a)] * std::conj(U[mi(nb, ik, i, b)]);
                rho0[mi(nb, ik, a, b)] = acc;
            }

    mark("T, rho0");
    // ---- stepper and time axis (const_dt_evolution.cpp:164-172, runge_kutta_init.cpp:356-370)
    // The adaptive loop runs Taylor or else RK4 (dyn_dt_evolution.cpp:51-58).
    if (in.taylor) { stepper = EDUS_STEPPER_TAYLOR; order = in.solver_order; }
    else if (in.adams_bashforth && !in.dyn_dt) { stepper = EDUS_STEPPER_AB; order = in.solver_order; }
    else { stepper = EDUS_STEPPER_RK4; order = 4; }
    if (stepper != EDUS_STEPPER_RK4 && (order < 1 || order > EDUS_MAX_ORDER))
        throw Error("solver order out of range 1.." + std::to_string(EDUS_MAX_ORDER));
    if (stepper == EDUS_STEPPER_TAYLOR) {
        if (order < 2) throw Error("TaylorOrder must be >= 2");
        tau_inv = taylor_tau_inverse(std::vector<double>(order, in.dt));  // dt_prev.fill(dt), main_MPI.cpp:366-372
    }
    double t_fin = in.t_fin;
    if (!(t_fin > 0)) {
        // Without t_fin the run lasts 8 sigma (gaussian pump) or the pump's cycles.
        const Pulse &p0 = in.pumps[0];
        double tf = p0.gaussian ? 8 * p0.sigma : p0.ncycle * p0.Period;
        int nstep = int(tf / in.dt);
        t_fin = nstep * in.dt;
        in.t_fin = t_fin;
    }
    t_start = -0.21 / Units::time_au_fs;
    itfi = (int)((t_fin - t_start) / in.dt);
    times.resize((size_t)itfi + 2);
    double t = t_start;
    for (int64_t it = 0; it <= itfi + 1; it++) { times[it] = t; t += in.dt; }
}

// Orbital positions for the Coulomb phases: the real diagonal of the dipole at
// crystal k = (0, 0.3, 0.1) (Fourier_Coulomb.cpp:43-65, "011" orientation).
std::vector<double> Setup::disp_at_reference_point() const {
    const double k[3] = {0.0, 0.3, 0.1};
    std::vector<double> d((size_t)nb * 3, 0.0);
    if (in.iMode == "TB") {
        TBModel tb;
        tb.init(in.TBtype, nch, nb, in.hopping, in.b);
        for (int i = 0; i < nb; i++)
            for (int x = 0; x < 3; x++) d[(size_t)i * 3 + x] = tb.dipole(i, i, x, k).real();
    } else if (in.iMode == "CY") {
        for (int i = 0; i < nb; i++)
            for (int x = 0; x < 3; x++) d[(size_t)i * 3 + x] = cy.dipole(i, i, x, k).real();
    } else {
        WannierModel w;
        std::string seed = in.ref_coord;
        w.read_tb((seed.size() && seed[0] == '/') ? seed + "_tb.dat" : in.base_dir + "/" + seed + "_tb.dat");
        const double twopi = 2 * Units::pi();
        for (int r = 0; r < w.nR; r++) {
            const double ph = twopi * (k[0] * w.R[3 * r] + k[1] * w.R[3 * r + 1] + k[2] * w.R[3 * r + 2]);
            const cplx e(std::cos(ph), std::sin(ph));
            for (int i = 0; i < nb; i++)
                for (int x = 0; x < 3; x++) d[(size_t)i * 3 + x] += (e * w.rR[(((size_t)r * nb + i) * nb + i) * 3 + x]).real();
        }
    }
    return d;
}

// Fourier_Coulomb.cpp:95-500: the G shell |G| < G_distance sqrt(det_b) scanned as
// there (offsets -4999..4999, first offset outer), the reference's fixed grids
// (4999 Simpson points per direction, 39999 for the q ~ 0 correction), the
// coefficients on the GPU, then the coefficient files under Output/Coulomb/ with
// the label G_num<n>_Ncut<ncut><labelInput>.
void Setup::generate_coulomb(const std::vector<double> &disp) {
    ncut = in.Ncut;
    const double gdist = in.G_distance * std::sqrt(det_b);
    std::vector<int32_t> g;
    const int bord = 4999;
    for (int ix = -bord; ix <= bord; ix++)
        for (int iy = -bord; iy <= bord; iy++) {
            const double crys[3] = {0.0, (double)ix, (double)iy};
            double cart[3];
            L.to_cart(crys, cart);
            if (std::sqrt(cart[0] * cart[0] + cart[1] * cart[1] + cart[2] * cart[2]) < gdist) {
                g.push_back(ix);
                g.push_back(iy);
            }
        }
    edus_coulomb_gen p;
    std::memset(&p, 0, sizeof p);
    p.nb = nb;
    p.ncut = ncut;
    p.n_sum = 4999;
    p.n_peak = 39999;
    p.rytova_keldysh = in.rytova_keldysh ? 1 : 0;
    p.n_g = (int32_t)(g.size() / 2);
    p.g = g.data();
    for (int x = 0; x < 3; x++) { p.b1[x] = in.b[1][x]; p.b2[x] = in.b[2][x]; }
    p.eps_area = in.epsilon_static * det_a;
    p.qtf = in.qTF;
    p.r0 = in.r0;
    p.q_peak = (in.qTF < 0.02) ? 0.02 : in.qTF;
    p.disp = disp.data();
    p.verbose = 1;
    const int nb2 = nb * nb, nmm = ncut * ncut;
    cA.assign((size_t)nmm * nb2, 0.);
    cB = cA; cC = cA; cD = cA;
    cscreen.assign(nb2, 0.);
    cvh.assign(nb2, 0.);
    int rc = edus_coulomb_coefficients(&p, reinterpret_cast<double *>(cA.data()), reinterpret_cast<double *>(cB.data()),
                                       reinterpret_cast<double *>(cC.data()), reinterpret_cast<double *>(cD.data()),
                                       reinterpret_cast<double *>(cscreen.data()), reinterpret_cast<double *>(cvh.data()));
    if (rc != 0) throw Error(std::string("Coulomb coefficients on the GPU failed: ") + edus_last_error());
    const std::string dir = in.base_dir + "/Output/Coulomb/";
    ::mkdir((in.base_dir + "/Output").c_str(), 0755);
    ::mkdir(dir.c_str(), 0755);
    const std::string label = "G_num" + std::to_string(p.n_g) + "_Ncut" + std::to_string(ncut) + in.coulomb_label;
    coulomb_label_written = label;
    {   // V_Hartree (Fourier_Coulomb.cpp:197-208)
        std::ofstream f(dir + "V_Hartree_" + label + ".txt");
        if (!f) throw Error("cannot write " + dir);
        for (int ic = 0; ic < nb; ic++) {
            for (int jc = 0; jc < nb; jc++) f << std::setprecision(20) << cvh[(size_t)ic * nb + jc].real() << "    ";
            f << std::endl;
        }
    }
    {   // A..D (Fourier_Coulomb.cpp:341-367)
        std::ofstream fa(dir + "A_coeff" + label + ".txt"), fb(dir + "B_coeff" + label + ".txt");
        std::ofstream fc(dir + "C_coeff" + label + ".txt"), fd(dir + "D_coeff" + label + ".txt");
        std::ofstream *fs[4] = {&fa, &fb, &fc, &fd};
        const std::vector<cplx> *vs[4] = {&cA, &cB, &cC, &cD};
        for (int mm = 0; mm < nmm; mm++) {
            for (int i = 0; i < nb2; i++)
                for (int q = 0; q < 4; q++) {
                    const cplx v = (*vs[q])[(size_t)mm * nb2 + i];
                    *fs[q] << std::setprecision(20) << std::scientific << v.real() << "    " << v.imag() << "    ";
                }
            for (int q = 0; q < 4; q++) *fs[q] << std::endl;
        }
    }
    {   // Screen_const (Fourier_Coulomb.cpp:469-478)
        std::ofstream f(dir + "Screen_const_" + label + ".txt");
        for (int i = 0; i < nb2; i++) f << std::setprecision(20) << cscreen[i].real() << " " << cscreen[i].imag() << std::endl;
    }
}

// Coefficient files of Fourier_Coulomb.cpp:505-622: V_Hartree (nb lines of nb
// reals), A/B/C/D (ncut^2 lines of 2 nb^2 numbers, re im pairs), Screen_const
// (nb^2 lines "re im"), all under Output/Coulomb/ next to the input file.
void Setup::read_coulomb() {
    ncut = in.Ncut;
    const int nb2 = nb * nb, nmm = ncut * ncut;
    const std::string dir = in.base_dir + "/Output/Coulomb/";
    auto numbers = [&](const std::string &path, size_t expect) {
        std::ifstream f(path);
        if (!f) throw Error("cannot open Coulomb coefficient file " + path);
        std::vector<double> v;
        double x;
        while (f >> x) v.push_back(x);
        if (v.size() < expect) throw Error("Coulomb coefficient file " + path + " is too short");
        return v;
    };
    std::vector<double> vh = numbers(dir + "V_Hartree_" + in.coulomb_label + ".txt", (size_t)nb2);
    cvh.resize(nb2);
    for (int i = 0; i < nb2; i++) cvh[i] = cplx(vh[i], 0.0);
    auto table = [&](const char *name, std::vector<cplx> &out) {
        std::vector<double> v = numbers(dir + name + in.coulomb_label + ".txt", (size_t)nmm * nb2 * 2);
        out.resize((size_t)nmm * nb2);
        for (size_t i = 0; i < out.size(); i++) out[i] = cplx(v[2 * i], v[2 * i + 1]);
    };
    table("A_coeff", cA);
    table("B_coeff", cB);
    table("C_coeff", cC);
    table("D_coeff", cD);
    std::vector<double> sc = numbers(dir + "Screen_const_" + in.coulomb_label + ".txt", (size_t)nb2 * 2);
    cscreen.resize(nb2);
    for (int i = 0; i < nb2; i++) cscreen[i] = cplx(sc[2 * i], sc[2 * i + 1]);
}

std::vector<double> taylor_tau_inverse(const std::vector<double> &dt_prev) {
    const int n = (int)dt_prev.size() - 1;
    std::vector<double> ti(n), tau((size_t)n * n);
    ti[0] = dt_prev[0];
    for (int i = 1; i < n; i++) ti[i] = ti[i - 1] + dt_prev[i - 1];
    const double scale = std::pow(10, 7);
    for (int i = 0; i < n; i++)
        for (int p = 0; p < n; p++) tau[(size_t)i * n + p] = scale * std::pow(ti[i], (p + 1));
    std::vector<double> inv = inverse_gauss_jordan(tau, n);
    for (double &v : inv) v *= scale;
    return inv;
}

// Pump sum and half probe at time t (RungeKutta.cpp:24-30).
void Setup::fields_at(double t, double ef[6]) const {
    ef[0] = ef[1] = ef[2] = 0.;
    for (const Pulse &p : in.pumps) {
        double e[3];
        p.field(t, e);
        for (int i = 0; i < 3; i++) ef[i] += e[i];
    }
    double e2[3];
    in.probe.field(t, e2);
    for (int i = 0; i < 3; i++) ef[3 + i] = .5 * e2[i];
}

static bool field_on(const double ef[6]) {
    double s = 0;
    for (int i = 0; i < 6; i++) s += ef[i] * ef[i];
    return std::sqrt(s) > 1e-20;
}

// Fields of step `it` and its gradient flag: RK4 evaluates at t, t+dt/2 and
// t+dt and gates the gradient with the previous step's last field
// (RungeKutta.cpp:20-30,44-61,113-128,151-162; initially true, main_MPI.cpp:274);
// the multistep methods evaluate at t and gate with it.
void Setup::step_fields(int64_t it, double *f, uint8_t *grad_on) const {
    const double tl = times[it];
    if (stepper == EDUS_STEPPER_RK4) {
        const double dt2 = in.dt / 2.0;
        fields_at(tl, f);
        fields_at(tl + dt2, f + 6);
        fields_at(tl + in.dt, f + 12);
        if (it == 0) *grad_on = 1;
        else {
            double prev[6];
            fields_at(times[it - 1] + in.dt, prev);
            *grad_on = field_on(prev) ? 1 : 0;
        }
    } else {
        fields_at(tl, f);
        *grad_on = field_on(f) ? 1 : 0;
    }
}

}  // namespace edus
