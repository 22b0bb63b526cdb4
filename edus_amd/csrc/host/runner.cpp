// Time loop and observable files of the drop-in.
//
// The loop is Source_Main/const_dt_evolution.cpp:164-217: steps it = 0..itfi,
// observables printed *before* the step whenever it % it_resolution == 0 with
// the field left by the previous step's last evaluation (zero at it = 0), then
// time_loop += dt.  The writers reproduce the stream formats of
// Observables_MPI.h (PrintEF 134-140, PrintLossesMPI_exciton 394-420,
// Current1 705-713, Current2 764-772, PrintTransientAbsMPI 584-596) and
// print_k_MPI (492-554).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "../../../include/edus.h"
#include "host.hpp"

namespace edus {

namespace {
void mkdirs(const std::string &p) {
    std::string cur;
    for (size_t i = 0; i < p.size(); i++) {
        cur += p[i];
        if (p[i] == '/' || i + 1 == p.size()) ::mkdir(cur.c_str(), 0755);
    }
}

struct Engine {
    edus_ctx *ctx = nullptr;
    ~Engine() { if (ctx) edus_destroy(ctx); }
};
}  // namespace

void fill_problem(const Setup &s, edus_problem &p);
void fill_coulomb(const Setup &s, edus_coulomb &c);

// The observable files of one run, written in the reference's stream formats.
struct Writers {
    const Setup &s;
    std::string od;
    std::ofstream fE, fL, fJ1, fJ2, fT;
    bool k_printed = false;
    std::vector<double> pop;
    double n_cond = 0.0;      // largest printed population (Observables_MPI.h:413-420)
    double p_cond_max = 0.0;  // largest excited population over k (Observables_MPI.h:330-347)
    // k-resolved snapshots (if_resolution.cpp:19-35, 232-273)
    std::ofstream fTEk;
    int icont = 0;
    std::vector<int> print_order;  // Create_sorted_indices_k_MPI (Observables_MPI.h:617-657)

    Writers(const Setup &st, const std::string &outdir) : s(st), od(outdir + "/Output/"), pop(st.nb) {
        mkdirs(od);
        fE.open(od + "EF.txt");
        fL.open(od + "Losses.txt");
        if (s.in.iCurrent) { fJ1.open(od + "J1.txt"); fJ2.open(od + "J2.txt"); }
        if (s.in.iTAbs) fT.open(od + "TransientAbs.txt");
        if (!fE || !fL) throw Error("cannot open output files under " + od);
        if (s.in.print_band_dispersion || s.in.print_population) {
            fTEk.open(od + "t_Ek.txt");
            if (s.in.print_population && s.ncc >= s.nb)
                throw Error("PrintPopulation needs a conduction band (valence/conduction snapshot elements)");
            sorted_order();
        }
    }

    // Rows of the grid sorted by one crystal coordinate, as the reference orders
    // the k-resolved files (Observables_MPI.h:617-657 with 38-60).
    void sorted_order() {
        const int *Nk = s.in.Nk;
        int nk_x = Nk[1] * Nk[0], nk_y = s.nk / (Nk[1] * Nk[0]), pol = 1;
        if (Nk[1] > Nk[2]) {
            pol = 0;
            nk_y = s.nk / (Nk[1] * Nk[2]);
            nk_x = Nk[1] * Nk[2];
        }
        std::vector<double> kt(nk_y, 0.0);
        for (int ik = 0; ik < s.nk; ik++)
            if (ik % nk_x == 0 && ik / nk_x < nk_y) kt[ik / nk_x] = s.kpt[(size_t)ik * 3 + pol];
        std::vector<int> kidx(nk_y);
        for (int r = 0; r < nk_y; r++) {
            int best = 0;
            double mn = 1e15;
            for (int i = 0; i < nk_y; i++)
                if (kt[i] < mn) { mn = kt[i]; best = i; }
            kt[best] = 1e15;
            kidx[r] = best;
        }
        print_order.resize(s.nk);
        for (int ik = 0; ik < s.nk; ik++) print_order[ik] = kidx[ik / nk_x] * nk_x + ik % nk_x;
    }

    void snapshot(edus_ctx *ctx, const char *label, int ic, int jc) {
        std::vector<double> v((size_t)2 * s.nk);
        if (edus_eigen_element(ctx, ic, jc, v.data()) != 0) throw Error(std::string("edus_eigen_element: ") + edus_last_error());
        std::ofstream f(od + label + "_" + std::to_string(ic) + std::to_string(jc) + "_" + std::to_string(icont) + ".txt");
        for (int ik = 0; ik < s.nk; ik++) {
            const int q = print_order[ik];
            f << v[2 * (size_t)q] << " " << v[2 * (size_t)q + 1] << '\n';
        }
    }

    void observe(edus_ctx *ctx, double tl, const double ef_last[6], int verbose) {
        const double tfs = Units::time_au_fs;
        double j1[6], j2[6], ta[6], pm[2];
        fE << std::setw(15) << std::setprecision(8) << tl * tfs;
        for (int i = 0; i < 3; i++) fE << std::setw(15) << std::setprecision(8) << ef_last[i];
        fE << '\n';
        bool print_k = false;
        if ((s.in.print_band_dispersion || s.in.print_population) && tl > s.in.print_start + icont * s.in.print_step &&
            tl < s.in.print_end) {
            print_k = true;
            icont++;
            fTEk << tl * tfs << '\n';
        }
        int rc = edus_observables(ctx, pop.data(), j1, j2, ta, pm);
        if (rc != 0) throw Error(std::string("edus_observables: ") + edus_last_error());
        double trace = -s.ncc;
        n_cond = 0.0;
        fL << std::setprecision(16) << tl * tfs << "   ";
        for (int ic = 0; ic < s.nb; ic++) {
            double v = pop[ic];
            trace += v;
            if (ic < s.ncc) v = 1.0 - v;
            if (v > n_cond) n_cond = v;
            fL << std::setprecision(16) << v << "   ";
        }
        fL << '\n';
        p_cond_max = pm[1];
        if (verbose && std::fabs(trace) > 1e-13) std::cout << "Preservation of particle num fail P_trace = " << trace << std::endl;
        if (!k_printed) {
            std::ofstream fk(od + "k_grid.txt");
            for (int ik = 0; ik < s.nk; ik++)
                fk << s.kpt[(size_t)ik * 3] << " " << s.kpt[(size_t)ik * 3 + 1] << " " << s.kpt[(size_t)ik * 3 + 2] << '\n';
            k_printed = true;
        }
        if (s.in.iCurrent) {
            fJ1 << std::setw(15) << std::setprecision(8) << tl * tfs;
            for (int x = 0; x < 3; x++) {
                fJ1 << std::setw(25) << std::setprecision(15) << std::scientific << j1[2 * x];
                fJ1 << std::setw(25) << std::setprecision(15) << std::scientific << j1[2 * x + 1];
            }
            fJ1 << '\n';
            fJ2 << std::setw(15) << std::setprecision(8) << tl * tfs;
            for (int x = 0; x < 3; x++) fJ2 << std::setw(25) << std::setprecision(15) << std::scientific << -2. * j2[2 * x + 1];
            fJ2 << '\n';
        }
        if (s.in.iTAbs) {
            fT << std::setw(15) << std::setprecision(8) << tl * tfs;
            for (int i = 0; i < 3; i++) fT << std::setw(15) << std::setprecision(8) << ef_last[3 + i];
            for (int x = 0; x < 3; x++) {
                fT << std::setw(15) << std::setprecision(8) << ta[2 * x] / (double)s.nk;
                fT << std::setw(15) << std::setprecision(8) << ta[2 * x + 1] / (double)s.nk;
            }
            fT << '\n';
        }
        if (print_k && s.in.coulomb) {  // Edia_<icont>.txt (if_resolution.cpp:40-55, Observables_MPI.h:440-477)
            std::vector<double> ek((size_t)s.nk * s.nb);
            if (edus_renorm_energies(ctx, ek.data()) != 0) throw Error(std::string("edus_renorm_energies: ") + edus_last_error());
            std::ofstream f(od + "Edia_" + std::to_string(icont) + ".txt");
            for (int ik = 0; ik < s.nk; ik++) {
                for (int ic = 0; ic < s.nb; ic++) f << std::setprecision(20) << ek[(size_t)ik * s.nb + ic] << " ";
                f << '\n';
            }
        }
        if (print_k && s.in.print_population) {
            snapshot(ctx, "P_val_cond_in_Eigenstate", s.ncc - 1, s.ncc);
            snapshot(ctx, "P_val_in_Eigenstate", s.ncc - 1, s.ncc - 1);
            snapshot(ctx, "P_cond_in_Eigenstate", s.ncc, s.ncc);
        }
    }
};

static void make_engine(const Setup &s, Engine &eng) {
    edus_problem p;
    fill_problem(s, p);
    edus_coulomb coul;
    if (s.in.coulomb) {
        fill_coulomb(s, coul);
        p.coulomb = &coul;
    }
    if (edus_create(&p, &eng.ctx) != 0) throw Error(std::string("edus_create: ") + edus_last_error());
}

// Constant dt (Source_Main/const_dt_evolution.cpp).
static int64_t run_const(const Setup &s, const std::string &outdir, int64_t max_steps, int verbose) {
    const double tfs = Units::time_au_fs;
    Writers w(s, outdir);
    Engine eng;
    make_engine(s, eng);
    const int nfe = s.fields_per_step();
    const int64_t last = (max_steps >= 0) ? std::min<int64_t>(s.itfi, max_steps - 1) : s.itfi;
    const int res = std::max(1, s.in.it_resolution);
    double ef_last[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> fields;
    std::vector<uint8_t> flags;
    int64_t it = 0;
    while (it <= last) {
        const double tl = s.times[it];
        if (it % res == 0) {
            if (verbose && it % (10 * res) == 0)
                std::cout << " it: " << it << " time: " << tl * tfs << " fs, time fin: " << s.in.t_fin * tfs << " fs" << std::endl;
            w.observe(eng.ctx, tl, ef_last, verbose);
        }
        // Steps up to the next observation point, as one engine call.
        int64_t next_obs = (it / res + 1) * res;
        int64_t n = std::min(next_obs, last + 1) - it;
        fields.resize((size_t)n * nfe * 6);
        flags.resize(n);
        for (int64_t j = 0; j < n; j++) s.step_fields(it + j, &fields[(size_t)j * nfe * 6], &flags[j]);
        int rc;
        if (s.stepper == EDUS_STEPPER_RK4) rc = edus_step_rk4(eng.ctx, (int32_t)n, fields.data(), flags.data());
        else if (s.stepper == EDUS_STEPPER_AB) rc = edus_step_ab(eng.ctx, (int32_t)n, fields.data(), flags.data());
        else rc = edus_step_taylor(eng.ctx, (int32_t)n, fields.data(), flags.data());
        if (rc != 0) throw Error(std::string("edus_step: ") + edus_last_error());
        const double *lastf = &fields[((size_t)(n - 1) * nfe + (nfe - 1)) * 6];
        for (int i = 0; i < 6; i++) ef_last[i] = lastf[i];
        it += n;
    }
    return last + 1;
}

// Adaptive dt (Source_Main/dyn_dt_evolution.cpp:1-183): while time_loop < t_fin,
// observables every it_resolution steps, one step of the current dt, the
// step-size measure, then dt *= eps/max_dP when max_dP > eps and dt > 0.01, else
// dt *= 1.02 while dt < dt0; Taylor shifts its dt history; time_loop += the new dt.
static int64_t run_dyn(const Setup &s, const std::string &outdir, int64_t max_steps, int verbose) {
    const double tfs = Units::time_au_fs;
    Writers w(s, outdir);
    Engine eng;
    make_engine(s, eng);
    const bool taylor = s.stepper == EDUS_STEPPER_TAYLOR;
    const int res = std::max(1, s.in.it_resolution);
    const double eps = s.in.epsStepAbs, dt0 = s.in.dt;
    double dt = dt0;
    std::vector<double> dt_prev(taylor ? s.order : 0, dt), tinv;
    double ef_last[6] = {0, 0, 0, 0, 0, 0};
    double f[18];
    bool E_non0 = true;  // main_MPI.cpp:274
    double time_loop = s.t_start;
    int64_t it = 0;
    while (time_loop < s.in.t_fin && (max_steps < 0 || it < max_steps)) {
        if (it % res == 0) w.observe(eng.ctx, time_loop, ef_last, verbose);
        uint8_t on;
        if (taylor) {
            s.fields_at(time_loop, f);
            double e2 = 0;
            for (int i = 0; i < 6; i++) e2 += f[i] * f[i];
            on = std::sqrt(e2) > 1e-20;
            tinv = taylor_tau_inverse(dt_prev);
            for (int i = 0; i < 6; i++) ef_last[i] = f[i];
        } else {
            s.fields_at(time_loop, f);
            s.fields_at(time_loop + dt / 2.0, f + 6);
            s.fields_at(time_loop + dt, f + 12);
            on = E_non0;
            double e2 = 0;
            for (int i = 0; i < 6; i++) e2 += f[12 + i] * f[12 + i];
            E_non0 = std::sqrt(e2) > 1e-20;
            for (int i = 0; i < 6; i++) ef_last[i] = f[12 + i];
        }
        double max_dp = 0;
        int rc = edus_step_dyn(eng.ctx, dt, f, on, taylor ? tinv.data() : nullptr, w.p_cond_max + w.n_cond, &max_dp);
        if (rc != 0) throw Error(std::string("edus_step_dyn: ") + edus_last_error());
        if (verbose && max_dp > 1.01 * eps)
            std::cout << "!!!!!!!! dt too large  max_dP = " << max_dp << " dtDynamic: " << dt << " a.u." << std::endl;
        if (max_dp > eps && dt > 0.01) dt *= eps / max_dp;
        else if (dt < dt0) dt *= 1.02;
        if (taylor) {
            for (size_t i = dt_prev.size() - 1; i > 0; i--) dt_prev[i] = dt_prev[i - 1];
            dt_prev[0] = dt;
        }
        if (verbose && it % (10 * res) == 0)
            std::cout << " dtDynamic: " << dt << " a.u. | max_dP = " << max_dp << " time_loop fs: " << time_loop * tfs << std::endl;
        time_loop += dt;
        it++;
    }
    return it;
}

int64_t run_loop(const Setup &s, const std::string &outdir, int64_t max_steps, int verbose) {
    return s.in.dyn_dt ? run_dyn(s, outdir, max_steps, verbose) : run_const(s, outdir, max_steps, verbose);
}

}  // namespace edus
