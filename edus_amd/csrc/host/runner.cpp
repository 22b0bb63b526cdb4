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

int64_t run_loop(const Setup &s, const std::string &outdir, int64_t max_steps, int verbose) {
    const double tfs = Units::time_au_fs;
    std::string od = outdir + "/Output/";
    mkdirs(od);
    std::ofstream fE(od + "EF.txt"), fL(od + "Losses.txt"), fJ1, fJ2, fT;
    if (s.in.iCurrent) { fJ1.open(od + "J1.txt"); fJ2.open(od + "J2.txt"); }
    if (s.in.iTAbs) fT.open(od + "TransientAbs.txt");
    if (!fE || !fL) throw Error("cannot open output files under " + od);

    edus_problem p;
    fill_problem(s, p);
    edus_coulomb coul;
    if (s.in.coulomb) {
        fill_coulomb(s, coul);
        p.coulomb = &coul;
    }
    Engine eng;
    if (edus_create(&p, &eng.ctx) != 0) throw Error(std::string("edus_create: ") + edus_last_error());

    const int nfe = s.fields_per_step();
    const int64_t last = (max_steps >= 0) ? std::min<int64_t>(s.itfi, max_steps - 1) : s.itfi;
    const int res = std::max(1, s.in.it_resolution);
    double ef_last[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> pop(s.nb);
    double j1[6], j2[6], ta[6], pm[2];
    std::vector<double> fields;
    std::vector<uint8_t> flags;
    bool k_printed = false;
    const double nktot_d = (double)s.nk;

    int64_t it = 0;
    while (it <= last) {
        const double tl = s.times[it];
        if (it % res == 0) {
            if (verbose && it % (10 * res) == 0)
                std::cout << " it: " << it << " time: " << tl * tfs << " fs, time fin: " << s.in.t_fin * tfs << " fs" << std::endl;
            fE << std::setw(15) << std::setprecision(8) << tl * tfs;
            for (int i = 0; i < 3; i++) fE << std::setw(15) << std::setprecision(8) << ef_last[i];
            fE << std::endl;
            int rc = edus_observables(eng.ctx, pop.data(), j1, j2, ta, pm);
            if (rc != 0) throw Error(std::string("edus_observables: ") + edus_last_error());
            double trace = -s.ncc;
            fL << std::setprecision(16) << tl * tfs << "   ";
            for (int ic = 0; ic < s.nb; ic++) {
                double v = pop[ic];
                trace += v;
                if (ic < s.ncc) v = 1.0 - v;
                fL << std::setprecision(16) << v << "   ";
            }
            fL << std::endl;
            if (verbose && std::fabs(trace) > 1e-13) std::cout << "Preservation of particle num fail P_trace = " << trace << std::endl;
            if (!k_printed) {
                std::ofstream fk(od + "k_grid.txt");
                for (int ik = 0; ik < s.nk; ik++)
                    fk << s.kpt[(size_t)ik * 3] << " " << s.kpt[(size_t)ik * 3 + 1] << " " << s.kpt[(size_t)ik * 3 + 2] << std::endl;
                k_printed = true;
            }
            if (s.in.iCurrent) {
                fJ1 << std::setw(15) << std::setprecision(8) << tl * tfs;
                for (int x = 0; x < 3; x++) {
                    fJ1 << std::setw(25) << std::setprecision(15) << std::scientific << j1[2 * x];
                    fJ1 << std::setw(25) << std::setprecision(15) << std::scientific << j1[2 * x + 1];
                }
                fJ1 << std::endl;
                fJ2 << std::setw(15) << std::setprecision(8) << tl * tfs;
                for (int x = 0; x < 3; x++) fJ2 << std::setw(25) << std::setprecision(15) << std::scientific << -2. * j2[2 * x + 1];
                fJ2 << std::endl;
            }
            if (s.in.iTAbs) {
                fT << std::setw(15) << std::setprecision(8) << tl * tfs;
                for (int i = 0; i < 3; i++) fT << std::setw(15) << std::setprecision(8) << ef_last[3 + i];
                for (int x = 0; x < 3; x++) {
                    fT << std::setw(15) << std::setprecision(8) << ta[2 * x] / nktot_d;
                    fT << std::setw(15) << std::setprecision(8) << ta[2 * x + 1] / nktot_d;
                }
                fT << std::endl;
            }
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

}  // namespace edus
