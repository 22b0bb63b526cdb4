// extern "C" surface of libedus_host.so (include/edus_host.h).
#include "../../../include/edus_host.h"

#include <cstdio>
#include <cstring>
#include <string>

#include "host.hpp"

struct edus_setup {
    edus::Setup s;
    edus_coulomb coul;
};

namespace edus {
int64_t run_loop(const Setup &s, const std::string &outdir, int64_t max_steps, int verbose);

void fill_problem(const Setup &s, edus_problem &p) {
    std::memset(&p, 0, sizeof p);
    p.nk = s.nk;
    p.nb = s.nb;
    p.nch = s.nch;
    p.ncc = s.ncc;
    p.n_nbr = (int32_t)(s.nbr_wb.size() / 3);  // b vectors over all shells (grid or list)
    p.stepper = s.stepper;
    p.order = s.order;
    p.nbr = s.nbr.data();
    p.nbr_wb = s.nbr_wb.data();
    p.H = reinterpret_cast<const double *>(s.H.data());
    p.D = reinterpret_cast<const double *>(s.D.data());
    p.U = reinterpret_cast<const double *>(s.U.data());
    p.T = s.T.data();
    p.rho0 = reinterpret_cast<const double *>(s.rho0.data());
    p.gradH = s.gradH.empty() ? nullptr : reinterpret_cast<const double *>(s.gradH.data());
    p.wx2 = s.in.probe.wl;
    p.dt = s.in.dt;
    double dk[3] = {1. / s.in.Nk[0], 1. / s.in.Nk[1], 1. / s.in.Nk[2]};
    p.j_norm = dk[0] * dk[1] * dk[2] / s.det_a;
    p.tabs_norm = std::fabs(s.L.jacobian) * dk[0] * dk[1] * dk[2];
    p.nktot = (double)s.nk;
    p.taylor_tau_inv = s.tau_inv.empty() ? nullptr : s.tau_inv.data();
    // An explicit k list (nkPT read) is no grid: no strip order over it.
    for (int i = 0; i < 3; i++) p.nk_dims[i] = s.in.kptread ? (i == 0 ? s.nk : 1) : s.in.Nk[i];
}

void fill_coulomb(const Setup &s, edus_coulomb &c) {
    std::memset(&c, 0, sizeof c);
    c.ncut = s.ncut;
    c.A = reinterpret_cast<const double *>(s.cA.data());
    c.B = reinterpret_cast<const double *>(s.cB.data());
    c.C = reinterpret_cast<const double *>(s.cC.data());
    c.D = reinterpret_cast<const double *>(s.cD.data());
    c.screen = reinterpret_cast<const double *>(s.cscreen.data());
    c.v_hartree = reinterpret_cast<const double *>(s.cvh.data());
    c.diag_basis = s.in.coulomb_diag_basis ? 1 : 0;
}
}  // namespace edus

static thread_local std::string g_err;

#define GUARD_BEGIN try {
#define GUARD_END(ret)                      \
    }                                       \
    catch (const std::exception &e) {       \
        g_err = e.what();                   \
        return ret;                         \
    }                                       \
    catch (...) {                           \
        g_err = "unknown host error";       \
        return ret;                         \
    }

extern "C" {

const char *edus_host_last_error(void) { return g_err.c_str(); }

int edus_setup_from_input(const char *input_path, int32_t flags, edus_setup **out) {
    GUARD_BEGIN
    if (!input_path || !out) { g_err = "null argument"; return EDUS_ERR_INVALID; }
    auto *h = new edus_setup();
    try {
        h->s.build(input_path, flags);
        if (h->s.in.coulomb) edus::fill_coulomb(h->s, h->coul);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    g_err.clear();
    return 0;
    GUARD_END(EDUS_ERR_INVALID)
}

void edus_setup_free(edus_setup *s) { delete s; }

int edus_setup_get_info(const edus_setup *h, edus_setup_info *info) {
    GUARD_BEGIN
    if (!h || !info) return EDUS_ERR_INVALID;
    const edus::Setup &s = h->s;
    std::memset(info, 0, sizeof *info);
    info->nk0 = s.in.Nk[0]; info->nk1 = s.in.Nk[1]; info->nk2 = s.in.Nk[2]; info->nk = s.nk;
    info->nb = s.nb; info->nch = s.nch; info->ncc = s.ncc;
    for (int i = 0; i < 3; i++) info->nbands[i] = s.in.Nb[i];
    info->n_nbr = (int32_t)(s.nbr_wb.size() / 3);
    info->n_shells = (int32_t)s.sh.bvec.size();
    info->stepper = s.stepper; info->order = s.order;
    info->it_resolution = s.in.it_resolution; info->i_current = s.in.iCurrent; info->i_tabs = s.in.iTAbs;
    info->n_pumps = (int32_t)s.in.pumps.size();
    info->model = (s.in.iMode == "W") ? 1 : (s.in.iMode == "CY") ? 2 : 0;
    info->n_diss = s.n_diss;
    info->itfi = s.itfi;
    info->dt = s.in.dt; info->t_fin = s.in.t_fin; info->t_start = s.t_start;
    info->det_a = s.det_a; info->det_b = s.det_b; info->jacobian = s.L.jacobian;
    edus_problem p;
    edus::fill_problem(s, p);
    info->j_norm = p.j_norm; info->tabs_norm = p.tabs_norm; info->nktot = p.nktot; info->wx2 = p.wx2;
    info->T1 = s.in.T1; info->T2 = s.in.T2; info->Tch = s.in.Tch;
    info->shell_residual = s.sh.residual;
    info->ncut = s.in.coulomb ? s.ncut : 0;
    info->dyn_dt = s.in.dyn_dt ? 1 : 0;
    info->eps_step = s.in.epsStepAbs;
    info->print_flags = (s.in.print_population ? 1 : 0) | (s.in.print_band_dispersion ? 2 : 0);
    info->print_start = s.in.print_start; info->print_end = s.in.print_end; info->print_step = s.in.print_step;
    return 0;
    GUARD_END(EDUS_ERR_INVALID)
}

const double *edus_setup_array(const edus_setup *h, const char *name, int64_t *count) {
    if (!h || !name) return nullptr;
    const edus::Setup &s = h->s;
    auto ret = [&](const double *p, size_t n) { if (count) *count = (int64_t)n; return p; };
    auto retc = [&](const auto &v) { return ret(reinterpret_cast<const double *>(v.data()), v.size() * 2); };
    std::string n = name;
    if (n == "kpt") return ret(s.kpt.data(), s.kpt.size());
    if (n == "a") return ret(&s.in.a[0][0], 9);
    if (n == "b") return ret(&s.in.b[0][0], 9);
    if (n == "kcart") return ret(s.kcart.data(), s.kcart.size());
    if (n == "H") return retc(s.H);
    if (n == "U") return retc(s.U);
    if (n == "D") return retc(s.D);
    if (n == "rho0") return retc(s.rho0);
    if (n == "gradH") return retc(s.gradH);
    if (n == "T") return ret(s.T.data(), s.T.size());
    if (n == "nbr_wb") return ret(s.nbr_wb.data(), s.nbr_wb.size());
    if (n == "bvec") return ret(s.bvec_flat.data(), s.bvec_flat.size());
    if (n == "weights") return ret(s.sh.weights.data(), s.sh.weights.size());
    if (n == "taylor_tau_inv") return ret(s.tau_inv.data(), s.tau_inv.size());
    if (n == "times") return ret(s.times.data(), s.times.size());
    if (n == "coul_A") return retc(s.cA);
    if (n == "coul_B") return retc(s.cB);
    if (n == "coul_C") return retc(s.cC);
    if (n == "coul_D") return retc(s.cD);
    if (n == "coul_screen") return retc(s.cscreen);
    if (n == "coul_vh") return retc(s.cvh);
    if (count) *count = 0;
    return nullptr;
}

const int32_t *edus_setup_nbr(const edus_setup *h, int64_t *count) {
    if (!h) return nullptr;
    if (count) *count = (int64_t)h->s.nbr.size();
    return h->s.nbr.data();
}

int edus_setup_pulse(const edus_setup *h, int32_t index, edus_pulse *out) {
    if (!h || !out) return EDUS_ERR_INVALID;
    const edus::Setup &s = h->s;
    const edus::Pulse *p;
    if (index >= 0 && index < (int)s.in.pumps.size()) p = &s.in.pumps[index];
    else if (index == (int)s.in.pumps.size()) p = &s.in.probe;
    else return EDUS_ERR_INVALID;
    out->gaussian = p->gaussian; out->sin2 = p->sin2; out->pw = p->PW; out->envelope = p->envelope;
    out->wl = p->wl; out->period = p->Period; out->intensity = p->Intensity; out->A0 = p->A0; out->E0 = p->E0;
    out->t0 = p->t0; out->tf = p->tf; out->w1 = p->w1; out->w2 = p->w2; out->ncycle = p->ncycle;
    out->sigma = p->sigma; out->phase = p->phase;
    for (int i = 0; i < 3; i++) { out->pol1[i] = p->pol1[i]; out->pol2[i] = p->pol2[i]; }
    return 0;
}

int edus_setup_model(const edus_setup *h, char *name, int32_t len, double *params) {
    GUARD_BEGIN
    if (!h || !name || len < 1) return EDUS_ERR_INVALID;
    const edus::Setup &s = h->s;
    if (s.in.iMode == "W") {
        std::snprintf(name, (size_t)len, "%s", s.in.ref_coord.c_str());
        return 1;
    }
    if (s.in.iMode == "CY") {
        std::snprintf(name, (size_t)len, "%s", "crystal");
        return 1;
    }
    edus::TBModel tb;
    tb.init(s.in.TBtype, s.nch, s.nb, s.in.hopping, s.in.b);
    std::snprintf(name, (size_t)len, "%s", tb.name.c_str());
    if (params) {
        const double v[8] = {tb.a, tb.g, tb.b, tb.t2, tb.phi0, tb.core, tb.rchv, tb.rchc};
        std::memcpy(params, v, sizeof v);
    }
    return 0;
    GUARD_END(EDUS_ERR_INVALID)
}

int edus_setup_eval_tb(const edus_setup *h, const double *kcrys, int32_t n, double *H, double *U, double *D) {
    GUARD_BEGIN
    if (!h || !kcrys || n < 0) return EDUS_ERR_INVALID;
    const edus::Setup &s = h->s;
    if (s.in.iMode != "TB") { g_err = "not a tight-binding setup"; return EDUS_ERR_INVALID; }
    edus::TBModel tb;
    tb.init(s.in.TBtype, s.nch, s.nb, s.in.hopping, s.in.b);
    const int nb = s.nb;
    for (int q = 0; q < n; q++)
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++) {
                const size_t o = ((size_t)q * nb + i) * nb + j;
                edus::cplx hv = tb.energy(i, j, kcrys + 3 * q), uv = tb.unitary(i, j, kcrys + 3 * q);
                if (H) { H[2 * o] = hv.real(); H[2 * o + 1] = hv.imag(); }
                if (U) { U[2 * o] = uv.real(); U[2 * o + 1] = uv.imag(); }
                if (D)
                    for (int x = 0; x < 3; x++) {
                        edus::cplx dv = tb.dipole(i, j, x, kcrys + 3 * q);
                        D[(o * 3 + x) * 2] = dv.real();
                        D[(o * 3 + x) * 2 + 1] = dv.imag();
                    }
            }
    return 0;
    GUARD_END(EDUS_ERR_INVALID)
}

int edus_setup_problem(const edus_setup *h, edus_problem *p) {
    GUARD_BEGIN
    if (!h || !p) return EDUS_ERR_INVALID;
    edus::fill_problem(h->s, *p);
    if (h->s.in.coulomb) p->coulomb = &h->coul;
    return 0;
    GUARD_END(EDUS_ERR_INVALID)
}

int edus_setup_fields(const edus_setup *h, int64_t it0, int64_t nsteps, double *fields, uint8_t *grad_on) {
    GUARD_BEGIN
    if (!h || !fields || !grad_on || it0 < 0 || nsteps < 0) return EDUS_ERR_INVALID;
    const edus::Setup &s = h->s;
    if (it0 + nsteps > s.itfi + 1) { g_err = "step range beyond the last step"; return EDUS_ERR_INVALID; }
    const int nfe = s.fields_per_step();
    for (int64_t j = 0; j < nsteps; j++) s.step_fields(it0 + j, fields + (size_t)j * nfe * 6, grad_on + j);
    return 0;
    GUARD_END(EDUS_ERR_INVALID)
}

double edus_setup_time(const edus_setup *h, int64_t it) {
    if (!h || it < 0 || it >= (int64_t)h->s.times.size()) return 0.0;
    return h->s.times[it];
}

int64_t edus_run(const edus_setup *h, const char *outdir, int64_t max_steps, int32_t verbose) {
    GUARD_BEGIN
    if (!h || !outdir) return EDUS_ERR_INVALID;
    return edus::run_loop(h->s, outdir, max_steps, verbose);
    GUARD_END(EDUS_ERR_INVALID)
}

int edus_write_synthetic_tb(const char *path, int32_t nw, int32_t rmax, double a_ang, double sigma_h_ev,
                            double sigma_r_ang, double decay_ang, uint64_t seed) {
    GUARD_BEGIN
    if (!path || nw < 1 || rmax < 0) return EDUS_ERR_INVALID;
    edus::write_synthetic_tb(path, nw, rmax, a_ang, sigma_h_ev, sigma_r_ang, decay_ang, seed);
    return 0;
    GUARD_END(EDUS_ERR_INVALID)
}

}  // extern "C"
