// Band models feeding the H, U, D tables.
//
// Tight binding: the hexagonal boron-nitride family of headers/TightBinding.h
// (init 81-89, set_hopping 153-173, energy 177-210 / 271-290 / 387-440,
// unitary 894-1004, dipoles 1256-1421).  Each matrix element is the model's
// closed form; the expressions keep the reference's operand order so the values
// agree to the last bit where libm agrees.
// Wannier: reader of wannier90 <seed>_tb.dat as in headers/Wannier.h:47-149 and a
// seeded synthetic generator writing the same format.
#include "host.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace edus {

namespace {
const cplx I1(0.0, 1.0);

bool is_haldane(const std::string &n) { return n == "Haldane_CoreBN_Nedge" || n == "Haldane_CoreBN_Bedge"; }
bool is_corebn(const std::string &n) { return n == "CoreBN_Nedge" || n == "CoreBN_Bedge"; }
}  // namespace

void TBModel::init(const std::string &nm, int nch, int ncv, const std::vector<double> &hopping) {
    name = nm;
    Nch = nch;
    Ncv = ncv;
    if (!supported()) throw Error("TightBinding type " + nm + " is not supported by this build");
    // Boron-nitride family defaults (TightBinding.h:81-89).
    a = 2.5 * Units::space_A_au();
    b = 3.16 * Units::energy_eV_au();
    c = 0;
    d = 0;
    g = 6.2 * Units::energy_eV_au();
    if (is_haldane(name)) {
        if (hopping.size() < 4) throw Error("missing Haldane hopping parameters");
        g = hopping[0]; b = hopping[1]; t2 = hopping[2]; phi0 = hopping[3];
    }
    if (name == "gen2d_hexagonal") {
        if (hopping.size() < 8) throw Error("missing gen2d_hexagonal parameters");
        a = hopping[0]; g = hopping[1]; b = hopping[2]; t2 = hopping[3];
        phi0 = hopping[4]; core = hopping[5]; rchv = hopping[6]; rchc = hopping[7];
    }
}

bool TBModel::supported() const {
    return name == "BoronNitride" || is_corebn(name) || is_haldane(name) || name == "gen2d_hexagonal";
}

cplx TBModel::energy(int ic, int jc, const double k[3]) const {
    const double pi = Units::pi();
    if (is_corebn(name)) {
        cplx f = 1. + std::exp(2. * pi * I1 * k[1]) + std::exp(2. * pi * I1 * k[2]);
        if (ic == 0) return (jc == 0) ? cplx(name == "CoreBN_Nedge" ? -15.06 : -6.91) : cplx(0.);
        if (ic == 1) return (jc == 1) ? cplx(-g / 2.) : (jc == 2 ? b * std::conj(f) : cplx(0.));
        if (ic == 2) return (jc == 1) ? b * f : (jc == 2 ? cplx(g / 2.) : cplx(0.));
        return 0.;
    }
    if (name == "BoronNitride") {
        if (ic == jc) return ic == 0 ? cplx(-g / 2) : (ic == 1 ? cplx(g / 2) : cplx(0));
        cplx f = std::exp(I1 * 2. * pi * k[1]) + 1. + std::exp(I1 * 2. * pi * k[2]);
        return ic == 0 ? std::conj(f) * b : f * b;
    }
    // Haldane_CoreBN_* and gen2d_hexagonal: core level, then a two-band Haldane
    // block with nearest (b) and next-nearest (t2, phi0) hopping.
    if (ic == 0) {
        if (jc != 0) return 0;
        if (name == "Haldane_CoreBN_Nedge") return -15.06;
        if (name == "Haldane_CoreBN_Bedge") return -6.91;
        return core;
    }
    cplx f2 = std::exp(-2. * pi * I1 * k[2] + 2. * pi * I1 * k[1]) + std::exp(2. * pi * I1 * k[2]) +
              std::exp(-2. * pi * I1 * k[1]);
    cplx f = 1. + std::exp(-2. * pi * I1 * k[1]) + std::exp(-2. * pi * I1 * k[2]);
    if (ic == 1) {
        if (jc == 1) return -g / 2. + 2. * std::real(t2 * std::exp(I1 * phi0) * f2);
        if (jc == 2) return b * (f);
        return 0.;
    }
    if (ic == 2) {
        if (jc == 1) return b * std::conj(f);
        if (jc == 2) return g / 2. + 2. * std::real(t2 * std::exp(-I1 * phi0) * f2);
        return 0.;
    }
    return 0.;
}

cplx TBModel::unitary(int ic, int jc, const double k[3]) const {
    const double pi = Units::pi();
    if (is_corebn(name)) {
        if (ic == 0) return jc == 0 ? cplx(1.) : cplx(0);
        if (jc == 0) return 0.;
        cplx f = 1. + std::exp(2. * pi * I1 * k[1]) + std::exp(2. * pi * I1 * k[2]);
        double E = std::sqrt(g * g + 4. * b * b * f * std::conj(f)).real() / 2.;
        double aa = std::sqrt((g + 2. * E) / (4. * E));
        double h = std::sqrt((2. * E - g) / (4. * E));
        if (ic == 1) return jc == 1 ? cplx(aa) : -h * std::conj(f) / std::abs(f);
        if (ic == 2) return jc == 1 ? cplx(h) : aa * std::conj(f) / std::abs(f);
        return 0.;
    }
    int o = 0;  // first band of the two-band block
    if (name != "BoronNitride") {
        if (ic == 0) return jc == 0 ? cplx(1.) : cplx(0);
        if (jc == 0) return 0.;
        o = 1;
    }
    double Bx = b * (1. + std::cos(2. * pi * k[1]) + std::cos(2. * pi * k[2]));
    double By = b * (std::sin(2. * pi * k[1]) + std::sin(2. * pi * k[2]));
    double Bz = -g / 2.;
    if (name != "BoronNitride")
        Bz = -g / 2. - 2. * t2 * std::sin(phi0) *
                           (std::sin(-2. * pi * k[2] + 2. * pi * k[1]) + std::sin(+2. * pi * k[2]) + std::sin(-2. * pi * k[1]));
    double B = std::sqrt(Bx * Bx + By * By + Bz * Bz);
    double phi = std::atan2(By, Bx) / 2.;
    double costh2 = std::sqrt((B + Bz) / 2. / B);
    double sinth2 = std::sqrt((B - Bz) / 2. / B);
    if (ic == o) return jc == o ? std::exp(+I1 * phi) * sinth2 : -std::exp(-I1 * phi) * costh2;
    if (ic == o + 1) return jc == o ? std::exp(+I1 * phi) * costh2 : std::exp(-I1 * phi) * sinth2;
    return 0.;
}

cplx TBModel::dipole(int ic, int jc, int x, const double k[3]) const {
    (void)k;
    if (name == "gen2d_hexagonal") {
        if (x == 0) {
            if ((ic == 0 && jc == 1) || (ic == 1 && jc == 0)) return rchv;
            if ((ic == 0 && jc == 2) || (ic == 2 && jc == 0)) return rchc;
            return 0.;
        }
        if (x == 1) {
            if (ic == 2 && jc == 2) return -a / std::sqrt(3.);
            if ((ic == 0 && jc == 0) && std::fabs(rchc) > 1.e-6) return -a / std::sqrt(3.);
            return 0.;
        }
        return 0.;
    }
    if (name == "CoreBN_Nedge" || name == "Haldane_CoreBN_Nedge") {
        if (x == 0) return ((ic == 0 && jc == 1) || (ic == 1 && jc == 0)) ? cplx(0.0697595094) : cplx(0.);
        if (x == 1) return (ic == 2 && jc == 2) ? cplx(a / std::sqrt(3.)) : cplx(0.);
        return 0.;
    }
    if (name == "CoreBN_Bedge" || name == "Haldane_CoreBN_Bedge") {
        if (x == 0) return ((ic == 0 && jc == 2) || (ic == 2 && jc == 0)) ? cplx(0.0869151662) : cplx(0.);
        if (x == 1) return ((ic == 2 && jc == 2) || (ic == 0 && jc == 0)) ? cplx(a / std::sqrt(3.)) : cplx(0.);
        return 0.;
    }
    if (name == "BoronNitride") {
        if (x == 1) return (ic == 1 && jc == 1) ? cplx(a / std::sqrt(3)) : cplx(0.);
        return 0.;
    }
    return 0.;
}

// ------------------------------------------------------------------ Wannier
void WannierModel::read_tb(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw Error("No path " + path + " file found.");
    std::string line;
    for (int i = 0; i < 4; i++) std::getline(f, line);
    f >> nw >> nR;
    if (!f || nw <= 0 || nR <= 0) throw Error("malformed wannier90 file " + path);
    ndegen.resize(nR);
    for (int i = 0; i < nR; i++) f >> ndegen[i];
    R.assign((size_t)nR * 3, 0);
    HR.assign((size_t)nR * nw * nw, 0.);
    rR.assign((size_t)nR * nw * nw * 3, 0.);
    const double eV = Units::energy_eV_au(), A = Units::space_A_au();
    // H block: per R the vector, then lines "m n Re Im" with m fastest.
    for (int iR = 0; iR < nR; iR++) {
        double Rv[3];
        f >> Rv[0] >> Rv[1] >> Rv[2];
        for (int c = 0; c < 3; c++) R[(size_t)iR * 3 + c] = (int32_t)std::lround(Rv[c]);
        for (int in = 0; in < nw; in++)
            for (int im = 0; im < nw; im++) {
                double m_, n_, hr, hi;
                f >> m_ >> n_ >> hr >> hi;
                HR[((size_t)iR * nw + im) * nw + in] = cplx(hr / ndegen[iR] * eV, hi / ndegen[iR] * eV);
            }
    }
    // Position block: per R the vector, then "m n xr xi yr yi zr zi".
    for (int iR = 0; iR < nR; iR++) {
        double Rv[3];
        f >> Rv[0] >> Rv[1] >> Rv[2];
        for (int in = 0; in < nw; in++)
            for (int im = 0; im < nw; im++) {
                double m_, n_, v[6];
                f >> m_ >> n_ >> v[0] >> v[1] >> v[2] >> v[3] >> v[4] >> v[5];
                for (int c = 0; c < 3; c++)
                    rR[(((size_t)iR * nw + im) * nw + in) * 3 + c] = cplx(v[2 * c] / ndegen[iR] * A, v[2 * c + 1] / ndegen[iR] * A);
            }
    }
    if (!f) throw Error("wannier90 file ended early: " + path);
}

namespace {
// Counter-based generator: splitmix64 of (seed, counter) -> uniform (0,1).
uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t seed, uint64_t ctr) {
    uint64_t z = mix64(seed * 0x100000001B3ull ^ mix64(ctr));
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
double normal(uint64_t seed, uint64_t ctr) {  // Box-Muller on two counters
    double u1 = uniform01(seed, 2 * ctr), u2 = uniform01(seed, 2 * ctr + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * Units::pi() * u2);
}
}  // namespace

void write_synthetic_tb(const std::string &path, int nw, int rmax, double a_ang, double sigma_h_ev,
                        double sigma_r_ang, double decay_ang, uint64_t seed) {
    const double pi = Units::pi();
    const int side = 2 * rmax + 1, nR = side * side;
    // Hexagonal in-plane lattice (as the gen2d_hexagonal built-in), Angstrom.
    const double a1[2] = {a_ang * std::cos(pi / 6.), a_ang * std::sin(pi / 6.)};
    const double a2[2] = {a_ang * std::cos(pi / 6.), -a_ang * std::sin(pi / 6.)};
    std::vector<cplx> H((size_t)nR * nw * nw), X((size_t)nR * nw * nw * 3);
    auto ridx = [&](int r1, int r2) { return (r1 + rmax) * side + (r2 + rmax); };
    uint64_t ctr = 0;
    for (int r1 = 0; r1 <= rmax; r1++)
        for (int r2 = -rmax; r2 <= rmax; r2++) {
            if (r1 == 0 && r2 < 0) continue;  // R > 0 lexicographically, and R = 0
            double y = r1 * a1[0] + r2 * a2[0], z = r1 * a1[1] + r2 * a2[1];
            double len = std::sqrt(y * y + z * z);
            double sh = sigma_h_ev * std::exp(-len / decay_ang) / std::sqrt(2.0);
            double sr = sigma_r_ang * std::exp(-len / decay_ang) / std::sqrt(2.0);
            int ip = ridx(r1, r2), im = ridx(-r1, -r2);
            for (int m = 0; m < nw; m++)
                for (int n = 0; n < nw; n++) {
                    cplx h(sh * normal(seed, ctr), sh * normal(seed, ctr + 1));
                    ctr += 2;
                    cplx xr[3];
                    for (int c = 0; c < 3; c++) {
                        xr[c] = cplx(sr * normal(seed, ctr), sr * normal(seed, ctr + 1));
                        ctr += 2;
                    }
                    if (r1 == 0 && r2 == 0) {  // Hermitian H(0), r(0) from the upper triangle
                        if (n < m) continue;
                        if (n == m) { h = h.real(); for (auto &v : xr) v = v.real(); }
                        H[((size_t)ip * nw + m) * nw + n] = h;
                        H[((size_t)ip * nw + n) * nw + m] = std::conj(h);
                        for (int c = 0; c < 3; c++) {
                            X[(((size_t)ip * nw + m) * nw + n) * 3 + c] = xr[c];
                            X[(((size_t)ip * nw + n) * nw + m) * 3 + c] = std::conj(xr[c]);
                        }
                    } else {  // H(-R) = H(R)^+
                        H[((size_t)ip * nw + m) * nw + n] = h;
                        H[((size_t)im * nw + n) * nw + m] = std::conj(h);
                        for (int c = 0; c < 3; c++) {
                            X[(((size_t)ip * nw + m) * nw + n) * 3 + c] = xr[c];
                            X[(((size_t)im * nw + n) * nw + m) * 3 + c] = std::conj(xr[c]);
                        }
                    }
                }
        }
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) throw Error("cannot write " + path);
    std::fprintf(fp, "synthetic wannier90 tb file, nw=%d rmax=%d seed=%llu\n", nw, rmax, (unsigned long long)seed);
    std::fprintf(fp, "  %.10f %.10f %.10f\n", 3.35, 0.0, 0.0);
    std::fprintf(fp, "  %.10f %.10f %.10f\n", 0.0, a1[0], a1[1]);
    std::fprintf(fp, "  %.10f %.10f %.10f\n", 0.0, a2[0], a2[1]);
    std::fprintf(fp, "%d\n%d\n", nw, nR);
    for (int i = 0; i < nR; i++) std::fprintf(fp, "%5d%s", 1, (i % 15 == 14 || i == nR - 1) ? "\n" : "");
    for (int r1 = -rmax; r1 <= rmax; r1++)
        for (int r2 = -rmax; r2 <= rmax; r2++) {
            int ip = ridx(r1, r2);
            std::fprintf(fp, "\n%5d%5d%5d\n", 0, r1, r2);
            for (int n = 0; n < nw; n++)
                for (int m = 0; m < nw; m++) {
                    cplx h = H[((size_t)ip * nw + m) * nw + n];
                    std::fprintf(fp, "%5d%5d %.17e %.17e\n", m + 1, n + 1, h.real(), h.imag());
                }
        }
    for (int r1 = -rmax; r1 <= rmax; r1++)
        for (int r2 = -rmax; r2 <= rmax; r2++) {
            int ip = ridx(r1, r2);
            std::fprintf(fp, "\n%5d%5d%5d\n", 0, r1, r2);
            for (int n = 0; n < nw; n++)
                for (int m = 0; m < nw; m++) {
                    const cplx *x = &X[(((size_t)ip * nw + m) * nw + n) * 3];
                    std::fprintf(fp, "%5d%5d %.17e %.17e %.17e %.17e %.17e %.17e\n", m + 1, n + 1, x[0].real(), x[0].imag(),
                                 x[1].real(), x[1].imag(), x[2].real(), x[2].imag());
                }
        }
    std::fclose(fp);
}

}  // namespace edus
