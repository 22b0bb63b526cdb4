// Band models feeding the H, U, D tables.
//
// Tight binding: the model families of headers/TightBinding.h (init 44-147,
// set_hopping 153-173, energy 175-459, unitary 892-1213, dipoles 1256-1421):
// the boron-nitride / Haldane / gen2d family, the graphene family (Graphene,
// CoreGraphene, Graphene_costD, Graphene_symmetricD, GrapheneZurron, DLGraphene),
// BN1b, Vampa2015 and GeS / GeS_HSE06.  Each matrix element is the model's
// closed form; the expressions keep the reference's operand order so the values
// agree to the last bit where libm agrees.
// Wannier: reader of wannier90 <seed>_tb.dat as in headers/Wannier.h:47-149 and a
// seeded synthetic generator writing the same format.
#include "host.hpp"

#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

namespace edus {

namespace {
const cplx I1(0.0, 1.0);

bool is_haldane(const std::string &n) { return n == "Haldane_CoreBN_Nedge" || n == "Haldane_CoreBN_Bedge"; }
bool is_corebn(const std::string &n) { return n == "CoreBN_Nedge" || n == "CoreBN_Bedge"; }
bool is_graphene_d(const std::string &n) { return n == "Graphene_costD" || n == "Graphene_symmetricD"; }
bool is_ges(const std::string &n) { return n == "GeS" || n == "GeS_HSE06"; }
bool is_graphene_family(const std::string &n) {
    return n == "Graphene" || n == "CoreGraphene" || n == "GrapheneZurron" || is_graphene_d(n) || n == "DLGraphene";
}

// Vampa2015 band coefficients (TightBinding.h:98-121): [band][direction][harmonic j] of
// sum_x sum_j c cos(j k_x); the y and z directions carry only j = 0, 1.
const double kVampa[2][3][6] = {
    {{-0.0928, 0.0705, 0.0200, -0.0012, 0.0029, 0.0006}, {-0.0307, 0.0307, 0, 0, 0, 0}, {-0.0059, 0.0059, 0, 0, 0, 0}},
    {{0.0898, -0.0814, -0.0024, -0.0048, -0.0003, -0.0009}, {0.1147, -0.1147, 0, 0, 0, 0}, {0.0435, -0.0435, 0, 0, 0, 0}}};

// Eigenvectors of a two-band block written as B0 + B.sigma, the form shared by
// BoronNitride, the Haldane block and GeS (TightBinding.h:958-1029): from
// (Bx, By, Bz), rows (e^{i phi} sin, -e^{-i phi} cos) and (e^{i phi} cos,
// e^{-i phi} sin) of the half polar angle, phi half the azimuth.
cplx bloch_unitary(int r, int col, double Bx, double By, double Bz) {
    double B = std::sqrt(Bx * Bx + By * By + Bz * Bz);
    double phi = std::atan2(By, Bx) / 2.;
    double costh2 = std::sqrt((B + Bz) / 2. / B);
    double sinth2 = std::sqrt((B - Bz) / 2. / B);
    if (r == 0) return col == 0 ? std::exp(+I1 * phi) * sinth2 : -std::exp(-I1 * phi) * costh2;
    if (r == 1) return col == 0 ? std::exp(+I1 * phi) * costh2 : std::exp(-I1 * phi) * sinth2;
    return 0.;
}
}  // namespace

void TBModel::init(const std::string &nm, int nch, int ncv, const std::vector<double> &hopping, const double b_rows[3][3]) {
    name = nm;
    Nch = nch;
    Ncv = ncv;
    if (!supported()) throw Error("TightBinding type " + nm + " is not supported by this library");
    if (b_rows)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) kcart[i][j] = b_rows[j][i];
    const double A = Units::space_A_au(), eV = Units::energy_eV_au();
    // Default parameters of TightBinding::init (TightBinding.h:44-147).
    if (is_graphene_family(name)) {
        a = 2.46 * A; b = 2.97 * eV; c = 0; d = 0; g = 0;
        if (name == "DLGraphene") { c = 3.35 * A; d = 0.39 * eV; }
    } else if (name == "Vampa2015") {
        std::memcpy(vampa, kVampa, sizeof vampa);
        g = 0.1213;
    } else if (name == "GeS") {
        g = 0.41 * eV; f = -2.33 * eV; t2 = 0.61 * eV; b = 0.13 * eV; c = 0.07 * eV; d = -0.09 * eV;
    } else if (name == "GeS_HSE06") {
        g = 0.62 * eV; f = 2.91 * eV; t2 = -0.797 * eV; b = -0.05 * eV; c = -0.14 * eV; d = 0.106 * eV;
    } else {  // boron-nitride family, BN1b, Haldane, gen2d
        a = 2.5 * A; b = 3.16 * eV; c = 0; d = 0; g = 6.2 * eV;
    }
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
    return name == "BoronNitride" || is_corebn(name) || is_haldane(name) || name == "gen2d_hexagonal" || name == "BN1b" ||
           is_graphene_family(name) || name == "Vampa2015" || is_ges(name);
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
    // Nearest-neighbour structure factor of the hexagonal lattice in crystal k.
    auto fhex = [&]() { return std::exp(I1 * 2. * pi * k[1]) + 1. + std::exp(I1 * 2. * pi * k[2]); };
    if (name == "BoronNitride" || name == "Graphene") {
        if (ic == jc) return name == "Graphene" ? cplx(0.) : (ic == 0 ? cplx(-g / 2) : (ic == 1 ? cplx(g / 2) : cplx(0)));
        cplx fk = fhex();
        return ic == 0 ? std::conj(fk) * b : fk * b;
    }
    if (is_graphene_d(name)) {  // two core levels, the graphene bands already diagonal
        if (ic == jc && ic < 2) return -10.892178;
        if (ic >= 2 && jc >= 2 && ic == jc) {
            double modf = std::abs(fhex());
            return ic == 2 ? cplx(-b * modf) : cplx(b * modf);
        }
        return 0.;
    }
    if (name == "CoreGraphene") {
        if (ic == jc && ic < 2) return -10.892178;
        if (ic >= 2 && jc >= 2 && ic != jc) {
            cplx fk = fhex();
            return ic == 2 ? std::conj(fk) * b : fk * b;
        }
        return 0.;
    }
    if (name == "GrapheneZurron") {  // the structure factor in cartesian k (TightBinding.h:251-264)
        if (ic == jc) return 0.;
        double kc[3];
        for (int i = 0; i < 3; i++) kc[i] = kcart[i][0] * k[0] + kcart[i][1] * k[1] + kcart[i][2] * k[2];
        cplx fz = std::exp(-I1 * a * kc[1] / std::sqrt(3.)) * (1. + 2. * std::exp(I1 * std::sqrt(3.) * a * kc[1] / 2.) * std::cos(a * kc[2] / 2.));
        return ic == 0 ? fz * b : std::conj(fz) * b;
    }
    if (name == "DLGraphene") {  // two graphene layers coupled by d between orbitals 0 and 3
        if (ic == jc) return 0.;
        cplx fk = fhex();
        if ((ic == 0 && jc == 1) || (ic == 2 && jc == 3)) return std::conj(fk) * b;
        if ((ic == 1 && jc == 0) || (ic == 3 && jc == 2)) return fk * b;
        if ((ic == 3 && jc == 0) || (ic == 0 && jc == 3)) return d;
        return 0.;
    }
    if (name == "BN1b") {  // the conduction band of BoronNitride alone
        cplx fk = fhex();
        return 0.5 * std::sqrt(g * g + 4 * b * b * fk * std::conj(fk)).real();
    }
    if (name == "Vampa2015") {
        if (ic != jc) return 0.;
        cplx e = 0.;
        for (int x = 0; x < 3; x++)
            for (int j = 0; j < 6; j++) e += vampa[ic][x][j] * std::cos(j * k[x]);
        if (ic == 1) e += g;
        return e;
    }
    if (is_ges(name)) {
        if (ic == jc) {
            double tor = -2. * c * (std::cos(2. * pi * k[1]) + std::cos(2. * pi * k[2])) - 2. * d * (std::cos(2. * pi * (k[1] - k[2])));
            return ic == 0 ? cplx(g + tor) : cplx(-g + tor);
        }
        cplx phix = std::exp(I1 * 2. * pi * k[1]) + std::exp(I1 * 2. * pi * k[2]);
        cplx tor = -f - t2 * phix - b * std::conj(phix);
        return ic == 0 ? std::conj(tor) : tor;
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
    cplx fm = 1. + std::exp(-2. * pi * I1 * k[1]) + std::exp(-2. * pi * I1 * k[2]);
    if (ic == 1) {
        if (jc == 1) return -g / 2. + 2. * std::real(t2 * std::exp(I1 * phi0) * f2);
        if (jc == 2) return b * (fm);
        return 0.;
    }
    if (ic == 2) {
        if (jc == 1) return b * std::conj(fm);
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
        cplx fk = 1. + std::exp(2. * pi * I1 * k[1]) + std::exp(2. * pi * I1 * k[2]);
        double E = std::sqrt(g * g + 4. * b * b * fk * std::conj(fk)).real() / 2.;
        double aa = std::sqrt((g + 2. * E) / (4. * E));
        double h = std::sqrt((2. * E - g) / (4. * E));
        if (ic == 1) return jc == 1 ? cplx(aa) : -h * std::conj(fk) / std::abs(fk);
        if (ic == 2) return jc == 1 ? cplx(h) : aa * std::conj(fk) / std::abs(fk);
        return 0.;
    }
    if (is_ges(name)) {
        cplx phix = std::exp(I1 * 2. * pi * k[1]) + std::exp(I1 * 2. * pi * k[2]);
        cplx tor = -f - t2 * phix - b * std::conj(phix);
        return bloch_unitary(ic, jc, std::real(tor), std::imag(tor), g);
    }
    // Graphene-like eigenvectors (1, +-f/|f|)/sqrt(2) with f of -k (TightBinding.h:1031-1071).
    auto fneg = [&]() { return std::exp(-I1 * 2. * pi * k[1]) + 1. + std::exp(-I1 * 2. * pi * k[2]); };
    if (name == "Graphene") {
        if (jc == 1) {
            cplx fk = fneg();
            return ic == 0 ? -fk / (std::sqrt(2.) * std::abs(fk)) : fk / (std::sqrt(2.) * std::abs(fk));
        }
        return 1. / std::sqrt(2.);
    }
    if (name == "CoreGraphene") {
        if (ic < 2) return ic == jc ? cplx(1.) : cplx(0.);
        if (jc == 3) {
            cplx fk = fneg();
            return ic == 2 ? -fk / (std::sqrt(2.) * std::abs(fk)) : +fk / (std::sqrt(2.) * std::abs(fk));
        }
        return jc == 2 ? cplx(1. / std::sqrt(2.)) : cplx(0.);
    }
    if (is_graphene_d(name) || name == "BN1b" || name == "Vampa2015") return (ic == jc || name == "BN1b") ? cplx(1.) : cplx(0.);
    if (name == "GrapheneZurron") {
        if (jc == 1) {
            double kc[3];
            for (int i = 0; i < 3; i++) kc[i] = kcart[i][0] * k[0] + kcart[i][1] * k[1] + kcart[i][2] * k[2];
            cplx fz = std::exp(-I1 * a * kc[1] / std::sqrt(3.)) * (1. + 2. * std::exp(I1 * std::sqrt(3.) * a * kc[1] / 2.) * std::cos(a * kc[2] / 2.));
            return ic == 0 ? -fz / (std::sqrt(2.) * std::abs(fz)) : fz / (std::sqrt(2.) * std::abs(fz));
        }
        return 1. / std::sqrt(2.);
    }
    if (name == "DLGraphene") {
        // Closed-form eigenvectors of the bilayer (TightBinding.h:1072-1138), with
        // S = sqrt(d^2 + 4 b^2 |f|^2) and the two normalisations N1, N2.
        cplx fk = std::exp(I1 * 2. * pi * k[1]) + 1. + std::exp(I1 * 2. * pi * k[2]);
        cplx S = std::sqrt(std::pow(d, 2) + 4 * std::pow(b, 2) * fk * std::conj(fk));
        cplx Sc = std::conj(S);
        double N1 = std::sqrt(8 + std::pow(std::abs((d - S) / (b * std::conj(fk))), 2) + 16 * std::pow(std::abs((b * fk) / (d + S)), 2));
        double N2 = std::sqrt(8 + 16 * std::pow(std::abs((b * fk) / (d - S)), 2) + std::pow(std::abs((d + S) / (b * std::conj(fk))), 2));
        if (ic == 0 || ic == 3) {
            const double s = ic == 0 ? -1. : 1.;
            if (jc == 0) return s * 2 / N1;
            if (jc == 1) return (-d + Sc) / (b * fk * N1);
            if (jc == 2) return (s * 4 * b * fk) / (N1 * (d + Sc));
            return 2 / N1;
        }
        const double s = ic == 1 ? 1. : -1.;
        if (jc == 0) return s * 2 / N2;
        if (jc == 1) return -((d + Sc) / (b * fk * N2));
        if (jc == 2) return (s * 4 * b * fk) / (N2 * (d - Sc));
        return 2 / N2;
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
    return bloch_unitary(ic - o, jc - o, Bx, By, Bz);
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
    // The graphene family and the other lattices (TightBinding.h:1256-1421).
    const double y_pz = -2.68341111;  // position of the second graphene site along y
    if (name == "Graphene") return (x == 1 && ic == 1 && jc == 1) ? cplx(y_pz) : cplx(0.);
    if (name == "CoreGraphene") {
        if (x == 0) return ((ic == 0 && jc == 2) || (ic == 1 && jc == 3) || (ic == 2 && jc == 0) || (ic == 3 && jc == 1)) ? cplx(0.0780088) : cplx(0.);
        if (x == 1) return (ic == jc && ic % 2 == 1) ? cplx(y_pz) : cplx(0.);
        return 0.;
    }
    if (is_graphene_d(name)) {
        if (x == 2 && name == "Graphene_costD") return 0.;
        if (x == 0) {
            if (ic <= 1 && jc > 1) {
                const double dx = 0.0780088 / std::sqrt(2.);
                return (ic == 0 || (ic == 1 && jc == 2)) ? cplx(dx) : cplx(-dx);
            }
            if (ic > 1 && jc <= 1) return std::conj(dipole(jc, ic, 0, k));
            return 0.;
        }
        // y, and z of Graphene_symmetricD (dipolez = dipoley)
        if (ic == 1 && jc == 1) return y_pz;
        if (ic > 1 && jc > 1) return ic == jc ? cplx(y_pz / 2) : cplx(-y_pz / 2);
        return 0.;
    }
    if (name == "DLGraphene") {
        if (ic != jc) return 0.;
        if (x == 0) return (ic == 2 || ic == 3) ? cplx(c) : cplx(0.);
        if (x == 1) return ic == 1 ? cplx(-a / std::sqrt(3)) : ((ic == 2 || ic == 3) ? cplx(a / (2 * std::sqrt(3))) : cplx(0.));
        return ic == 2 ? cplx(a / 2.) : (ic == 3 ? cplx(-a / 2.) : cplx(0.));
    }
    if (name == "Vampa2015") return ic != jc ? cplx(x == 2 ? 3.94 : 3.46) : cplx(0.);
    if (is_ges(name)) return (x == 1 && ic == 1 && jc == 1) ? cplx(0.52 * Units::space_A_au()) : cplx(0.);
    return 0.;  // BN1b, GrapheneZurron
}

// ------------------------------------------------------------------ Wannier
namespace {
// The numbers of a file read whole and taken one at a time (std::from_chars:
// correctly rounded like stream extraction, locale-independent, and an order of
// magnitude faster on the 10^7 numbers of a large seedname_tb.dat).
struct NumberStream {
    std::string buf;
    const char *p = nullptr, *end = nullptr;
    bool ok = true;
    void skip_lines(int n) {
        for (; n > 0 && p < end; n--) {
            const char *q = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
            p = q ? q + 1 : end;
        }
    }
    double next() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
        if (p < end && *p == '+') p++;
        double v = 0.0;
        const auto r = std::from_chars(p, end, v);
        if (r.ec != std::errc() || r.ptr == p) {
            ok = false;
            return 0.0;
        }
        p = r.ptr;
        return v;
    }
};
}  // namespace

void WannierModel::read_tb(const std::string &path) {
    NumberStream f;
    {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw Error("No path " + path + " file found.");
        f.buf.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    f.p = f.buf.data();
    f.end = f.p + f.buf.size();
    f.skip_lines(4);
    const double nw_d = f.next(), nR_d = f.next();
    nw = (int)nw_d;
    nR = (int)nR_d;
    if (!f.ok || nw <= 0 || nR <= 0) throw Error("malformed wannier90 file " + path);
    ndegen.resize(nR);
    for (int i = 0; i < nR; i++) ndegen[i] = f.next();
    R.assign((size_t)nR * 3, 0);
    HR.assign((size_t)nR * nw * nw, 0.);
    rR.assign((size_t)nR * nw * nw * 3, 0.);
    const double eV = Units::energy_eV_au(), A = Units::space_A_au();
    // H block: per R the vector, then lines "m n Re Im" with m fastest.
    for (int iR = 0; iR < nR; iR++) {
        for (int c = 0; c < 3; c++) R[(size_t)iR * 3 + c] = (int32_t)std::lround(f.next());
        for (int in = 0; in < nw; in++)
            for (int im = 0; im < nw; im++) {
                f.next();
                f.next();
                const double hr = f.next(), hi = f.next();
                HR[((size_t)iR * nw + im) * nw + in] = cplx(hr / ndegen[iR] * eV, hi / ndegen[iR] * eV);
            }
    }
    // Position block: per R the vector, then "m n xr xi yr yi zr zi".
    for (int iR = 0; iR < nR; iR++) {
        for (int c = 0; c < 3; c++) f.next();
        for (int in = 0; in < nw; in++)
            for (int im = 0; im < nw; im++) {
                f.next();
                f.next();
                double v[6];
                for (int q = 0; q < 6; q++) v[q] = f.next();
                for (int c = 0; c < 3; c++)
                    rR[(((size_t)iR * nw + im) * nw + in) * 3 + c] = cplx(v[2 * c] / ndegen[iR] * A, v[2 * c + 1] / ndegen[iR] * A);
            }
    }
    if (!f.ok) throw Error("wannier90 file ended early: " + path);
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

namespace edus {

// ---------------------------------------------------------------- crystal interface
// Crystal_interface.h:19-304 with Interpolation.h (FittedData, 2D case): tabulated
// H(k) and Berry connections on an nkf x nkf grid of crystal (k1, k2) in
// [-0.5, 0.5] (spacing 1/(nkf-1)), interpolated by periodic bicubic convolution
// (Catmull-Rom weights in each direction) at any k.
namespace {
std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : s) {
        if (ch != ' ' && ch != '\t' && ch != '\r') cur += ch;
        else if (!cur.empty()) { out.push_back(cur); cur.clear(); }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}
bool last_is(const std::string &line, const char *word) {
    auto t = split_ws(line);
    if (t.empty()) return false;
    std::string a = t.back(), b = word;
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}
}  // namespace

void CrystalModel::read(const std::string &dir, double a[3][3], double b[3][3]) {
    auto open = [&](const std::string &name, std::ifstream &f) {
        f.open(dir + "/" + name);
        if (!f) throw Error("cannot open " + dir + "/" + name + " (Crystal interface)");
    };
    std::ifstream f;
    open("kmesh_data.dat", f);
    std::string line;
    auto next = [&](std::ifstream &s) {
        if (!std::getline(s, line)) throw Error("unexpected end of a Crystal interface file");
        return split_ws(line);
    };
    auto two = [&](double *row) {
        auto t = next(f);
        if (t.size() < 2) throw Error("kmesh_data.dat: expected two lattice components");
        row[1] = std::atof(t[0].c_str());
        row[2] = std::atof(t[1].c_str());
    };
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = b[i][j] = 0.0;
    next(f); two(a[1]);
    next(f); two(a[2]);
    a[0][0] = 15.;
    next(f); two(b[1]);
    next(f); two(b[2]);
    b[0][0] = 1. / 15.;
    do next(f); while (!last_is(line, "basis"));
    norb = std::atoi(next(f).at(0).c_str());
    do next(f); while (!last_is(line, "direction"));
    nkf = std::atoi(next(f).at(0).c_str());
    if (norb < 1 || nkf < 2) throw Error("kmesh_data.dat: bad orbital or k-point count");
    f.close();

    const size_t g = (size_t)nkf * nkf;
    h.assign((size_t)norb * norb * g, 0.0);
    bc.assign((size_t)norb * norb * 2 * g, 0.0);
    std::ifstream fh;
    open("hk_orthonormal.dat", fh);
    for (int i1 = 0; i1 < nkf; i1++)
        for (int i2 = 0; i2 < nkf; i2++) {
            const size_t kk = (size_t)i1 * nkf + i2;
            for (int part = 0; part < 2; part++) {
                for (int ic = 0; ic < norb; ic++) {
                    auto t = next(fh);
                    for (int jc = 0; jc <= ic; jc++) {
                        const double v = std::atof(t.at(jc).c_str());
                        cplx &e = h[((size_t)ic * norb + jc) * g + kk];
                        e += part ? cplx(0.0, v) : cplx(v, 0.0);
                    }
                }
                next(fh);
                if (part) next(fh);
            }
        }
    std::ifstream fr, fi;
    open("bc_betabetap_real.dat", fr);
    open("bc_betabetap_imag.dat", fi);
    for (int i1 = 0; i1 < nkf; i1++)
        for (int i2 = 0; i2 < nkf; i2++) {
            const size_t kk = (size_t)i1 * nkf + i2;
            for (int x = 0; x < 2; x++) {
                for (int ic = 0; ic < norb; ic++) {
                    auto tr = next(fr);
                    auto ti = next(fi);
                    for (int jc = 0; jc <= ic; jc++)
                        bc[(((size_t)ic * norb + jc) * 2 + x) * g + kk] = cplx(std::atof(tr.at(jc).c_str()), std::atof(ti.at(jc).c_str()));
                }
                next(fr); next(fi);
                if (x) { next(fr); next(fi); }
            }
        }
}

// Periodic bicubic convolution of a [nkf][nkf] table at crystal (k1, k2).
cplx CrystalModel::interp(const cplx *grid, const double k[3]) const {
    const double spacing = 1. / (nkf - 1), xmin = -0.5;
    double dy = std::fmod((k[1] - xmin) / spacing, nkf), dz = std::fmod((k[2] - xmin) / spacing, nkf);
    if (dy < 0) dy += nkf;
    if (dz < 0) dz += nkf;
    const int yi = (int)std::floor(dy), zi = (int)std::floor(dz);
    dy -= yi;
    dz -= zi;
    auto weights = [](double t, double w[4]) {
        const double t2 = t * t, t3 = t2 * t;
        w[0] = 0.5 * (-t3 + 2 * t2 - t);
        w[1] = 0.5 * (3 * t3 - 5 * t2 + 2);
        w[2] = 0.5 * (-3 * t3 + 4 * t2 + t);
        w[3] = 0.5 * (t3 - t2);
    };
    double wy[4], wz[4];
    weights(dy, wy);
    weights(dz, wz);
    auto wrap = [&](int i) { return ((i % nkf) + nkf) % nkf; };
    cplx v = 0.0;
    for (int r = 0; r < 4; r++) {
        cplx row = 0.0;
        const int iy = wrap(yi - 1 + r);
        for (int q = 0; q < 4; q++) row += wz[q] * grid[(size_t)iy * nkf + wrap(zi - 1 + q)];
        v += wy[r] * row;
    }
    return v;
}

cplx CrystalModel::energy(int ic, int jc, const double k[3]) const {
    const size_t g = (size_t)nkf * nkf;
    if (jc <= ic) return interp(&h[((size_t)ic * norb + jc) * g], k);
    return std::conj(interp(&h[((size_t)jc * norb + ic) * g], k));
}

// Component x of the reference's dipole (Crystal::dipole): 0 <- the third Berry
// component (not in the files: zero), 1 <- the first, 2 <- the second; real diagonal.
cplx CrystalModel::dipole(int ic, int jc, int x, const double k[3]) const {
    if (x == 0) return 0.0;
    const size_t g = (size_t)nkf * nkf;
    const int c = x - 1;
    cplx v = (jc <= ic) ? interp(&bc[(((size_t)ic * norb + jc) * 2 + c) * g], k)
                        : std::conj(interp(&bc[(((size_t)jc * norb + ic) * 2 + c) * g], k));
    if (ic == jc) v.imag(0.0);
    return v;
}

}  // namespace edus
