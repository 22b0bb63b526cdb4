// input.txt reader and laser pulses.
//
// Grammar and side effects follow headers/ReadInput.h:1-1234 (sections with a
// '{' on the header line, keys one per line, '}' closing) and the setters of
// headers/Laser.h:29-106; the pulse field is Laser::E (Laser.h:128-144).
#include "host.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <strings.h>

namespace edus {

// ------------------------------------------------------------------ Pulse
// Setter arithmetic of Laser.h:29-33: wavelength <-> frequency through
// 2*pi*c*0.0529, E0 = sqrt(I/3.51e16) (not the Constants.h intensity unit).
void Pulse::set_freq(double a) {
    wl = a;
    lambdanm = 2. * Units::pi() * Units::c_light * 0.0529 / wl;
    Period = 2 * Units::pi() / wl;
    if (Intensity != 0) { A0 = std::sqrt(Intensity / (3.51e+16)) / wl; E0 = A0 * wl; }
    if (ncycle != 0) tf = t0 + ncycle * Period;
}
void Pulse::set_lambdanm(double a) {
    lambdanm = a;
    wl = 2. * Units::pi() * Units::c_light * 0.0529 / lambdanm;
    Period = 2 * Units::pi() / wl;
    if (Intensity != 0) { A0 = std::sqrt(Intensity / (3.51e+16)) / wl; E0 = A0 * wl; }
    if (ncycle != 0) tf = t0 + ncycle * Period;
}
void Pulse::set_intensity_wcm2(double a) {
    Intensity = a;
    if (wl != 0) { A0 = std::sqrt(Intensity / (3.51e+16)) / wl; E0 = A0 * wl; }
}
void Pulse::set_ncycle_sigma(double a) {
    ncycle = a;
    sigma = a;
    if (Period != 0) tf = t0 + ncycle * Period;
}
void Pulse::set_t0(double a) {
    t0 = a;
    if (tf != 0.) tf += t0;
}
// Gram-Schmidt of the second polarisation vector against the first (Laser.h:82-106).
void Pulse::set_pol(double a0, double a1, double a2, double b0, double b1, double b2) {
    if ((b0 != 0. && b1 != 0. && b2 != 0.) &&
        (std::fabs(a0 / (b0 + 1e-11) - a1 / (b1 + 1e-11)) < 1e-9 &&
         std::fabs(a1 / (b1 + 1e-11) - a2 / (b2 + 1e-11)) < 1e-9))
        throw Error("care! you defined a plane with two parallel vectors!");
    if (std::fabs(a0 * b0 + a1 * b1 + a2 * b2) >= 1e-7) {
        double proj = (a0 * b0 + a1 * b1 + a2 * b2) / (a0 * a0 + a1 * a1 + a2 * a2);
        b0 -= proj * a0;
        b1 -= proj * a1;
        b2 -= proj * a2;
        double nrm = std::sqrt(b0 * b0 + b1 * b1 + b2 * b2);
        b0 /= nrm; b1 /= nrm; b2 /= nrm;
        if (std::fabs(a0 * b0 + a1 * b1 + a2 * b2) >= 1e-7)
            std::fprintf(stderr, "WARNING! unsuccessful orthogonalization!! care!!\n");
    }
    pol1 = {a0, a1, a2};
    pol2 = {b0, b1, b2};
}

void Pulse::field(double t, double out[3]) const {
    double env;
    if (gaussian) {
        // Gaussian envelope truncated where (t-t0)^2 >= 40 sigma^2 (Laser.h:56-60).
        double g = 0.;
        if (std::pow(t - t0, 2) < 40 * sigma * sigma) g = std::exp(-std::pow(t - t0, 2) / (2 * sigma * sigma));
        env = E0 * g;
    } else if (sin2) {
        double s = 0.;
        if (t >= t0 && t <= tf) s = std::pow(std::sin(Units::pi() * (t - t0) / (tf - t0)), 2);
        env = E0 * s;
    } else if (PW) {
        double tr = 0.;
        if (t >= t0 && t < w1) tr = (t - t0) / (w1 - t0);
        else if (t >= w1 && t < w2) tr = 1.;
        else if (t >= w2 && t <= tf) tr = (t - tf) / (w2 - tf);
        env = E0 * tr;
    } else {
        env = 0.;
    }
    double p1 = envelope ? 1. : std::sin(wl * (t - t0) + phase);
    double p2 = envelope ? 1. : std::cos(wl * (t - t0) + phase);
    for (int i = 0; i < 3; i++) out[i] = env * (p1 * pol1[i] + p2 * pol2[i]);
}

// ------------------------------------------------------------------ tokeniser
namespace {

bool ieq(const std::string &a, const char *b) { return strcasecmp(a.c_str(), b) == 0; }

// Tokens of s[0:pos) separated by blanks (ReadInput.h:1-33 splits on ' ' only).
std::vector<std::string> tokens(const std::string &s, size_t pos) {
    std::vector<std::string> out;
    std::string cur;
    size_t n = std::min(pos, s.size());
    for (size_t w = 0; w < n; w++) {
        char ch = s[w];
        if (ch != ' ' && ch != '\t') cur += ch;
        else if (!cur.empty()) { out.push_back(cur); cur.clear(); }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

bool blank(const std::string &s) {
    for (char ch : s)
        if (ch != ' ' && ch != '\t') return false;
    return true;
}

struct Reader {
    std::ifstream f;
    bool getline(std::string &s) {
        if (!std::getline(f, s)) return false;
        if (!s.empty() && s.back() == '\r') s.pop_back();
        return true;
    }
};

[[noreturn]] void fail(const char *fmt, const std::string &a = "", const std::string &b = "") {
    char buf[512];
    std::snprintf(buf, sizeof buf, fmt, a.c_str(), b.c_str());
    throw Error(buf);
}

// End_Section (ReadInput.h:54-81): the next non-blank line must hold '}'.
void end_section(Reader &r, const std::string &name) {
    std::string s;
    while (r.getline(s)) {
        if (s.empty() || blank(s)) continue;
        if (s.find('}') == std::string::npos) fail("Cannot find end of %s. Correct your input.", name);
        return;
    }
}

double tok_d(const std::vector<std::string> &t, size_t i) { return i < t.size() ? std::atof(t[i].c_str()) : 0.0; }

// Intensity with optional unit (ReadInput.h:86-103).
void read_intensity(const std::vector<std::string> &t, Pulse &p) {
    if (t.size() == 2 || (t.size() > 2 && ieq(t[2], "wcm2"))) p.set_intensity_wcm2(tok_d(t, 1));
    else if (ieq(t[2], "au") || ieq(t[2], "atomicunit")) p.set_intensity_wcm2(tok_d(t, 1) * Units::intensity_au_Wcm2);
    else fail("Cannot understand %s as units.", t[2]);
}

void read_wavelength(const std::vector<std::string> &t, Pulse &p) {
    if (t.size() == 2 || (t.size() > 2 && ieq(t[2], "nm"))) p.set_lambdanm(tok_d(t, 1));
    else if (ieq(t[2], "au") || ieq(t[2], "atomicunit")) p.set_lambdanm(tok_d(t, 1) * 10. * Units::space_au_A);
    else fail("Cannot understand %s as units.", t[2]);
}

// Linear or circular polarisation, normalised then Gram-Schmidt (ReadInput.h:135-159).
void read_polarization(const std::vector<std::string> &t, Pulse &p) {
    if (t.size() > 1 && ieq(t[1], "circular")) {
        if (t.size() < 8) fail("polarization circular needs 6 numbers%s%s");
        double v1 = tok_d(t, 2), v2 = tok_d(t, 3), v3 = tok_d(t, 4), v4 = tok_d(t, 5), v5 = tok_d(t, 6), v6 = tok_d(t, 7);
        double n = std::sqrt(v1 * v1 + v2 * v2 + v3 * v3); v1 /= n; v2 /= n; v3 /= n;
        n = std::sqrt(v4 * v4 + v5 * v5 + v6 * v6); v4 /= n; v5 /= n; v6 /= n;
        p.set_pol(v1, v2, v3, v4, v5, v6);
    } else if ((t.size() > 1 && ieq(t[1], "linear")) || t.size() == 4) {
        size_t o = ieq(t[1], "linear") ? 2 : 1;
        double v1 = tok_d(t, o), v2 = tok_d(t, o + 1), v3 = tok_d(t, o + 2);
        double n = std::sqrt(v1 * v1 + v2 * v2 + v3 * v3); v1 /= n; v2 /= n; v3 /= n;
        p.set_pol(v1, v2, v3, 0., 0., 0.);
    }
}

double read_sigma(const std::vector<std::string> &t) {
    double s = tok_d(t, 1);
    if (t.size() >= 3 && ieq(t[2], "fs")) s *= Units::time_fs_au();
    return s;
}

// One laserpump section (ReadInput.h:236-326).  Returns gaussian flag and sigma
// through the arguments, as the probe's delay uses the last pump's.
Pulse read_pump(Reader &r, const std::vector<std::string> &hdr, const std::string &line, bool &gaussian, double &sigma) {
    Pulse p;
    bool pw = false, s2 = false;
    double t0 = 0;
    gaussian = false;
    if (hdr.size() < 2) fail("in %s you have to specify laser type: gaussian or sin2. \n", line);
    if (ieq(hdr[1], "gaussian")) gaussian = true;
    else if (hdr[1] == "sin2") s2 = true;
    else if (hdr[1] == "pw") pw = true;
    else fail("Allowed types for laserpump: gaussian, sin2. Your type \" %s \" is not available.\n", hdr[1]);
    p.envelope = false; p.gaussian = gaussian; p.PW = pw; p.sin2 = s2;
    std::string s;
    if (!r.getline(s)) fail("Can't find the end of laserpump section.\n");
    while (s.find('}') == std::string::npos || s.empty()) {
        auto t = tokens(s, s.size());
        if (!t.empty()) {
            const std::string &k = t[0];
            if (ieq(k, "frequency")) {
                if (t.size() == 2) p.set_freq(tok_d(t, 1));
                else fail("Cannot understand %s as units.", t.size() > 2 ? t[2] : "");
            } else if (ieq(k, "phase")) p.phase = tok_d(t, 1);
            else if (ieq(k, "window")) {
                if (!pw) fail("You defined a window for a pulse that is not pseudoPW!! Error.\n");
                double w1 = tok_d(t, 1), w2 = tok_d(t, 2);
                if (t.size() > 3 && ieq(t[3], "fs")) { w1 *= Units::time_fs_au(); w2 *= Units::time_fs_au(); }
                p.w1 = w1; p.w2 = w2;
            } else if (ieq(k, "wavelength")) read_wavelength(t, p);
            else if (ieq(k, "intensity")) read_intensity(t, p);
            else if (ieq(k, "polarization")) read_polarization(t, p);
            else if (ieq(k, "sigma")) {
                if (s2) fail("You can't define sigma for sin2 wave. Change your input.\n");
                sigma = read_sigma(t);
                p.set_ncycle_sigma(sigma);
            } else if (ieq(k, "cycles")) {
                if (gaussian) fail("You can't define cycles for gaussian wave. Change your input.\n");
                sigma = tok_d(t, 1);
                p.set_ncycle_sigma(sigma);
            } else if (ieq(k, "t0_pulse")) {
                t0 = tok_d(t, 1);
                if (t.size() > 2 && ieq(t[2], "fs")) t0 *= Units::time_fs_au();
            } else {
                fail("cannot recognize key %s", s);
            }
        }
        if (!r.getline(s)) fail("Can't find the end of laserpump section.\n");
    }
    if (gaussian) t0 += 4. * sigma;   // the gaussian peak sits 4 sigma after t0_pulse
    p.set_t0(t0);
    return p;
}

}  // namespace

// ------------------------------------------------------------------ parse_input
Input parse_input(const std::string &path) {
    Input in;
    Reader r;
    r.f.open(path);
    if (!r.f.is_open()) throw Error("error opening file " + path);
    {
        size_t sl = path.find_last_of('/');
        in.base_dir = (sl == std::string::npos) ? std::string(".") : path.substr(0, sl);
    }
    r.f >> in.ref_coord;
    in.pumps.resize(1);
    int i_pump = 0;
    bool gaussian1 = false;
    double sigma1 = 0, sigma2 = 0;
    std::string s;
    while (r.getline(s)) {
        if (s.empty()) continue;
        size_t pos = s.find('{');
        if (pos == std::string::npos) {
            if (blank(s)) continue;
            fail("Can't find the character { in the string \" %s \". \nPlease, review your input.\n", s);
        }
        auto str = tokens(s, pos);
        if (str.empty()) fail("Can't find a section name in \" %s \".\n", s);
        const std::string &key = str[0];

        if (ieq(key, "unitcell")) {
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    if (!(r.f >> in.a[i][j])) fail("Cannot read unit cell because the file ended. Correct your input.%s%s");
            reciprocal_from_direct(in.a, in.b);
            if (str.size() == 2) {
                if (ieq(str[1], "angstrom")) {
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) { in.a[i][j] *= Units::space_A_au(); in.b[i][j] /= Units::space_A_au(); }
                } else if (!(ieq(str[1], "au") || ieq(str[1], "atomicunit"))) {
                    fail("Can't recognize the units %s in the line %s, please modify your input.\nAllowed units: angstrom, au / atomicunit", str[1], s);
                }
            }
            end_section(r, key);
        } else if (ieq(key, "nkpt")) {
            if (!(r.f >> in.Nk[0] >> in.Nk[1] >> in.Nk[2])) fail("Cannot read Nkpt because the file ended. Correct your input.%s%s");
            if (str.size() == 2) {
                if (ieq(str[1], "read")) in.kptread = true;
                else std::printf("%s can't have units %s. Please modify your input.", str[0].c_str(), str[1].c_str());
            }
            if (in.kptread) {  // the k points follow the sizes, three crystal coordinates each (ReadInput.h:432-439)
                const long long n = (long long)in.Nk[0] * in.Nk[1] * in.Nk[2];
                if (n < 1) fail("Cannot read Nkpt because the file ended. Correct your input.%s%s");
                in.kpt_list.resize((size_t)n * 3);
                for (long long i = 0; i < 3 * n; i++)
                    if (!(r.f >> in.kpt_list[(size_t)i])) fail("Cannot read the k points of nkPT read: the list ended early.%s%s");
            }
            end_section(r, key);
        } else if (ieq(key, "bands")) {
            s.clear();
            while (s.empty()) if (!r.getline(s)) fail("Can't find the end of bands section.\n");
            while (s.find('}') == std::string::npos || s.empty()) {
                auto t = tokens(s, s.size());
                if (!t.empty()) {
                    if (ieq(t[0], "conduction")) in.Nb[2] = std::atoi(t.size() > 1 ? t[1].c_str() : "0");
                    else if (ieq(t[0], "valence")) in.Nb[1] = std::atoi(t.size() > 1 ? t[1].c_str() : "0");
                    else if (ieq(t[0], "core")) in.Nb[0] = std::atoi(t.size() > 1 ? t[1].c_str() : "0");
                    else if (ieq(t[0], "conduction_shift")) in.change_gap_constant = tok_d(t, 1);
                    else fail("Can't recognize band type %s. Possible band types: conduction, valence, core.\n", t[0]);
                }
                if (!r.getline(s)) fail("Can't find the end of bands section.\n");
            }
        } else if (ieq(key, "laserpump") && i_pump < in.num_pump_pulses) {
            in.pumps[i_pump] = read_pump(r, str, s, gaussian1, sigma1);
            i_pump++;
        } else if (ieq(key, "laserprobe")) {
            // ReadInput.h:492-726: constant envelope (no carrier), delay keyword,
            // unknown keys are reported but tolerated.
            Pulse &p = in.probe;
            bool g2 = false, s2 = false, pw = false;
            if (str.size() < 2) fail("in %s you have to specify laser type: gaussian or sin2. \n", s);
            if (ieq(str[1], "gaussian")) g2 = true;
            else if (str[1] == "sin2") s2 = true;
            else if (str[1] == "pw") pw = true;
            else fail("Allowed types for laserpump: gaussian, sin2. Your type \" %s \" is not available.\n", str[1]);
            if (!r.getline(s)) fail("Can't find the end of laserprobe section.\n");
            p.envelope = true; p.gaussian = g2; p.PW = pw; p.sin2 = s2;
            while (s.find('}') == std::string::npos || s.empty()) {
                auto t = tokens(s, s.size());
                if (!t.empty()) {
                    const std::string &k = t[0];
                    if (ieq(k, "frequency")) {
                        if (t.size() == 2) p.set_freq(tok_d(t, 1));
                        else fail("Cannot understand %s as units.", t.size() > 2 ? t[2] : "");
                    } else if (ieq(k, "wavelength")) read_wavelength(t, p);
                    else if (ieq(k, "window")) {
                        if (!pw) fail("You defined a window for a pulse that is not pseudoPW!! Error.\n");
                        double w1 = tok_d(t, 1), w2 = tok_d(t, 2);
                        if (t.size() > 3 && ieq(t[3], "fs")) { w1 *= Units::time_fs_au(); w2 *= Units::time_fs_au(); }
                        p.w1 = w1; p.w2 = w2;
                    } else if (ieq(k, "intensity")) read_intensity(t, p);
                    else if (ieq(k, "polarization")) read_polarization(t, p);
                    else if (ieq(k, "sigma")) {
                        if (s2) fail("You can't define sigma for sin2 wave. Change your input.\n");
                        if (t.size() < 3) sigma2 = tok_d(t, 1);
                        else if (ieq(t[2], "fs")) sigma2 = tok_d(t, 1) * Units::time_fs_au();
                        else sigma2 = tok_d(t, 1);
                        p.set_ncycle_sigma(sigma2);
                    } else if (ieq(k, "delay")) {
                        if (t.size() == 2 || (t.size() > 2 && ieq(t[2], "fs"))) in.DELAY = tok_d(t, 1) * Units::time_fs_au();
                        else if (ieq(t[2], "au") || ieq(t[2], "atomicunit")) in.DELAY = tok_d(t, 1);
                        else fail("Error in reading time delay. Please review your input.\n");
                    } else if (ieq(k, "cycles")) {
                        if (g2) fail("You can't define cycles for gaussian wave. Change your input.\n");
                        sigma2 = tok_d(t, 1);
                        p.set_ncycle_sigma(sigma2);
                    } else {
                        std::fprintf(stderr, "cannot recognize key %s in laser probe.\n", s.c_str());
                    }
                }
                if (!r.getline(s)) fail("Can't find the end of laserprobe section.\n");
            }
            // Probe centre: after the (last read) gaussian pump's 4 sigma, or at the
            // middle of the first pump minus half the probe's "cycles" (ReadInput.h:721-724).
            double t0;
            if (gaussian1) t0 = 4. * sigma1 + in.DELAY;
            else t0 = (in.pumps[0].tf - in.pumps[0].t0) / 2. + in.DELAY - p.ncycle / 2. * p.Period;
            p.set_t0(t0);
        } else if (ieq(key, "TDSE")) {
            if (!r.getline(s)) fail("Can't find the end of tdse section.\n");
            while (s.find('}') == std::string::npos) {
                auto t = tokens(s, s.size());
                if (!t.empty()) {
                    const std::string &k = t[0];
                    if (ieq(k, "Wannier")) in.iMode = "W";
                    else if (ieq(k, "Crystal")) in.iMode = "CY";
                    else if (ieq(k, "tightbinding")) {
                        if (t.size() < 2) fail("Please specify the TightBinding type.%s%s");
                        in.TBtype = t[1];
                        const double pi = Units::pi();
                        const std::string &tb = in.TBtype;
                        auto set_a = [&](double a00, double a11, double a12, double a21, double a22) {
                            for (auto &row : in.a) row[0] = row[1] = row[2] = 0.;
                            in.a[0][0] = a00; in.a[1][1] = a11; in.a[1][2] = a12; in.a[2][1] = a21; in.a[2][2] = a22;
                        };
                        // Built-in lattices and band counts of ReadInput.h:762-888 (BN1b, DLGraphene,
                        // Vampa2015, GeS: 800-835).
                        if (ieq(tb, "graphene") || ieq(tb, "GrapheneZurron")) {
                            double L = 2.46 * Units::space_A_au();
                            set_a(6.3303100000, L * std::cos(pi / 6.), L * std::sin(pi / 6.), L * std::cos(pi / 6.), -L * std::sin(pi / 6.));
                            in.Nb[0] = 0; in.Nb[1] = 1; in.Nb[2] = 1;
                        } else if (ieq(tb, "CoreGraphene") || ieq(tb, "Graphene_costD") || ieq(tb, "Graphene_symmetricD")) {
                            double L = 2.46 * Units::space_A_au();
                            set_a(6.3303100000, L * std::cos(pi / 6.), L * std::sin(pi / 6.), L * std::cos(pi / 6.), -L * std::sin(pi / 6.));
                            in.Nb[0] = 2; in.Nb[1] = 1; in.Nb[2] = 1;
                        } else if (ieq(tb, "BoronNitride")) {
                            double L = 2.5 * Units::space_A_au();
                            set_a(6.3303100000, L * std::cos(pi / 6.), L * std::sin(pi / 6.), L * std::cos(pi / 6.), -L * std::sin(pi / 6.));
                            in.Nb[0] = 0; in.Nb[1] = 1; in.Nb[2] = 1;
                        } else if (ieq(tb, "CoreBN_Nedge") || ieq(tb, "CoreBN_Bedge")) {
                            double L = 2.5 * Units::space_A_au();
                            set_a(6.3303100000, L * std::cos(pi / 6.), L * std::sin(pi / 6.), L * std::cos(pi / 6.), -L * std::sin(pi / 6.));
                            in.Nb[0] = 1; in.Nb[1] = 1; in.Nb[2] = 1;
                        } else if (ieq(tb, "Haldane_CoreBN_Bedge") || ieq(tb, "Haldane_CoreBN_Nedge")) {
                            double L = 2.5 * Units::space_A_au();
                            double det = 1.0;
                            set_a(6.3303100000, L * std::cos(det * pi / 6.), L * std::sin(det * pi / 6.), L * std::cos(det * pi / 6.),
                                  -L * std::sin(det * pi / 6.));
                            in.Nb[0] = 1; in.Nb[1] = 1; in.Nb[2] = 1;
                            in.hopping.assign(4, 0.);
                            for (size_t i = 2; i < t.size(); i++) {
                                double v = (i + 1 < t.size()) ? std::atof(t[i + 1].c_str()) : 0.;
                                if (ieq(t[i], "gap=")) in.hopping[0] = v * Units::energy_eV_au();
                                else if (ieq(t[i], "t1=")) in.hopping[1] = v * Units::energy_eV_au();
                                if (ieq(t[i], "t2=")) in.hopping[2] = v * Units::energy_eV_au();
                                else if (ieq(t[i], "phi0=")) in.hopping[3] = v;
                            }
                        } else if (ieq(tb, "gen2d_hexagonal")) {
                            in.hopping.assign(8, 0.);
                            for (size_t i = 2; i < t.size(); i++) {
                                double v = (i + 1 < t.size()) ? std::atof(t[i + 1].c_str()) : 0.;
                                if (ieq(t[i], "a=")) in.hopping[0] = v * Units::space_A_au();
                                else if (ieq(t[i], "gap=")) in.hopping[1] = v * Units::energy_eV_au();
                                else if (ieq(t[i], "t1=")) in.hopping[2] = v * Units::energy_eV_au();
                                else if (ieq(t[i], "t2=")) in.hopping[3] = v * Units::energy_eV_au();
                                else if (ieq(t[i], "phi0=")) in.hopping[4] = v;
                                else if (ieq(t[i], "core=")) in.hopping[5] = v * Units::energy_eV_au();
                                else if (ieq(t[i], "chv_dipole=")) in.hopping[6] = v;
                                else if (ieq(t[i], "chc_dipole=")) in.hopping[7] = v;
                            }
                            double A = in.hopping[0];
                            set_a(6.3303100000, A * std::cos(pi / 6.), A * std::sin(pi / 6.), A * std::cos(pi / 6.), -A * std::sin(pi / 6.));
                            in.Nb[0] = 1; in.Nb[1] = 1; in.Nb[2] = 1;
                        } else if (ieq(tb, "BN1b")) {  // the reference writes cos for both in-plane components
                            double L = 2.5 * Units::space_A_au();
                            set_a(6.3303100000, L * std::cos(pi / 6.), L * std::cos(pi / 6.), L * std::cos(pi / 6.), -L * std::cos(pi / 6.));
                            in.Nb[0] = 0; in.Nb[1] = 1; in.Nb[2] = 0;
                        } else if (ieq(tb, "DLGraphene")) {
                            set_a(6.3303100000, 4.0257407400, -2.3242630000, 4.0257407400, 2.3242630000);
                            in.Nb[0] = 0; in.Nb[1] = 2; in.Nb[2] = 2;
                        } else if (ieq(tb, "Vampa2015")) {
                            set_a(5.3200000000, 6.1400000000, 0.0, 0.0, 9.8300000000);
                            in.Nb[0] = 0; in.Nb[1] = 1; in.Nb[2] = 1;
                        } else if (ieq(tb, "GeS") || ieq(tb, "GeS_HSE06")) {
                            double ax = (4.53 / 2.) * Units::space_A_au(), ay = (3.63 / 2.) * Units::space_A_au();
                            set_a(6.3303100000, ax, ay, ax, -ay);
                            in.Nb[0] = 0; in.Nb[1] = 1; in.Nb[2] = 1;
                        } else {
                            throw Error("TightBinding type " + tb + " is not supported by this library");
                        }
                        reciprocal_from_direct(in.a, in.b);
                        in.iMode = "TB";
                    } else if (ieq(k, "FermiEnergy")) {
                        if (t.size() < 3) in.FermiE = tok_d(t, 1);
                        else throw Error("FermiEnergy takes one value");
                    } else if (ieq(k, "dt") || ieq(k, "t_fin")) {
                        // value [fs|au]; a bare number is taken in a.u.
                        double v = tok_d(t, 1);
                        if (t.size() >= 3) {
                            if (ieq(t[2], "fs")) v *= Units::time_fs_au();
                            else if (!(ieq(t[2], "au") || ieq(t[2], "atomicunit")))
                                fail("Units %s aren't allowed for dt.\n", t[2]);
                        }
                        if (ieq(k, "dt")) in.dt = v; else in.t_fin = v;
                    } else if (ieq(k, "epsStepAbs")) in.epsStepAbs = tok_d(t, 1);
                    else if (ieq(k, "dynamical_dt_evolution")) { in.const_dt = false; in.dyn_dt = true; }
                    else if (ieq(k, "No_Vectorization")) in.vectorization = false;
                    else if (ieq(k, "Taylor")) in.taylor = true;
                    else if (ieq(k, "num_pump_pulses")) {
                        in.num_pump_pulses = (int)tok_d(t, 1);
                        if (in.num_pump_pulses < 1) throw Error("num_pump_pulses must be >= 1");
                        in.pumps.resize(in.num_pump_pulses);
                    } else if (ieq(k, "TaylorOrder")) in.solver_order = (int)tok_d(t, 1);
                    else if (ieq(k, "Adams_Bashforth")) {
                        in.adams_bashforth = true;
                        in.taylor = false;
                        if (t.size() > 1) in.solver_order = (int)tok_d(t, 1);
                        if (in.solver_order > 5) in.solver_order = 5;
                    } else if (ieq(k, "Gap_correction")) in.gap_correction = tok_d(t, 1);
                }
                if (!r.getline(s)) fail("Can't find the end of tdse section.\n");
            }
        } else if (ieq(key, "observables")) {
            if (!r.getline(s)) fail("Can't find the end of observables section.\n");
            while (s.find('}') == std::string::npos) {
                auto t = tokens(s, s.size());
                if (!t.empty()) {
                    const std::string &k = t[0];
                    if (ieq(k, "wfd")) in.iWFDs = true;
                    else if (ieq(k, "Current")) in.iCurrent = true;
                    else if (ieq(k, "it_resolution")) in.it_resolution = (int)tok_d(t, 1);
                    else if (ieq(k, "PrintPopulation")) {
                        in.print_population = true;
                        if (t.size() > 1) {
                            in.print_start = tok_d(t, 1);
                            in.print_end = tok_d(t, 2);
                            in.print_step = tok_d(t, 3);
                        }
                    }
                    else if (ieq(k, "Coulomb_band_reconstruction")) in.coulomb = true;
                    else if (ieq(k, "TransientAbsorption") || ieq(k, "TAbs")) in.iTAbs = true;
                    else if (ieq(k, "kresolvedTAbs") || ieq(k, "kTAbs") || ieq(k, "kTransientAbsorption")) in.iTAbsK = true;
                    else fail("option %s not available in the observables.\n", k);
                }
                if (!r.getline(s)) fail("Can't find the end of observables section.\n");
            }
        } else if (ieq(key, "TAbsK")) {
            if (!in.iTAbsK) fail("Error in the input. You defined the section TAbsK without putting it in the observables.\n");
            s.clear();
            while (s.empty()) if (!r.getline(s)) fail("Can't find the end of TAbsK section.\n");
            auto t = tokens(s, s.size());
            if (t.size() != 1) fail("Error in TAbsK. you didn't define in the first line the number of k points.");
            int n = std::atoi(t[0].c_str());
            for (int i = 0; i < n; i++) r.getline(s);
            end_section(r, "TAbsK");
        } else if (ieq(key, "Coulomb")) {
            in.coulomb = true;
            if (!r.getline(s)) fail("Can't find the end of Coulomb section.\n");
            while (s.find('}') == std::string::npos) {
                auto t = tokens(s, s.size());
                if (!t.empty()) {
                    const std::string &k = t[0];
                    if (ieq(k, "qTF")) in.qTF = tok_d(t, 1);
                    if (ieq(k, "r0")) in.r0 = tok_d(t, 1);
                    if (ieq(k, "G_distance")) in.G_distance = tok_d(t, 1);
                    if (ieq(k, "Rytova_Keldysh")) in.rytova_keldysh = true;
                    if (ieq(k, "epsilon_static")) in.epsilon_static = tok_d(t, 1);
                    if (ieq(k, "Coulomb_diag_basis")) in.coulomb_diag_basis = true;
                    if (ieq(k, "Read_Coulomb_from_files")) in.coulomb_read = true;
                    if (ieq(k, "labelInput") && t.size() > 1) in.coulomb_label = t[1];
                    if (ieq(k, "Ncut")) in.Ncut = (int)tok_d(t, 1);
                    if (ieq(k, "Print_new_band_dispersion")) in.print_band_dispersion = true;
                }
                if (!r.getline(s)) fail("Can't find the end of Coulomb section.\n");
            }
        } else if (ieq(key, "decoherence")) {
            if (!r.getline(s)) fail("Can't find the end of decoherence section.\n");
            while (s.find('}') == std::string::npos) {
                auto t = tokens(s, s.size());
                if (!t.empty()) {
                    if (t.size() < 2) fail("You have to specify a number for the decoherence term.%s%s");
                    double v = tok_d(t, 1);
                    if (t.size() == 3) {
                        if (ieq(t[2], "meV")) v *= Units::energy_eV_au() / 1000.;
                        else if (!(ieq(t[2], "au") || ieq(t[2], "atomicunit"))) fail("Can't understand units %s in decoherence.\n", t[2]);
                    }
                    if (ieq(t[0], "diagonal")) in.T1 = v;
                    else if (ieq(t[0], "offdiagonal")) in.T2 = v;
                    else if (ieq(t[0], "corehole")) in.Tch = v;
                    else fail("Unknow key %s in decoherence.", t[0]);
                }
                if (!r.getline(s)) fail("Can't find the end of tdse section.\n");
            }
        }
    }
    return in;
}

}  // namespace edus
