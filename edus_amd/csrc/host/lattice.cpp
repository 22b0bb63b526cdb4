// Lattice, k-grid, gradient shells and neighbour tables.
//
// Reciprocal vectors: ReadInput.h:388-391 / 894-897.  Crystal<->cartesian:
// Coordinate.h:68-149.  k-grid ordering: InitParameters.h:13-40.  Shells and
// weights (Mostofi et al., CPC 178, 685 (2008), sec. 3.2): CalculateWeigths.h:1-138.
// Neighbours: the reference searches k+b over the whole grid
// (CalculateWeigths.h:145-226, O(Nk^2)); on a uniform grid whose b-vectors are
// grid points the answer is an integer offset per shell vector, computed here in
// O(Nk) and checked against the brute-force form (neighbours_bruteforce).
#include "host.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <unordered_map>

namespace edus {

static void cross(const double *x, const double *y, double *out) {
    out[0] = x[1] * y[2] - x[2] * y[1];
    out[1] = x[2] * y[0] - x[0] * y[2];
    out[2] = x[0] * y[1] - x[1] * y[0];
}

void reciprocal_from_direct(const double a[3][3], double b[3][3]) {
    const double pi = Units::pi();
    cross(a[1], a[2], b[0]);
    cross(a[2], a[0], b[1]);
    cross(a[0], a[1], b[2]);
    for (int i = 0; i < 3; i++) {
        double v = (a[i][0] * b[i][0] + a[i][1] * b[i][1] + a[i][2] * b[i][2]) / 2. / pi;
        b[i][0] /= v; b[i][1] /= v; b[i][2] /= v;
    }
}

void Lattice::set_from_b(const double b[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = b[j][i];
    double det;
    inv3(M, MI, &det);
    jacobian = std::fabs(det);
}

void Lattice::to_cart(const double crys[3], double cart[3]) const {
    for (int i = 0; i < 3; i++) {
        cart[i] = 0.;
        for (int j = 0; j < 3; j++) cart[i] += M[i][j] * crys[j];
    }
}

void Lattice::to_crys(const double cart[3], double crys[3]) const {
    for (int i = 0; i < 3; i++) crys[i] = MI[i][0] * cart[0] + MI[i][1] * cart[1] + MI[i][2] * cart[2];
}

std::vector<double> make_kgrid(const int Nk[3]) {
    int nk = Nk[0] * Nk[1] * Nk[2];
    std::vector<double> kpt((size_t)nk * 3);
    double dk[3];
    for (int c = 0; c < 3; c++) dk[c] = 1. / Nk[c];
    for (int ik = 0; ik < nk; ik++) {
        int ii[3] = {(ik / (Nk[2] * Nk[1])) % Nk[0], (ik / Nk[2]) % Nk[1], ik % Nk[2]};
        for (int c = 0; c < 3; c++) kpt[(size_t)ik * 3 + c] = (dk[c] != 1.) ? -0.5 + dk[c] * ii[c] : 0.;
    }
    return kpt;
}

static double norm3(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// Shells of grid points nearest Gamma, added until the B1 + fourth-moment system
// A w = q (21 rows) is solved to 1e-7 (CalculateWeigths.h:18-133).
Shells find_shells(const std::vector<double> &kpt, const int Nk[3], const Lattice &L, bool grid) {
    for (int c = 0; c < 3 && grid; c++)
        if (Nk[c] > 1 && (Nk[c] % 2) != 0)
            throw Error("k-grid sizes must be even (no Gamma point on the grid otherwise)");
    int plane = -1;
    for (int i = 0; i < 3; i++)
        if (Nk[i] == 1) plane = i;
    const int nk = (int)(kpt.size() / 3);
    std::vector<double> norms(nk);
    std::vector<std::array<double, 3>> cart(nk);
    for (int ik = 0; ik < nk; ik++) {
        double c[3];
        L.to_cart(&kpt[(size_t)ik * 3], c);
        cart[ik] = {c[0], c[1], c[2]};
        norms[ik] = norm3(c);
    }
    Shells sh;
    std::vector<double> mins{1.e-08};
    std::vector<std::vector<int>> shell_members;
    for (int count = 0;; count++) {
        double m = 1000.;
        for (int ik = 0; ik < nk; ik++) {
            bool in_prev = false;
            for (double mm : mins) in_prev = in_prev || (norms[ik] <= mm + 1.e-11);
            if (!in_prev && norms[ik] < m) m = norms[ik];
        }
        mins.push_back(m);
        std::vector<std::array<double, 3>> bv;
        std::vector<int> mem;
        for (int ik = 0; ik < nk; ik++)
            if (std::fabs(norms[ik] - m) < 1.e-11) { bv.push_back(cart[ik]); mem.push_back(ik); }
        sh.bvec.push_back(bv);
        shell_members.push_back(mem);
        const int ns = (int)sh.bvec.size();
        // A (21 x ns): second moments of the in-plane directions, their cross
        // moment, then from row 5 the fourth moments (idir<=jdir<=ldir<=mdir).
        std::vector<double> A(21 * (size_t)ns, 0.0);
        for (int s = 0; s < ns; s++) {
            int row = 0;
            for (int i = 0; i < 3; i++)
                if (i != plane) {
                    for (auto &v : sh.bvec[s]) A[row * ns + s] += v[i] * v[i];
                    row++;
                }
            for (int i = 0; i < 3; i++)
                for (int j = i + 1; j < 3; j++)
                    if (i != plane && j != plane) {
                        for (auto &v : sh.bvec[s]) A[row * ns + s] += v[i] * v[j];
                        row++;
                    }
            row = 5;
            for (int i = 0; i < 3; i++)
                for (int j = i; j < 3; j++)
                    for (int l = j; l < 3; l++)
                        for (int mm = l; mm < 3; mm++) {
                            for (auto &v : sh.bvec[s]) A[row * ns + s] += v[i] * v[j] * v[l] * v[mm];
                            row++;
                        }
        }
        std::vector<double> q(21, 0.0);
        q[0] = 1; q[1] = 1; q[2] = !(plane + 1);
        std::vector<double> w = lstsq(A, 21, ns, q);
        double r2 = 0;
        for (int row = 0; row < 21; row++) {
            double acc = 0;
            for (int s = 0; s < ns; s++) acc += A[row * ns + s] * w[s];
            r2 += (acc - q[row]) * (acc - q[row]);
        }
        sh.weights = w;
        sh.residual = std::sqrt(r2);
        if (sh.residual < 1.e-07) break;
        if (count + 1 >= 32) throw Error("More than 32 shells needed. The program could be very slow!");
    }
    // Integer grid offsets of each b-vector: b is the grid point ik relative to Gamma.
    if (!grid) return sh;
    for (auto &mem : shell_members)
        for (int ik : mem) {
            int ii[3] = {(ik / (Nk[2] * Nk[1])) % Nk[0], (ik / Nk[2]) % Nk[1], ik % Nk[2]};
            std::array<int, 3> o;
            for (int c = 0; c < 3; c++) o[c] = (Nk[c] == 1) ? 0 : ii[c] - Nk[c] / 2;
            sh.offsets.push_back(o);
        }
    return sh;
}

std::vector<int32_t> neighbours_closed_form(const int Nk[3], const Shells &sh) {
    const int nk = Nk[0] * Nk[1] * Nk[2];
    const int nn = (int)sh.offsets.size();
    std::vector<int32_t> nbr((size_t)nk * nn);
    for (int ik = 0; ik < nk; ik++) {
        int ii[3] = {(ik / (Nk[2] * Nk[1])) % Nk[0], (ik / Nk[2]) % Nk[1], ik % Nk[2]};
        for (int n = 0; n < nn; n++) {
            int jj[3];
            for (int c = 0; c < 3; c++) {
                int v = (ii[c] + sh.offsets[n][c]) % Nk[c];
                jj[c] = v < 0 ? v + Nk[c] : v;
            }
            nbr[(size_t)ik * nn + n] = (jj[0] * Nk[1] + jj[1]) * Nk[2] + jj[2];
        }
    }
    return nbr;
}

// k + b in an explicit list: the reference's matching rule (k+b in cartesian, back
// to crystal, wrapped with fmod(x+0.5,1)-0.5; a list point matches when every
// coordinate differs by less than 1e-8 modulo 1; the first match in list order)
// with the list hashed on a 1e-6 lattice: each point is entered with its
// images at +-1 in every coordinate, and a query looks at the 27 cells around it.
std::vector<int32_t> neighbours_search(const std::vector<double> &kpt, const Shells &sh, const Lattice &L) {
    const int nk = (int)(kpt.size() / 3);
    std::vector<std::array<double, 3>> bv;
    for (auto &s : sh.bvec)
        for (auto &v : s) bv.push_back(v);
    const int nn = (int)bv.size();
    const double q = 1.e6;
    auto key = [](long long a, long long b, long long c) {
        return (uint64_t)(a * 73856093LL) ^ (uint64_t)(b * 19349663LL) ^ (uint64_t)(c * 83492791LL);
    };
    std::unordered_multimap<uint64_t, int> cells;
    cells.reserve((size_t)nk * 27);
    for (int j = 0; j < nk; j++)
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++)
                for (int c = -1; c <= 1; c++)
                    cells.emplace(key(std::llround((kpt[(size_t)j * 3] + a) * q), std::llround((kpt[(size_t)j * 3 + 1] + b) * q),
                                      std::llround((kpt[(size_t)j * 3 + 2] + c) * q)),
                                  j);
    std::vector<int32_t> nbr((size_t)nk * nn, -1);
    for (int ik = 0; ik < nk; ik++) {
        double kc[3];
        L.to_cart(&kpt[(size_t)ik * 3], kc);
        for (int n = 0; n < nn; n++) {
            double pc[3] = {kc[0] + bv[n][0], kc[1] + bv[n][1], kc[2] + bv[n][2]};
            double cr[3];
            L.to_crys(pc, cr);
            for (int c = 0; c < 3; c++) cr[c] = std::fmod(cr[c] + 0.5, 1.) - 0.5;
            const long long c0 = std::llround(cr[0] * q), c1 = std::llround(cr[1] * q), c2 = std::llround(cr[2] * q);
            int found = -1;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++)
                    for (int c = -1; c <= 1; c++) {
                        auto range = cells.equal_range(key(c0 + a, c1 + b, c2 + c));
                        for (auto it = range.first; it != range.second; ++it) {
                            const int j = it->second;
                            if (found >= 0 && j >= found) continue;
                            bool ok = true;
                            for (int d = 0; d < 3 && ok; d++) {
                                const double dd = cr[d] - kpt[(size_t)j * 3 + d];
                                ok = std::fabs(dd) < 1.e-08 || std::fabs(dd - 1.) < 1.e-08 || std::fabs(dd + 1.) < 1.e-08;
                            }
                            if (ok) found = j;
                        }
                    }
            if (found < 0) {
                char msg[256];
                std::snprintf(msg, sizeof msg, "k + b = (%.6e, %.6e, %.6e) of k point %d is not in the k list", cr[0], cr[1], cr[2], ik);
                throw Error(msg);
            }
            nbr[(size_t)ik * nn + n] = found;
        }
    }
    return nbr;
}

// The reference's search: k+b in cartesian, back to crystal, wrap with
// fmod(x+0.5,1)-0.5, first grid point within 1e-8 (modulo 1) in every direction.
std::vector<int32_t> neighbours_bruteforce(const std::vector<double> &kpt, const Shells &sh, const Lattice &L) {
    const int nk = (int)(kpt.size() / 3);
    std::vector<std::array<double, 3>> bv;
    for (auto &s : sh.bvec)
        for (auto &v : s) bv.push_back(v);
    const int nn = (int)bv.size();
    std::vector<int32_t> nbr((size_t)nk * nn, -1);
    for (int ik = 0; ik < nk; ik++) {
        double kc[3];
        L.to_cart(&kpt[(size_t)ik * 3], kc);
        for (int n = 0; n < nn; n++) {
            double pc[3] = {kc[0] + bv[n][0], kc[1] + bv[n][1], kc[2] + bv[n][2]};
            double cr[3];
            L.to_crys(pc, cr);
            for (int c = 0; c < 3; c++) cr[c] = std::fmod(cr[c] + 0.5, 1.) - 0.5;
            int found = -1;
            for (int j = 0; j < nk && found < 0; j++) {
                bool ok = true;
                for (int c = 0; c < 3 && ok; c++) {
                    double d = cr[c] - kpt[(size_t)j * 3 + c];
                    ok = std::fabs(d) < 1.e-08 || std::fabs(d - 1.) < 1.e-08 || std::fabs(d + 1.) < 1.e-08;
                }
                if (ok) found = j;
            }
            nbr[(size_t)ik * nn + n] = found;
        }
    }
    return nbr;
}

}  // namespace edus
