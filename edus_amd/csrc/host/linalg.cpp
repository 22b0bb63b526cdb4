// Small dense linear algebra for the host setup.  The reference calls Armadillo
// (inv, det, pinv, eig_sym: Coordinate.h:81-91, CalculateWeigths.h:91-100,
// Taylor_DE_solver_deriv_Polynom.cpp:117, Wannier.h:510); these are textbook
// replacements whose results agree with LAPACK's to rounding.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace edus {

void inv3(const double A[3][3], double Ainv[3][3], double *det) {
    double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    double d = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    if (d == 0.0) throw Error("singular 3x3 lattice matrix");
    Ainv[0][0] = c00 / d;
    Ainv[1][0] = c01 / d;
    Ainv[2][0] = c02 / d;
    Ainv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / d;
    Ainv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / d;
    Ainv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / d;
    Ainv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / d;
    Ainv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / d;
    Ainv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / d;
    if (det) *det = d;
}

// Least squares min ||A w - q|| for a full-column-rank m x n A (m >= n), by
// Householder QR.  Equals pinv(A) q when A has full column rank.
// Minimum-norm least squares w = pinv(A) q (the reference calls Armadillo's pinv,
// CalculateWeigths.h:91-100): singular value decomposition by one-sided Jacobi
// rotations of A's columns (A V = U S), then w = V S^+ U^T q with singular values
// at or below max(m, n) * max(s) * eps treated as zero (pinv's default
// tolerance).  The shell systems that are rank-deficient (lattices with shells of
// two vectors, the orthorhombic and rhombic ones) need the minimum-norm solution.
std::vector<double> lstsq(const std::vector<double> &A0, int m, int n, const std::vector<double> &q0) {
    std::vector<double> U = A0, V((size_t)n * n, 0.0);
    for (int j = 0; j < n; j++) V[(size_t)j * n + j] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 100; sweep++) {
        bool rotated = false;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double a = 0, b = 0, g = 0;
                for (int i = 0; i < m; i++) {
                    const double up = U[(size_t)i * n + p], uq = U[(size_t)i * n + q];
                    a += up * up;
                    b += uq * uq;
                    g += up * uq;
                }
                if (g == 0.0 || std::fabs(g) <= eps * std::sqrt(a * b)) continue;
                rotated = true;
                const double zeta = (b - a) / (2.0 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (int i = 0; i < m; i++) {
                    const double up = U[(size_t)i * n + p], uq = U[(size_t)i * n + q];
                    U[(size_t)i * n + p] = c * up - sn * uq;
                    U[(size_t)i * n + q] = sn * up + c * uq;
                }
                for (int i = 0; i < n; i++) {
                    const double vp = V[(size_t)i * n + p], vq = V[(size_t)i * n + q];
                    V[(size_t)i * n + p] = c * vp - sn * vq;
                    V[(size_t)i * n + q] = sn * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    std::vector<double> sv(n, 0.0);
    double smax = 0;
    for (int j = 0; j < n; j++) {
        double s2 = 0;
        for (int i = 0; i < m; i++) s2 += U[(size_t)i * n + j] * U[(size_t)i * n + j];
        sv[j] = std::sqrt(s2);
        smax = std::max(smax, sv[j]);
    }
    const double tol = std::max(m, n) * smax * eps;
    std::vector<double> w(n, 0.0);
    for (int j = 0; j < n; j++) {
        if (!(sv[j] > tol)) continue;
        double uq = 0;  // (U_j / s_j)^T q
        for (int i = 0; i < m; i++) uq += U[(size_t)i * n + j] * q0[i];
        const double coef = uq / (sv[j] * sv[j]);
        for (int r = 0; r < n; r++) w[r] += V[(size_t)r * n + j] * coef;
    }
    return w;
}

std::vector<double> inverse_gauss_jordan(const std::vector<double> &A0, int n) {
    std::vector<double> A = A0, I((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) I[i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
        if (A[p * n + c] == 0.0) throw Error("singular matrix in inverse");
        if (p != c)
            for (int k = 0; k < n; k++) { std::swap(A[p * n + k], A[c * n + k]); std::swap(I[p * n + k], I[c * n + k]); }
        double piv = A[c * n + c];
        for (int k = 0; k < n; k++) { A[c * n + k] /= piv; I[c * n + k] /= piv; }
        for (int r = 0; r < n; r++)
            if (r != c) {
                double f = A[r * n + c];
                if (f == 0.0) continue;
                for (int k = 0; k < n; k++) { A[r * n + k] -= f * A[c * n + k]; I[r * n + k] -= f * I[c * n + k]; }
            }
    }
    return I;
}

// Cyclic complex Jacobi: rotations zero |H_pq| until off-diagonal norm is at
// rounding level; eigenvalues sorted ascending; each eigenvector (column of V)
// phased so its largest-modulus component is real positive.
void eigh_jacobi(int n, const cplx *H0, double *w, cplx *V) {
    std::vector<cplx> A(H0, H0 + (size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int i = 0; i < n; i++) {
        A[i * n + i] = cplx(A[i * n + i].real(), 0.0);
        for (int j = i + 1; j < n; j++) {  // enforce exact Hermiticity from the upper triangle
            cplx h = 0.5 * (A[i * n + j] + std::conj(A[j * n + i]));
            A[i * n + j] = h;
            A[j * n + i] = std::conj(h);
        }
    }
    double fro = 0;
    for (int i = 0; i < n * n; i++) fro += std::norm(A[i]);
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += std::norm(A[p * n + q]);
        if (off <= 1e-28 * fro || off == 0.0) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                cplx apq = A[p * n + q];
                double mag = std::abs(apq);
                if (mag == 0.0) continue;
                double app = A[p * n + p].real(), aqq = A[q * n + q].real();
                double theta = 0.5 * (aqq - app) / mag;
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                cplx ph = apq / mag;  // e^{i phi}
                // Rotation J with J_pp = c, J_qq = c, J_pq = s*ph, J_qp = -s*conj(ph); A <- J^+ A J
                for (int k = 0; k < n; k++) {  // columns p, q
                    cplx akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * std::conj(ph) * akq;
                    A[k * n + q] = s * ph * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {  // rows p, q
                    cplx apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * ph * aqk;
                    A[q * n + k] = s * std::conj(ph) * apk + c * aqk;
                }
                A[p * n + q] = 0.0;
                A[q * n + p] = 0.0;
                A[p * n + p] = cplx(A[p * n + p].real(), 0.0);
                A[q * n + q] = cplx(A[q * n + q].real(), 0.0);
                for (int k = 0; k < n; k++) {
                    cplx vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * std::conj(ph) * vkq;
                    V[k * n + q] = s * ph * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::vector<double> ev(n);
    for (int i = 0; i < n; i++) ev[i] = A[i * n + i].real();
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return ev[x] < ev[y]; });
    std::vector<cplx> Vs((size_t)n * n);
    for (int c = 0; c < n; c++) {
        int src = idx[c];
        w[c] = ev[src];
        int best = 0;
        double bm = -1;
        for (int r = 0; r < n; r++) {
            double m = std::abs(V[r * n + src]);
            if (m > bm + 1e-12) { bm = m; best = r; }
        }
        cplx ph = std::conj(V[best * n + src]) / std::abs(V[best * n + src]);
        for (int r = 0; r < n; r++) Vs[r * n + c] = V[r * n + src] * ph;
    }
    std::copy(Vs.begin(), Vs.end(), V);
}

}  // namespace edus
