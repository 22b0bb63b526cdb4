// Small dense linear algebra for the host setup.  The reference calls Armadillo
// (inv, det, pinv, eig_sym: Coordinate.h:81-91, CalculateWeigths.h:91-100,
// Taylor_DE_solver_deriv_Polynom.cpp:117, Wannier.h:510); these are textbook
// replacements whose results agree with LAPACK's to rounding.
#include "host.hpp"

#include <algorithm>
#include <cmath>

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
std::vector<double> lstsq(const std::vector<double> &A0, int m, int n, const std::vector<double> &q0) {
    std::vector<double> A = A0, q = q0;
    for (int j = 0; j < n; j++) {
        double nrm = 0;
        for (int i = j; i < m; i++) nrm += A[i * n + j] * A[i * n + j];
        nrm = std::sqrt(nrm);
        if (nrm == 0) continue;
        double alpha = A[j * n + j] > 0 ? -nrm : nrm;
        std::vector<double> v(m, 0.0);
        for (int i = j; i < m; i++) v[i] = A[i * n + j];
        v[j] -= alpha;
        double vv = 0;
        for (int i = j; i < m; i++) vv += v[i] * v[i];
        if (vv == 0) continue;
        for (int c = j; c < n; c++) {
            double s = 0;
            for (int i = j; i < m; i++) s += v[i] * A[i * n + c];
            s = 2 * s / vv;
            for (int i = j; i < m; i++) A[i * n + c] -= s * v[i];
        }
        double s = 0;
        for (int i = j; i < m; i++) s += v[i] * q[i];
        s = 2 * s / vv;
        for (int i = j; i < m; i++) q[i] -= s * v[i];
    }
    std::vector<double> w(n, 0.0);
    for (int j = n - 1; j >= 0; j--) {
        double s = q[j];
        for (int c = j + 1; c < n; c++) s -= A[j * n + c] * w[c];
        double d = A[j * n + j];
        w[j] = (std::fabs(d) < 1e-300) ? 0.0 : s / d;
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
