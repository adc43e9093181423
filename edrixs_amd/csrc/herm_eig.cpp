// Dense complex Hermitian eigensolver on the host, for the full-diagonalization
// path of the ED driver (replaces LAPACK zheev, src/full_diag.f90:1-32, used
// when the dimension is below min_ndim).  Householder reduction to Hermitian
// tridiagonal form, a diagonal phase transform to a real symmetric tridiagonal
// matrix, implicit QL (tridiag_ql), and back-transformation.
#include <cmath>
#include <complex>
#include <vector>

namespace edx {

int tridiag_ql(int n, const double *d, const double *e, double *w, double *z, double *zlast);

// A: n*n row-major Hermitian (overwritten).  w: eigenvalues ascending.
// Z: n*n row-major, column j is eigenvector j (may be nullptr).
int herm_eig(int n, std::complex<double> *A, double *w, std::complex<double> *Z) {
    using cd = std::complex<double>;
    std::vector<cd> Q;
    if (Z) {
        Q.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    }
    std::vector<cd> u(n), p(n), q(n), qu(n);
    for (int k = 0; k + 2 < n; ++k) {
        double xn2 = 0.0;
        for (int i = k + 1; i < n; ++i) xn2 += std::norm(A[(size_t)i * n + k]);
        const double xn = std::sqrt(xn2);
        if (xn == 0.0) continue;
        const cd x0 = A[(size_t)(k + 1) * n + k];
        const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1.0, 0.0);
        const cd alpha = -ph * xn;
        for (int i = 0; i < n; ++i) u[i] = 0.0;
        for (int i = k + 1; i < n; ++i) u[i] = A[(size_t)i * n + k];
        u[k + 1] -= alpha;
        double un2 = 0.0;
        for (int i = k + 1; i < n; ++i) un2 += std::norm(u[i]);
        if (un2 == 0.0) continue;
        const double un = std::sqrt(un2);
        for (int i = k + 1; i < n; ++i) u[i] /= un;
        // p = A u, K = u^H p, q = p - K u ; A -= 2 u q^H + 2 q u^H
        for (int i = 0; i < n; ++i) {
            cd s = 0.0;
            for (int j = k + 1; j < n; ++j) s += A[(size_t)i * n + j] * u[j];
            p[i] = s;
        }
        cd K = 0.0;
        for (int i = k + 1; i < n; ++i) K += std::conj(u[i]) * p[i];
        const double Kr = K.real();
        for (int i = 0; i < n; ++i) q[i] = p[i] - Kr * u[i];
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                A[(size_t)i * n + j] -= 2.0 * (u[i] * std::conj(q[j]) + q[i] * std::conj(u[j]));
        if (Z) {   // Q <- Q (I - 2 u u^H)
            for (int i = 0; i < n; ++i) {
                cd s = 0.0;
                for (int j = k + 1; j < n; ++j) s += Q[(size_t)i * n + j] * u[j];
                qu[i] = s;
            }
            for (int i = 0; i < n; ++i)
                for (int j = k + 1; j < n; ++j) Q[(size_t)i * n + j] -= 2.0 * qu[i] * std::conj(u[j]);
        }
    }
    std::vector<double> d(n), e(n > 1 ? n - 1 : 1);
    std::vector<cd> phase(n);
    phase[0] = 1.0;
    for (int k = 0; k < n; ++k) d[k] = A[(size_t)k * n + k].real();
    for (int k = 0; k + 1 < n; ++k) {
        const cd t = A[(size_t)(k + 1) * n + k];
        const double at = std::abs(t);
        e[k] = at;
        phase[k + 1] = at > 0 ? phase[k] * (t / at) : phase[k];
    }
    std::vector<double> zt(Z ? (size_t)n * n : 0);
    if (tridiag_ql(n, d.data(), e.data(), w, Z ? zt.data() : nullptr, nullptr)) return -1;
    if (Z) {   // eigvec j = Q D z_j
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                cd s = 0.0;
                for (int c = 0; c < n; ++c)
                    s += Q[(size_t)i * n + c] * phase[c] * zt[(size_t)j * n + c];
                Z[(size_t)i * n + j] = s;
            }
    }
    return 0;
}

}  // namespace edx
