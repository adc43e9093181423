// Dense complex Hermitian eigensolver on the host, for the full-diagonalization
// path of the ED driver (replaces LAPACK zheev, src/full_diag.f90:1-32, used
// when the dimension is below min_ndim) and the projected matrices of the
// block Lanczos solver.  Householder reduction to Hermitian tridiagonal form
// (Golub & Van Loan, "Matrix Computations", sec. 8.3.1), a diagonal phase
// transform to a real symmetric tridiagonal matrix, implicit QL (tridiag_ql),
// and back-transformation of the requested eigenvectors by the stored
// reflections.  Each reduction step touches only the trailing block, with the
// complex arithmetic of the O(n^3) loops written out on real and imaginary
// parts.
#include <cmath>
#include <complex>
#include <vector>

namespace edx {

int tridiag_ql(int n, const double *d, const double *e, double *w, double *z, double *zlast);

namespace {

using cd = std::complex<double>;

// A (n x n row-major Hermitian, overwritten) -> real tridiagonal (d, e) with
// A = Q D T D^H Q^H, Q = H_0 ... H_{n-3}, H_k = I - 2 u_k u_k^H (u_k stored in
// row k of U, entries k+1..n-1; zero row: H_k = I), D = diag(phase).
void reduce(int n, cd *A, std::vector<cd> &U, std::vector<double> &d, std::vector<double> &e,
            std::vector<cd> &phase) {
    U.assign((size_t)n * n, cd(0.0, 0.0));
    std::vector<double> ur(n), ui(n), pr(n), pi(n), qr(n), qi(n);
    double *a = reinterpret_cast<double *>(A);   // interleaved (re, im)
    for (int k = 0; k + 2 < n; ++k) {
        const int j0 = k + 1;
        double xn2 = 0.0;
        for (int i = j0; i < n; ++i) xn2 += std::norm(A[(size_t)i * n + k]);
        if (xn2 == 0.0) continue;
        const double xn = std::sqrt(xn2);
        const cd x0 = A[(size_t)j0 * n + k];
        const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1.0, 0.0);
        const cd alpha = -ph * xn;
        double un2 = 0.0;
        for (int i = j0; i < n; ++i) {
            cd v = A[(size_t)i * n + k];
            if (i == j0) v -= alpha;
            ur[i] = v.real();
            ui[i] = v.imag();
            un2 += v.real() * v.real() + v.imag() * v.imag();
        }
        if (un2 == 0.0) continue;
        const double inv = 1.0 / std::sqrt(un2);
        for (int i = j0; i < n; ++i) {
            ur[i] *= inv;
            ui[i] *= inv;
            U[(size_t)k * n + i] = cd(ur[i], ui[i]);
        }
        // p = A22 u
        for (int i = j0; i < n; ++i) {
            const double *row = a + 2 * ((size_t)i * n);
            double sr = 0.0, si = 0.0;
            for (int j = j0; j < n; ++j) {
                const double xr = row[2 * j], xi = row[2 * j + 1];
                sr += xr * ur[j] - xi * ui[j];
                si += xr * ui[j] + xi * ur[j];
            }
            pr[i] = sr;
            pi[i] = si;
        }
        double K = 0.0;   // u^H p, real for Hermitian A22
        for (int i = j0; i < n; ++i) K += ur[i] * pr[i] + ui[i] * pi[i];
        for (int i = j0; i < n; ++i) {
            qr[i] = pr[i] - K * ur[i];
            qi[i] = pi[i] - K * ui[i];
        }
        // A22 -= 2 (u q^H + q u^H)
        for (int i = j0; i < n; ++i) {
            double *row = a + 2 * ((size_t)i * n);
            const double uir = 2.0 * ur[i], uii = 2.0 * ui[i];
            const double qir = 2.0 * qr[i], qii = 2.0 * qi[i];
            for (int j = j0; j < n; ++j) {
                // u_i conj(q_j) + q_i conj(u_j)
                row[2 * j] -= uir * qr[j] + uii * qi[j] + qir * ur[j] + qii * ui[j];
                row[2 * j + 1] -= uii * qr[j] - uir * qi[j] + qii * ur[j] - qir * ui[j];
            }
        }
        // column k becomes alpha e_{k+1}
        A[(size_t)j0 * n + k] = alpha;
        A[(size_t)k * n + j0] = std::conj(alpha);
        for (int i = j0 + 1; i < n; ++i) {
            A[(size_t)i * n + k] = 0.0;
            A[(size_t)k * n + i] = 0.0;
        }
    }
    d.assign(n, 0.0);
    e.assign(n > 1 ? n - 1 : 1, 0.0);
    phase.assign(n, cd(1.0, 0.0));
    for (int k = 0; k < n; ++k) d[k] = A[(size_t)k * n + k].real();
    for (int k = 0; k + 1 < n; ++k) {
        const cd t = A[(size_t)(k + 1) * n + k];
        const double at = std::abs(t);
        e[k] = at;
        phase[k + 1] = at > 0 ? phase[k] * (t / at) : phase[k];
    }
}

}  // namespace

// Eigenvalues w (ascending, all n) and the eigenvectors of the nvec lowest:
// Z n x nvec row-major, column j = eigenvector j (Z may be null when nvec == 0).
int herm_eig_lowest(int n, std::complex<double> *A, double *w, int nvec, std::complex<double> *Z) {
    if (n <= 0) return 0;
    std::vector<cd> U, phase;
    std::vector<double> d, e;
    reduce(n, A, U, d, e, phase);
    std::vector<double> zt(nvec > 0 ? (size_t)n * n : 0);   // row j: eigenvector j of (d, e)
    if (tridiag_ql(n, d.data(), e.data(), w, nvec > 0 ? zt.data() : nullptr, nullptr)) return -1;
    if (nvec <= 0) return 0;
    // Y = D z_j for the requested columns, then Y <- H_k Y for k = n-3 .. 0
    std::vector<double> yr((size_t)n * nvec), yi((size_t)n * nvec);   // row-major n x nvec
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < nvec; ++j) {
            const double z = zt[(size_t)j * n + i];
            yr[(size_t)i * nvec + j] = phase[i].real() * z;
            yi[(size_t)i * nvec + j] = phase[i].imag() * z;
        }
    std::vector<double> sr(nvec), si(nvec);
    for (int k = n - 3; k >= 0; --k) {
        const cd *u = U.data() + (size_t)k * n;
        for (int j = 0; j < nvec; ++j) sr[j] = si[j] = 0.0;
        for (int i = k + 1; i < n; ++i) {   // s = u^H Y
            const double ar = u[i].real(), ai = -u[i].imag();
            if (ar == 0.0 && ai == 0.0) continue;
            const double *r = yr.data() + (size_t)i * nvec, *m = yi.data() + (size_t)i * nvec;
            for (int j = 0; j < nvec; ++j) {
                sr[j] += ar * r[j] - ai * m[j];
                si[j] += ar * m[j] + ai * r[j];
            }
        }
        for (int i = k + 1; i < n; ++i) {   // Y -= 2 u s
            const double ar = 2.0 * u[i].real(), ai = 2.0 * u[i].imag();
            if (ar == 0.0 && ai == 0.0) continue;
            double *r = yr.data() + (size_t)i * nvec, *m = yi.data() + (size_t)i * nvec;
            for (int j = 0; j < nvec; ++j) {
                r[j] -= ar * sr[j] - ai * si[j];
                m[j] -= ar * si[j] + ai * sr[j];
            }
        }
    }
    for (size_t i = 0; i < (size_t)n * nvec; ++i) Z[i] = cd(yr[i], yi[i]);
    return 0;
}

// A: n*n row-major Hermitian (overwritten).  w: eigenvalues ascending.
// Z: n*n row-major, column j is eigenvector j (may be nullptr).
int herm_eig(int n, std::complex<double> *A, double *w, std::complex<double> *Z) {
    return herm_eig_lowest(n, A, w, Z ? n : 0, Z);
}

}  // namespace edx
