// Symmetric tridiagonal eigensolver on the host (replaces LAPACK dsteqr,
// src/lanczos.f90:169-206).  Implicit QL with Wilkinson shifts (Golub & Van
// Loan, "Matrix Computations", sec. 8.3).  Eigenvalues come out ascending.
//
// z (n*n, column j = eigenvector j) and zlast (last component of every
// eigenvector, the Lanczos residual weights) are both optional; zlast alone
// costs O(n) per rotation sweep instead of O(n^2).
#include <cmath>
#include <cstring>
#include <vector>

namespace edx {

int tridiag_ql(int n, const double *d_in, const double *e_in, double *w, double *z,
               double *zlast) {
    if (n <= 0) return 0;
    std::vector<double> d(d_in, d_in + n), e(n + 1, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
    if (z) {
        std::memset(z, 0, sizeof(double) * (size_t)n * n);
        for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
    }
    std::vector<double> zl;
    if (zlast) {
        zl.assign(n, 0.0);
        zl[n - 1] = 1.0;   // row n-1 of the identity
    }
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd || std::fabs(e[m]) < 1e-300) break;
            }
            if (m != l) {
                if (++iter > 60 * n + 100) return -1;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                bool underflow = false;
                for (i = m - 1; i >= l; --i) {
                    const double f = s * e[i], b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        underflow = true;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    if (z) {
                        double *zi = z + (size_t)i * n, *zi1 = z + (size_t)(i + 1) * n;
                        for (int k = 0; k < n; ++k) {
                            const double t = zi1[k];
                            zi1[k] = s * zi[k] + c * t;
                            zi[k] = c * zi[k] - s * t;
                        }
                    }
                    if (zlast) {
                        const double t = zl[i + 1];
                        zl[i + 1] = s * zl[i] + c * t;
                        zl[i] = c * zl[i] - s * t;
                    }
                }
                if (underflow) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    // ascending order, carrying vectors
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            if (z)
                for (int q = 0; q < n; ++q) std::swap(z[(size_t)i * n + q], z[(size_t)k * n + q]);
            if (zlast) std::swap(zl[i], zl[k]);
        }
    }
    for (int i = 0; i < n; ++i) w[i] = d[i];
    if (zlast)
        for (int i = 0; i < n; ++i) zlast[i] = zl[i];
    return 0;
}

}  // namespace edx
