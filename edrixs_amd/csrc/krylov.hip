// Host drivers of the Krylov solvers (part of edx_kernels.hip).
//
//   edx_build_krylov  <- build_krylov_mp, src/lanczos.f90:210-303
//   edx_minres        <- pminres_csr,     src/linsys.f90:3-128
//   edx_lanczos_eigs  <- diag_ham_lanczos, src/lanczos.f90:6-167, with full
//                        reorthogonalization (classical Gram-Schmidt, two passes)
//                        instead of the eigenvector re-run, and deflated
//                        restarts so degenerate levels are all found.
//
// The Krylov and MINRES recurrences keep their scalars in device memory: one
// iteration is a fixed sequence of launches (SpMV with fused epilogue and dot,
// one-block scalar kernels, fused vector updates) gated by a per-iteration run
// flag, and the host only synchronises every few iterations to see whether the
// recurrence has stopped.

namespace edx {
int tridiag_ql(int n, const double *d, const double *e, double *w, double *z, double *zlast);
int herm_eig_lowest(int n, std::complex<double> *A, double *w, int nvec, std::complex<double> *Z);
}

namespace {

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { edx::dfree(p); }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
    int alloc(size_t bytes, const char *what) {
        p = edx::dmalloc(bytes, what);
        return p ? EDX_OK : EDX_ERR_NOMEM;
    }
};

int check_square(const edx_csr *h, int precision) {
    EDX_CHECK(h, EDX_ERR_ARG, "null matrix");
    EDX_CHECK(h->m == h->n, EDX_ERR_ARG, "matrix is %lld x %lld, not square", (long long)h->m,
              (long long)h->n);
    return edx_csr_prepare(const_cast<edx_csr *>(h), precision);
}

// y = c0*H x + c1*z with optional dot partials; converts x for the C64 path.
int apply(const edx_csr *h, int precision, const double2 *x, float2 *x32, const SpmvEpi &e,
          int *nb, const int *run) {
    if (precision == EDX_C64) {
        EDX_TRY(edx::to_c64(h->n, x, x32, run));
        return edx::spmv_launch(h, precision, x32, e, 0, nb);
    }
    return edx::spmv_launch(h, precision, x, e, 0, nb);
}

// Device memory of one Krylov run beyond its input and outputs: two vectors,
// the dot partials, the device state, the (alpha, beta) history and the
// per-step run flags.  krylov_scratch_bytes gives the size of one block that
// holds them all, carved by krylov_scratch_carve at 256-byte boundaries.
struct KrylovScratch {
    double2 *v;        // 2 n
    double2 *part;     // max(partials_len, vec_blocks(n))
    KrylovDev *S;
    double *ab;        // 2 maxn + 2
    int *run;          // maxn + 2
};

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t krylov_scratch_bytes(int64_t n, int64_t partials_len, int maxn) {
    const int64_t np = std::max<int64_t>(partials_len, vec_blocks(n));
    return align256(2 * (size_t)n * sizeof(double2)) + align256((size_t)np * sizeof(double2)) +
           align256(sizeof(KrylovDev)) + align256((2 * (size_t)maxn + 2) * sizeof(double)) +
           align256(((size_t)maxn + 2) * sizeof(int));
}

KrylovScratch krylov_scratch_carve(char *base, int64_t n, int64_t partials_len, int maxn) {
    KrylovScratch k;
    const int64_t np = std::max<int64_t>(partials_len, vec_blocks(n));
    k.v = (double2 *)base;
    base += align256(2 * (size_t)n * sizeof(double2));
    k.part = (double2 *)base;
    base += align256((size_t)np * sizeof(double2));
    k.S = (KrylovDev *)base;
    base += align256(sizeof(KrylovDev));
    k.ab = (double *)base;
    base += align256((2 * (size_t)maxn + 2) * sizeof(double));
    k.run = (int *)base;
    return k;
}

// Zero nbytes (a multiple of 8, 8-byte aligned) with vector stores; used
// instead of hipMemsetAsync where the caller may capture the launches.
__global__ void k_zero_words(int64_t nwords, uint64_t *p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nwords;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// build_krylov_mp (src/lanczos.f90:210-303) over any product
// prod(x, epilogue, &nblocks) that writes y = c0 A x + c1 z with dot partials;
// the run's device memory comes in k (see KrylovScratch).
template <class Prod>
int krylov_loop_in(int64_t n, Prod prod, const KrylovScratch &k, const double2 *v0, int maxn,
                   int *neff, double *alpha, double *beta, double *norm) {
    int nbv = vec_blocks(n);
    double2 *v[2] = {k.v, k.v + n};
    double *d_alpha = k.ab, *d_beta = k.ab + maxn + 1;
    int *run = k.run;
    KrylovDev *S = k.S;
    hipStream_t s = edx::stream();
    // clear the vectors, the history and the run flags (fill kernels, no memset nodes)
    const int64_t nz[3] = {4 * n, 2 * (int64_t)maxn + 2, ((int64_t)maxn + 2 + 1) / 2};
    uint64_t *pz[3] = {(uint64_t *)k.v, (uint64_t *)k.ab, (uint64_t *)k.run};
    for (int q = 0; q < 3; ++q)
        hipLaunchKernelGGL(k_zero_words, dim3(vec_blocks(nz[q])), dim3(VB), 0, s, nz[q], pz[q]);
    EDX_HIP(hipMemcpyAsync(v[0], v0, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    double2 *part_p = k.part;
    hipLaunchKernelGGL(k_krylov_init, dim3(1), dim3(1), 0, s, S, run);
    // norm = <v0|v0>; v1 = v0 / sqrt(norm)   (src/lanczos.f90:256-262)
    hipLaunchKernelGGL(k_dot, dim3(nbv), dim3(VB), 0, s, n, v[0], v[0], part_p, nullptr);
    hipLaunchKernelGGL(k_krylov_norm, dim3(1), dim3(VB), 0, s, part_p, nbv, S);
    hipLaunchKernelGGL(k_div_real, dim3(nbv), dim3(VB), 0, s, n, v[0], &S->sqrt_norm, nullptr);
    EDX_LAUNCH_CHECK();
    int curr = 0, prev = 1;
    int checked = 0;
    for (int i = 1; i <= maxn; ++i) {
        // prev = -beta_{i-1} prev + H curr ; partial <curr|prev>
        SpmvEpi e{};
        e.y = v[prev];
        e.z = v[prev];
        e.c0 = &S->coef[0];
        e.c1 = &S->coef[1];
        e.dvec = v[curr];
        e.partials = part_p;
        e.run = run + (i - 1);
        int nb = 0;
        EDX_TRY(prod((const double2 *)v[curr], e, &nb, run + (i - 1)));
        hipLaunchKernelGGL(k_krylov_alpha, dim3(1), dim3(VB), 0, s, part_p, nb, S,
                           d_alpha, i, run);
        hipLaunchKernelGGL(k_axpy_norm, dim3(nbv), dim3(VB), 0, s, n, v[prev], v[curr],
                           d_alpha + (i - 1), part_p, run + (i - 1));
        hipLaunchKernelGGL(k_krylov_beta, dim3(1), dim3(VB), 0, s, part_p, nbv, S,
                           d_beta, i, run);
        hipLaunchKernelGGL(k_div_real, dim3(nbv), dim3(VB), 0, s, n, v[prev], d_beta + i, run + i);
        EDX_LAUNCH_CHECK();
        std::swap(curr, prev);
        if (i - checked >= 32 && i < maxn) {   // stop launching once the recurrence broke down
            int r = 1;
            EDX_HIP(edx::copy_sync(&r, run + i, sizeof(int), hipMemcpyDeviceToHost));
            checked = i;
            if (!r) break;
        }
    }
    KrylovDev hs;
    EDX_HIP(edx::copy_sync(&hs, S, sizeof(KrylovDev), hipMemcpyDeviceToHost));
    EDX_HIP(edx::copy_sync(alpha, d_alpha, maxn * sizeof(double), hipMemcpyDeviceToHost));
    // beta_1 .. beta_maxn
    EDX_HIP(edx::copy_sync(beta, d_beta + 1, maxn * sizeof(double), hipMemcpyDeviceToHost));
    *neff = hs.neff;
    *norm = hs.norm;
    return EDX_OK;
}

// krylov_loop_in with its device memory allocated for the run.
template <class Prod>
int krylov_loop(int64_t n, Prod prod, int64_t partials_len, const double2 *v0, int maxn,
                int *neff, double *alpha, double *beta, double *norm) {
    DevBuf buf;
    EDX_TRY(buf.alloc(krylov_scratch_bytes(n, partials_len, maxn), "krylov scratch"));
    const KrylovScratch k = krylov_scratch_carve(buf.as<char>(), n, partials_len, maxn);
    return krylov_loop_in(n, prod, k, v0, maxn, neff, alpha, beta, norm);
}

int krylov_run(const edx_csr *h, int precision, const double2 *v0, int maxn, int *neff,
               double *alpha, double *beta, double *norm) {
    DevBuf x32;
    if (precision == EDX_C64) EDX_TRY(x32.alloc(h->n * sizeof(float2), "krylov x32"));
    float2 *x32p = x32.as<float2>();
    auto prod = [&](const double2 *x, const SpmvEpi &e, int *nb, const int *run) {
        return apply(h, precision, x, x32p, e, nb, run);
    };
    return krylov_loop(h->n, prod, edx::spmv_partials_len(h), v0, maxn, neff, alpha, beta, norm);
}

int minres_run(const edx_csr *a, int precision, const double2 *b, double2 *x, double tol, int maxit,
               int *iters, double *final_res) {
    const int64_t n = a->n;
    DevBuf vb, x32, part, st, runf;
    EDX_TRY(vb.alloc(5 * n * sizeof(double2), "minres vectors"));
    if (precision == EDX_C64) EDX_TRY(x32.alloc(n * sizeof(float2), "minres x32"));
    int nbv = vec_blocks(n);
    EDX_TRY(part.alloc((size_t)std::max<int64_t>(edx::spmv_partials_len(a), nbv) * sizeof(double2),
                       "partials"));
    EDX_TRY(st.alloc(sizeof(MinresDev), "minres state"));
    const int nflags = (maxit > 0 ? maxit : 0) + 2;
    EDX_TRY(runf.alloc((size_t)nflags * sizeof(int), "run flags"));
    double2 *v[2] = {vb.as<double2>(), vb.as<double2>() + n};
    double2 *w[2] = {vb.as<double2>() + 2 * n, vb.as<double2>() + 3 * n};
    double2 *tmp = vb.as<double2>() + 4 * n;
    int *run = runf.as<int>();
    MinresDev *S = st.as<MinresDev>();
    hipStream_t s = edx::stream();
    EDX_HIP(hipMemsetAsync(vb.p, 0, 5 * n * sizeof(double2), s));
    EDX_HIP(hipMemsetAsync(x, 0, n * sizeof(double2), s));
    EDX_HIP(hipMemsetAsync(run, 0, (size_t)nflags * sizeof(int), s));
    EDX_HIP(hipMemcpyAsync(v[0], b, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_dot, dim3(nbv), dim3(VB), 0, s, n, v[0], v[0], part.as<double2>(), nullptr);
    hipLaunchKernelGGL(k_minres_init, dim3(1), dim3(VB), 0, s, part.as<double2>(), nbv, S, run, tol,
                       maxit);
    EDX_LAUNCH_CHECK();
    int curr = 0, prev = 1, checked = 0;
    for (int i = 0; i < maxit; ++i) {
        hipLaunchKernelGGL(k_minres_divbeta, dim3(nbv), dim3(VB), 0, s, n, v[curr], S, run, i);
        SpmvEpi e{};
        e.y = tmp;
        e.dvec = v[curr];
        e.partials = part.as<double2>();
        e.run = run + i;
        int nb = 0;
        EDX_TRY(apply(a, precision, v[curr], x32.as<float2>(), e, &nb, run + i));
        hipLaunchKernelGGL(k_minres_alpha, dim3(1), dim3(VB), 0, s, part.as<double2>(), nb, S, run, i);
        hipLaunchKernelGGL(k_minres_vupdate, dim3(nbv), dim3(VB), 0, s, n, tmp, v[curr], v[prev], S,
                           part.as<double2>(), run, i);
        hipLaunchKernelGGL(k_minres_scalars, dim3(1), dim3(VB), 0, s, part.as<double2>(), nbv, S,
                           run, i);
        hipLaunchKernelGGL(k_minres_wx, dim3(nbv), dim3(VB), 0, s, n, v[curr], w[prev], w[curr], x,
                           S, run, i);
        EDX_LAUNCH_CHECK();
        std::swap(curr, prev);
        if (i + 1 - checked >= 16) {
            int r = 1;
            EDX_HIP(edx::copy_sync(&r, run + i + 1, sizeof(int), hipMemcpyDeviceToHost));
            checked = i + 1;
            if (!r) break;
        }
    }
    // final true residual |A x - b| (src/linsys.f90:117-125)
    SpmvEpi e{};
    e.y = tmp;
    EDX_TRY(apply(a, precision, x, x32.as<float2>(), e, nullptr, nullptr));
    hipLaunchKernelGGL(k_sub_vec, dim3(nbv), dim3(VB), 0, s, n, tmp, b);
    hipLaunchKernelGGL(k_dot, dim3(nbv), dim3(VB), 0, s, n, tmp, tmp, part.as<double2>(), nullptr);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(VB), 0, s, part.as<double2>(), nbv,
                       part.as<double2>() + nbv, nullptr);
    EDX_LAUNCH_CHECK();
    double2 r2;
    EDX_HIP(edx::copy_sync(&r2, part.as<double2>() + nbv, sizeof(double2), hipMemcpyDeviceToHost));
    MinresDev hs;
    EDX_HIP(edx::copy_sync(&hs, S, sizeof(MinresDev), hipMemcpyDeviceToHost));
    if (iters) *iters = hs.iters;
    if (final_res) *final_res = sqrt(fabs(r2.x));
    return EDX_OK;
}

// ------------------------------------------------------------- Lanczos

constexpr int GEMV_NBX = 256;   // row workgroups per column group of the projections

struct Lanczos {
    const edx_csr *h;
    int precision;
    int64_t n;
    double2 *basis;        // columns: [locked vectors | Krylov vectors]
    int64_t ld;
    int cap_cols;
    float2 *x32;
    double2 *w;
    double2 *part;
    double2 *hvec;
    int nbv;
};

int host_dot(Lanczos &L, const double2 *a, const double2 *b, double2 *out) {
    hipLaunchKernelGGL(k_dot, dim3(L.nbv), dim3(VB), 0, edx::stream(), L.n, a, b, L.part, nullptr);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(VB), 0, edx::stream(), L.part, L.nbv,
                       L.part + L.nbv, nullptr);
    EDX_LAUNCH_CHECK();
    EDX_HIP(edx::copy_sync(out, L.part + L.nbv, sizeof(double2), hipMemcpyDeviceToHost));
    return EDX_OK;
}

// w -= Q Q^H w over the first k basis columns; returns the coefficients in hk (host).
// The projections run on up to GEMV_NBX workgroups per group of 8 columns, so a
// basis of a few dozen columns still gives the GPU thousands of workgroups.
int orthogonalize(Lanczos &L, int k, double2 *w, std::vector<double2> *hk) {
    if (k <= 0) return EDX_OK;
    const int nbx = std::min(L.nbv, GEMV_NBX);
    dim3 grid(nbx, (k + 7) / 8);
    hipLaunchKernelGGL(k_gemv_ct, grid, dim3(VB), 0, edx::stream(), L.n, L.basis, L.ld, k, w, L.part);
    hipLaunchKernelGGL(k_reduce_cols, dim3((k + 255) / 256), dim3(256), 0, edx::stream(), L.part,
                       nbx, k, L.hvec);
    hipLaunchKernelGGL(k_gemv_n_sub, dim3(L.nbv), dim3(VB), 0, edx::stream(), L.n, L.basis, L.ld, k,
                       L.hvec, w);
    EDX_LAUNCH_CHECK();
    if (hk) {
        hk->resize(k);
        EDX_HIP(edx::copy_sync(hk->data(), L.hvec, k * sizeof(double2), hipMemcpyDeviceToHost));
    }
    return EDX_OK;
}

// One Lanczos run started from basis column `first` (already normalised and
// orthogonal to the locked columns [0, first)).  Returns Ritz pairs of the run.
// One Lanczos step from basis column first+j: w = H v_j - alpha v_j - beta v_{j-1},
// two classical Gram-Schmidt passes against columns [0, first+j], alpha
// corrected by the v_j component; appends alpha_j and beta_{j+1} = |w|.
int lanczos_one(Lanczos &L, int first, int j, std::vector<double> &al, std::vector<double> &be) {
    hipStream_t s = edx::stream();
    double2 *vj = L.basis + (int64_t)(first + j) * L.ld;
    SpmvEpi e{};
    e.y = L.w;
    EDX_TRY(apply(L.h, L.precision, vj, L.x32, e, nullptr, nullptr));
    double2 a;
    EDX_TRY(host_dot(L, vj, L.w, &a));
    double alpha = a.x;
    const double2 *vprev = j > 0 ? L.basis + (int64_t)(first + j - 1) * L.ld : nullptr;
    hipLaunchKernelGGL(k_axpby_sub, dim3(L.nbv), dim3(VB), 0, s, L.n, L.w, vj, alpha, vprev,
                       j > 0 ? be[j - 1] : 0.0);
    EDX_LAUNCH_CHECK();
    std::vector<double2> h1;
    EDX_TRY(orthogonalize(L, first + j + 1, L.w, &h1));
    EDX_TRY(orthogonalize(L, first + j + 1, L.w, nullptr));
    alpha += h1[first + j].x;
    double2 nb2;
    EDX_TRY(host_dot(L, L.w, L.w, &nb2));
    al.push_back(alpha);
    be.push_back(sqrt(fabs(nb2.x)));
    return EDX_OK;
}

// maxn fully reorthogonalized steps from column first (already normalised);
// stops early when beta < 1e-10 as src/lanczos.f90:290-293 does.
int lanczos_steps(Lanczos &L, int first, int maxn, std::vector<double> &al,
                  std::vector<double> &be, int &iters) {
    const int maxm = std::min(maxn, L.cap_cols - first - 1);
    for (int j = 0; j < maxm; ++j) {
        EDX_TRY(lanczos_one(L, first, j, al, be));
        ++iters;
        const double beta = be.back();
        if (beta < 1e-10) break;
        if (j + 1 < maxm) {
            hipLaunchKernelGGL(k_copy_div, dim3(L.nbv), dim3(VB), 0, edx::stream(), L.n, L.w, beta,
                               L.basis + (int64_t)(first + j + 1) * L.ld);
            EDX_LAUNCH_CHECK();
        }
    }
    return EDX_OK;
}

int lanczos_run(Lanczos &L, int first, int nev, int maxiter, double tol, std::vector<double> &ritz,
                std::vector<double> &S, int &m_out, int &iters) {
    std::vector<double> al, be;   // be[j] couples v_j and v_{j+1}
    const int maxm = std::min(maxiter, L.cap_cols - first - 1);
    int m = 0;
    hipStream_t s = edx::stream();
    for (int j = 0; j < maxm; ++j) {
        EDX_TRY(lanczos_one(L, first, j, al, be));
        const double beta = be.back();
        m = j + 1;
        ++iters;
        bool last = (beta < 1e-10) || (m == maxm);
        bool check = last || (m >= nev && (m % 5 == 0));
        if (check) {
            std::vector<double> w(m), zl(m);
            EDX_CHECK(edx::tridiag_ql(m, al.data(), be.data(), w.data(), nullptr, zl.data()) == 0,
                      EDX_ERR_NUMERIC, "tridiagonal eigensolver did not converge");
            int nconv = 0;
            for (int k = 0; k < std::min(nev, m); ++k)
                if (fabs(beta * zl[k]) < tol) ++nconv;
            if (nconv >= std::min(nev, m) || last) break;
        }
        if (beta < 1e-10) break;
        hipLaunchKernelGGL(k_copy_div, dim3(L.nbv), dim3(VB), 0, s, L.n, L.w, beta,
                           L.basis + (int64_t)(first + j + 1) * L.ld);
        EDX_LAUNCH_CHECK();
    }
    ritz.assign(m, 0.0);
    S.assign((size_t)m * m, 0.0);
    EDX_CHECK(edx::tridiag_ql(m, al.data(), be.data(), ritz.data(), S.data(), nullptr) == 0,
              EDX_ERR_NUMERIC, "tridiagonal eigensolver did not converge");
    m_out = m;
    return EDX_OK;
}

// Thick-restart Lanczos (K. Wu and H. Simon, SIAM J. Matrix Anal. Appl. 22,
// 602 (2000)) over basis columns [first, first + m]: the role of ARPACK's
// implicitly restarted Lanczos in diag_ham_arpack (src/arpack.f90:2-112,
// which='SR').  A cycle extends the basis to m fully reorthogonalized Lanczos
// vectors; the projected matrix T is the arrow matrix of the kept Ritz values
// and their couplings, followed by the tridiagonal of the new steps.  Ritz
// value i has converged when |beta_m y_{m,i}| <= tol max(eps^(2/3), |theta_i|)
// (ARPACK's test).  Unconverged, the cycle keeps the `keep` lowest Ritz vectors
// V y_i as the new first columns, the residual w / beta_m as the next, and
// T = diag(theta) with couplings beta_m y_{m,i}.  At the end the `take` lowest
// Ritz vectors are written to X (n x take) and their values to ritz.
int trl_run(Lanczos &L, int first, int nev, int m, int max_cycles, double tol,
            std::vector<double> &ritz, double2 *X, double2 *tmp, double *dS, int &take,
            int &iters, int &cycles, int &nconv) {
    using cd = std::complex<double>;
    hipStream_t s = edx::stream();
    const int64_t n = L.n;
    m = std::min<int64_t>(m, n - first);
    std::vector<double> T((size_t)m * m, 0.0);
    int k = 0;   // kept Ritz vectors at the front of the run's columns
    const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
    for (cycles = 1;; ++cycles) {
        int me = m;   // columns of this cycle (fewer on an invariant subspace)
        double beta = 0.0;
        for (int j = k; j < m; ++j) {
            double2 *vj = L.basis + (int64_t)(first + j) * L.ld;
            SpmvEpi e{};
            e.y = L.w;
            EDX_TRY(apply(L.h, L.precision, vj, L.x32, e, nullptr, nullptr));
            double2 a;
            EDX_TRY(host_dot(L, vj, L.w, &a));
            double alpha = a.x;
            // three-term part inside a tridiagonal stretch; after a restart the
            // arrow couplings go with the Gram-Schmidt passes
            const bool tri = j > k;
            const double2 *vprev = tri ? L.basis + (int64_t)(first + j - 1) * L.ld : nullptr;
            hipLaunchKernelGGL(k_axpby_sub, dim3(L.nbv), dim3(VB), 0, s, n, L.w, vj, alpha, vprev,
                               tri ? T[(size_t)(j - 1) * m + j] : 0.0);
            EDX_LAUNCH_CHECK();
            std::vector<double2> h1;
            EDX_TRY(orthogonalize(L, first + j + 1, L.w, &h1));
            EDX_TRY(orthogonalize(L, first + j + 1, L.w, nullptr));
            alpha += h1[first + j].x;
            double2 nb2;
            EDX_TRY(host_dot(L, L.w, L.w, &nb2));
            beta = sqrt(fabs(nb2.x));
            ++iters;
            T[(size_t)j * m + j] = alpha;
            if (beta < 1e-10) {
                me = j + 1;
                break;
            }
            if (j + 1 < m) {
                T[(size_t)j * m + j + 1] = T[(size_t)(j + 1) * m + j] = beta;
                hipLaunchKernelGGL(k_copy_div, dim3(L.nbv), dim3(VB), 0, s, n, L.w, beta,
                                   L.basis + (int64_t)(first + j + 1) * L.ld);
                EDX_LAUNCH_CHECK();
            }
        }
        // Ritz pairs of the me x me projected matrix
        std::vector<cd> A((size_t)me * me), Z((size_t)me * me);
        for (int i = 0; i < me; ++i)
            for (int j = 0; j < me; ++j) A[(size_t)i * me + j] = T[(size_t)i * m + j];
        ritz.assign(me, 0.0);
        EDX_CHECK(edx::herm_eig_lowest(me, A.data(), ritz.data(), me, Z.data()) == 0, EDX_ERR_NUMERIC,
                  "projected eigenproblem did not converge");
        const int want = std::min(nev, me);
        nconv = 0;
        for (int i = 0; i < want; ++i)
            if (fabs(beta * Z[(size_t)(me - 1) * me + i].real()) <=
                tol * std::max(eps23, fabs(ritz[i])))
                ++nconv;
        const bool done = nconv >= want || beta < 1e-10 || cycles >= max_cycles || me < m;
        // keep: the wanted values and half the rest of the cycle (at least one new vector)
        const int keep = done ? want : std::min(m - 1, want + (m - want) / 2);
        std::vector<double> S((size_t)keep * me);   // column i: Ritz vector i's coefficients
        for (int i = 0; i < keep; ++i)
            for (int c = 0; c < me; ++c) S[(size_t)i * me + c] = Z[(size_t)c * me + i].real();
        EDX_HIP(edx::copy_sync(dS, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice));
        const double2 *V = L.basis + (int64_t)first * L.ld;
        if (done) {
            hipLaunchKernelGGL(k_combine, dim3(L.nbv), dim3(VB), 0, s, n, V, L.ld, me, dS, keep, X, n);
            EDX_LAUNCH_CHECK();
            EDX_HIP(hipStreamSynchronize(s));
            take = keep;
            return EDX_OK;
        }
        hipLaunchKernelGGL(k_combine, dim3(L.nbv), dim3(VB), 0, s, n, V, L.ld, me, dS, keep, tmp, n);
        EDX_LAUNCH_CHECK();
        EDX_HIP(hipMemcpyAsync(L.basis + (int64_t)first * L.ld, tmp, (size_t)n * keep * sizeof(double2),
                               hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_copy_div, dim3(L.nbv), dim3(VB), 0, s, n, L.w, beta,
                           L.basis + (int64_t)(first + keep) * L.ld);
        EDX_LAUNCH_CHECK();
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < keep; ++i) {
            T[(size_t)i * m + i] = ritz[i];
            T[(size_t)i * m + keep] = T[(size_t)keep * m + i] = beta * Z[(size_t)(me - 1) * me + i].real();
        }
        k = keep;
    }
}

}  // namespace

extern "C" {

int edx_build_krylov(const edx_csr *h, int precision, const void *d_v0, int maxn, int *neff,
                     double *alpha, double *beta, double *norm) {
    EDX_TRY(check_square(h, precision));
    EDX_CHECK(d_v0 && neff && alpha && beta && norm && maxn >= 1, EDX_ERR_ARG, "bad argument");
    int rc = krylov_run(h, precision, (const double2 *)d_v0, maxn, neff, alpha, beta, norm);
    if (rc) return rc;
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    return EDX_OK;
}

int edx_minres(const edx_csr *a, int precision, const void *d_b, void *d_x, double tol, int maxit,
               int *iters, double *final_res) {
    EDX_TRY(check_square(a, precision));
    EDX_CHECK(d_b && d_x && maxit >= 0, EDX_ERR_ARG, "bad argument");
    int rc = minres_run(a, precision, (const double2 *)d_b, (double2 *)d_x, tol, maxit, iters,
                        final_res);
    if (rc) return rc;
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    return EDX_OK;
}

}  // extern "C"

namespace {

__global__ void k_sum_columns(int64_t n, const double2 *S, int64_t ld, int k, double2 *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double2 s = make_double2(0, 0);
        for (int c = 0; c < k; ++c) s = cadd(s, S[(int64_t)c * ld + i]);
        out[i] = s;
    }
}

// Lowest nev eigenpairs.  ncv == 0: one Lanczos run of up to maxiter steps
// per round; ncv > 0: thick-restart runs (trl_run) over a basis of m >= ncv
// columns and at most maxiter cycles.  With deflate, further rounds start from
// random vectors orthogonal to the locked pairs until a round finds nothing
// below the nev-th value (degenerate levels a single start vector misses).
int lanczos_impl(const edx_csr *h, int precision, int nev, int maxiter, double tol, uint64_t seed,
                 int deflate, int nvec, double *evals, void *d_evecs, int *niter_out,
                 const double2 *start, int nstart, int ncv = 0, int *nconv_out = nullptr) {
    EDX_TRY(check_square(h, precision));
    EDX_CHECK(nev >= 1 && maxiter >= 1 && evals && nvec >= 0 && nvec <= nev &&
                  (nvec == 0 || d_evecs) && ncv >= 0,
              EDX_ERR_ARG, "bad argument (nev=%d maxiter=%d nvec=%d ncv=%d)", nev, maxiter, nvec, ncv);
    const int64_t n = h->n;
    if (nev > n) nev = (int)n;
    if (nvec > nev) nvec = nev;
    // thick restart: at least 2 nev + 16 columns, so a cycle adds enough new directions
    // for clustered spectra (ARPACK's default ncv = nev + 2 stalls there)
    const int m_trl = ncv > 0 ? (int)std::min<int64_t>(n, std::max(ncv, 2 * nev + 16)) : 0;
    if (!ncv && maxiter > n) maxiter = (int)n;
    const int max_rounds = deflate ? 2 * nev + 2 : 1;
    const int cap_cols = nev + (ncv ? m_trl : maxiter) + 2;
    DevBuf basis, x32, w, part, hv, tmpS, evec_all, trl_tmp;
    EDX_TRY(basis.alloc((size_t)n * cap_cols * sizeof(double2), "lanczos basis"));
    if (precision == EDX_C64) EDX_TRY(x32.alloc(n * sizeof(float2), "lanczos x32"));
    EDX_TRY(w.alloc(n * sizeof(double2), "lanczos w"));
    const int nbv = vec_blocks(n);
    EDX_TRY(part.alloc(((size_t)cap_cols * GEMV_NBX + nbv + 64) * sizeof(double2), "partials"));
    EDX_TRY(hv.alloc((size_t)cap_cols * sizeof(double2), "coefficients"));
    EDX_TRY(tmpS.alloc((size_t)cap_cols * cap_cols * sizeof(double), "ritz vectors"));
    EDX_TRY(evec_all.alloc((size_t)n * nev * sizeof(double2), "eigenvectors"));
    if (ncv) EDX_TRY(trl_tmp.alloc((size_t)n * m_trl * sizeof(double2), "thick-restart vectors"));
    Lanczos L{h, precision, n, basis.as<double2>(), n, cap_cols, x32.as<float2>(), w.as<double2>(),
              part.as<double2>(), hv.as<double2>(), nbv};
    hipStream_t s = edx::stream();
    std::vector<double> locked_vals;   // eigenvalues of locked columns [0, nlock)
    int nlock = 0, iters = 0, nconv = 0;
    for (int round = 0; round < max_rounds; ++round) {
        double2 *v0 = L.basis + (int64_t)nlock * n;
        if (round == 0 && start)
            hipLaunchKernelGGL(k_sum_columns, dim3(nbv), dim3(VB), 0, s, n, start, n, nstart, v0);
        else
            hipLaunchKernelGGL(k_random_real, dim3(nbv), dim3(VB), 0, s, n,
                               seed + 0x9e37ull * (uint64_t)round, v0);
        EDX_LAUNCH_CHECK();
        EDX_TRY(orthogonalize(L, nlock, v0, nullptr));
        EDX_TRY(orthogonalize(L, nlock, v0, nullptr));
        double2 nn;
        EDX_TRY(host_dot(L, v0, v0, &nn));
        if (sqrt(nn.x) < 1e-12) break;   // nothing left outside the locked space
        hipLaunchKernelGGL(k_scale_real, dim3(nbv), dim3(VB), 0, s, n, v0, 1.0 / sqrt(nn.x));
        std::vector<double> ritz, S;
        int take = 0;
        if (ncv) {
            int cycles = 0, nc = 0;
            EDX_TRY(trl_run(L, nlock, nev, m_trl, maxiter, tol, ritz, evec_all.as<double2>(),
                            trl_tmp.as<double2>(), tmpS.as<double>(), take, iters, cycles, nc));
            if (round == 0) nconv = nc;
        } else {
            int m = 0;
            EDX_TRY(lanczos_run(L, nlock, nev, maxiter, tol, ritz, S, m, iters));
            // Ritz vectors of this run for its lowest values
            take = std::min(nev, m);
            std::vector<double> Sk((size_t)m * take);
            for (int j = 0; j < take; ++j)
                for (int c = 0; c < m; ++c) Sk[(size_t)j * m + c] = S[(size_t)j * m + c];
            EDX_HIP(edx::copy_sync(tmpS.p, Sk.data(), Sk.size() * sizeof(double), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_combine, dim3(nbv), dim3(VB), 0, s, n, L.basis + (int64_t)nlock * n,
                               n, m, tmpS.as<double>(), take, evec_all.as<double2>(), n);
            EDX_LAUNCH_CHECK();
        }
        // merge: locked (already lowest of previous rounds) with this run's Ritz values
        std::vector<std::pair<double, int>> cand;   // (value, source: <0 locked k, >=0 new j)
        for (int k = 0; k < nlock; ++k) cand.push_back({locked_vals[k], -1 - k});
        for (int j = 0; j < take; ++j) cand.push_back({ritz[j], j});
        std::stable_sort(cand.begin(), cand.end(),
                         [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                             return a.first < b.first;
                         });
        const int keep = std::min(nev, (int)cand.size());
        int new_taken = 0;
        for (int q = 0; q < keep; ++q)
            if (cand[q].second >= 0) ++new_taken;
        if (round > 0 && new_taken == 0) break;   // no missed level below the nev-th
        // rebuild locked set [0, keep) in value order
        DevBuf keepbuf;
        EDX_TRY(keepbuf.alloc((size_t)n * keep * sizeof(double2), "locked copy"));
        for (int q = 0; q < keep; ++q) {
            const double2 *src = cand[q].second >= 0
                                     ? evec_all.as<double2>() + (int64_t)cand[q].second * n
                                     : L.basis + (int64_t)(-1 - cand[q].second) * n;
            EDX_HIP(hipMemcpyAsync(keepbuf.as<double2>() + (int64_t)q * n, src, n * sizeof(double2),
                                   hipMemcpyDeviceToDevice, s));
        }
        EDX_HIP(hipMemcpyAsync(L.basis, keepbuf.p, (size_t)n * keep * sizeof(double2),
                               hipMemcpyDeviceToDevice, s));
        locked_vals.assign(keep, 0.0);
        for (int q = 0; q < keep; ++q) locked_vals[q] = cand[q].first;
        nlock = keep;
        if (!deflate) break;
        if (nlock + 2 >= cap_cols || nlock >= n) break;
    }
    for (int k = 0; k < nev; ++k) evals[k] = k < nlock ? locked_vals[k] : 0.0;
    if (nvec > 0)
        EDX_HIP(hipMemcpyAsync(d_evecs, L.basis, (size_t)n * std::min(nvec, nlock) * sizeof(double2),
                               hipMemcpyDeviceToDevice, s));
    EDX_HIP(hipStreamSynchronize(s));
    if (niter_out) *niter_out = iters;
    if (nconv_out) *nconv_out = nconv;
    return EDX_OK;
}

// The reference's two-pass procedure (diag_ham_lanczos, src/lanczos.f90:6-167)
// without a stored basis: pass 1 runs the three-term recurrence with no
// reorthogonalization from the seeded start vector, keeps alpha and beta, and
// every 10 steps (and at the last) tests the nev lowest Ritz pairs of the
// tridiagonal matrix by their residual estimates |beta_m y_{m,k}| < tol;
// pass 2 runs the same recurrence again from the same start and accumulates
// x_k = sum_j y_{j,k} v_j.  Three vectors of n plus nvec outputs; one copy of
// each level (the Krylov space of one vector holds one).
int lanczos_rerun_impl(const edx_csr *h, int precision, int nev, int maxiter, double tol,
                       uint64_t seed, int nvec, double *evals, void *d_evecs, int *niter_out) {
    EDX_TRY(check_square(h, precision));
    EDX_CHECK(nev >= 1 && maxiter >= 1 && evals && nvec >= 0 && nvec <= nev &&
                  (nvec == 0 || d_evecs),
              EDX_ERR_ARG, "bad argument (nev=%d maxiter=%d nvec=%d)", nev, maxiter, nvec);
    const int64_t n = h->n;
    if (nev > n) nev = (int)n;
    if (nvec > nev) nvec = nev;
    if (maxiter > n) maxiter = (int)n;
    const int nbv = vec_blocks(n);
    DevBuf vb, x32, part;
    EDX_TRY(vb.alloc((size_t)n * 3 * sizeof(double2), "lanczos vectors"));
    if (precision == EDX_C64) EDX_TRY(x32.alloc(n * sizeof(float2), "lanczos x32"));
    EDX_TRY(part.alloc(((size_t)nbv + 64) * sizeof(double2), "partials"));
    Lanczos L{h, precision, n, nullptr, n, 0, x32.as<float2>(), nullptr, part.as<double2>(), nullptr,
              nbv};
    hipStream_t s = edx::stream();
    double2 *buf[3] = {vb.as<double2>(), vb.as<double2>() + n, vb.as<double2>() + 2 * n};
    double2 *X = (double2 *)d_evecs;
    // v_1: the seeded start vector, normalised
    auto start = [&](double2 *v) -> int {
        hipLaunchKernelGGL(k_random_real, dim3(nbv), dim3(VB), 0, s, n, seed, v);
        EDX_LAUNCH_CHECK();
        double2 nn;
        EDX_TRY(host_dot(L, v, v, &nn));
        EDX_CHECK(nn.x > 0, EDX_ERR_NUMERIC, "zero start vector");
        hipLaunchKernelGGL(k_scale_real, dim3(nbv), dim3(VB), 0, s, n, v, 1.0 / sqrt(nn.x));
        EDX_LAUNCH_CHECK();
        return EDX_OK;
    };
    // w = H v_j - alpha_j v_j - beta_{j-1} v_{j-1}; alpha from the device when
    // known < 0, else given
    auto step = [&](double2 *vprev, double2 *vj, double2 *w, double bprev, double &alpha,
                    bool measure) -> int {
        SpmvEpi e{};
        e.y = w;
        EDX_TRY(apply(L.h, L.precision, vj, L.x32, e, nullptr, nullptr));
        if (measure) {
            double2 a;
            EDX_TRY(host_dot(L, vj, w, &a));
            alpha = a.x;
        }
        hipLaunchKernelGGL(k_axpby_sub, dim3(nbv), dim3(VB), 0, s, n, w, vj, alpha, vprev, bprev);
        EDX_LAUNCH_CHECK();
        return EDX_OK;
    };
    std::vector<double> al, be, ritz;
    int m = 0;
    {   // pass 1
        EDX_TRY(start(buf[1]));
        double2 *v[3] = {buf[0], buf[1], buf[2]};   // v[0] = v_{j-1}, v[1] = v_j, v[2] = w
        for (int j = 0; j < maxiter; ++j) {
            double alpha = 0.0;
            EDX_TRY(step(j > 0 ? v[0] : nullptr, v[1], v[2], j > 0 ? be[j - 1] : 0.0, alpha, true));
            double2 nb2;
            EDX_TRY(host_dot(L, v[2], v[2], &nb2));
            const double beta = sqrt(fabs(nb2.x));
            al.push_back(alpha);
            be.push_back(beta);
            m = j + 1;
            const bool last = beta < 1e-10 || m == maxiter;
            if (last || (m >= nev && m % 10 == 0)) {
                std::vector<double> wv(m), zl(m);
                EDX_CHECK(edx::tridiag_ql(m, al.data(), be.data(), wv.data(), nullptr, zl.data()) == 0,
                          EDX_ERR_NUMERIC, "tridiagonal eigensolver did not converge");
                int nconv = 0;
                for (int k = 0; k < std::min(nev, m); ++k)
                    if (fabs(beta * zl[k]) < tol) ++nconv;
                if (nconv >= std::min(nev, m) || last) break;
            }
            // v_{j+1} = w / beta over v_{j-1}, which the next step no longer needs
            hipLaunchKernelGGL(k_copy_div, dim3(nbv), dim3(VB), 0, s, n, v[2], beta, v[0]);
            EDX_LAUNCH_CHECK();
            std::swap(v[0], v[1]);
        }
    }
    ritz.assign(m, 0.0);
    std::vector<double> S((size_t)m * m);   // row k: eigenvector k of the tridiagonal matrix
    EDX_CHECK(edx::tridiag_ql(m, al.data(), be.data(), ritz.data(), S.data(), nullptr) == 0,
              EDX_ERR_NUMERIC, "tridiagonal eigensolver did not converge");
    for (int k = 0; k < nev; ++k) evals[k] = k < m ? ritz[k] : 0.0;
    if (nvec > 0) {   // pass 2
        EDX_HIP(hipMemsetAsync(X, 0, (size_t)n * nvec * sizeof(double2), s));
        EDX_TRY(start(buf[1]));
        double2 *v[3] = {buf[0], buf[1], buf[2]};   // v[0] = v_{j-1}, v[1] = v_j, v[2] = w
        for (int j = 0; j < m; ++j) {
            for (int k = 0; k < nvec && k < m; ++k)   // x_k += y_{j,k} v_j
                hipLaunchKernelGGL(k_axpby_sub, dim3(nbv), dim3(VB), 0, s, n, X + (int64_t)k * n, v[1],
                                   -S[(size_t)k * m + j], nullptr, 0.0);
            EDX_LAUNCH_CHECK();
            if (j + 1 == m) break;
            double alpha = al[j];
            EDX_TRY(step(j > 0 ? v[0] : nullptr, v[1], v[2], j > 0 ? be[j - 1] : 0.0, alpha, false));
            hipLaunchKernelGGL(k_copy_div, dim3(nbv), dim3(VB), 0, s, n, v[2], be[j], v[0]);
            EDX_LAUNCH_CHECK();
            std::swap(v[0], v[1]);   // v_j becomes v_{j-1}, the new vector v_{j+1} sits in v[1]
        }
        for (int k = 0; k < nvec && k < m; ++k) {   // unit norm
            double2 nn;
            EDX_TRY(host_dot(L, X + (int64_t)k * n, X + (int64_t)k * n, &nn));
            if (nn.x > 0)
                hipLaunchKernelGGL(k_scale_real, dim3(nbv), dim3(VB), 0, s, n, X + (int64_t)k * n,
                                   1.0 / sqrt(nn.x));
        }
        EDX_LAUNCH_CHECK();
    }
    EDX_HIP(hipStreamSynchronize(s));
    if (niter_out) *niter_out = m;
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_lanczos_eigs(const edx_csr *h, int precision, int nev, int maxiter, double tol,
                     uint64_t seed, int deflate, int nvec, double *evals, void *d_evecs,
                     int *niter_out) {
    return lanczos_impl(h, precision, nev, maxiter, tol, seed, deflate, nvec, evals, d_evecs,
                        niter_out, nullptr, 0);
}

int edx_lanczos_ground(const edx_csr *h, int precision, int nev, int maxiter, double tol,
                       uint64_t seed, int nvec, double *evals, void *d_evecs, int *niter_out) {
    return lanczos_rerun_impl(h, precision, nev, maxiter, tol, seed, nvec, evals, d_evecs, niter_out);
}

int edx_trlan_eigs(const edx_csr *h, int precision, int nev, int ncv, int maxiter, double tol,
                   uint64_t seed, int deflate, int nvec, double *evals, void *d_evecs,
                   int *nmatvec_out, int *nconv_out) {
    EDX_CHECK(ncv >= 1, EDX_ERR_ARG, "ncv must be positive (got %d)", ncv);
    return lanczos_impl(h, precision, nev, maxiter, tol, seed, deflate, nvec, evals, d_evecs,
                        nmatvec_out, nullptr, 0, ncv, nconv_out);
}

int edx_build_krylov_reorth(const edx_csr *h, int precision, const void *d_v0, int maxn,
                            int *neff, double *alpha, double *beta, double *norm) {
    EDX_TRY(check_square(h, precision));
    EDX_CHECK(d_v0 && neff && alpha && beta && norm && maxn >= 1, EDX_ERR_ARG, "bad argument");
    const int64_t n = h->n;
    if (maxn > n) maxn = (int)n;
    const int cap_cols = maxn + 2;
    DevBuf basis, x32, w, part, hv;
    EDX_TRY(basis.alloc((size_t)n * cap_cols * sizeof(double2), "krylov basis"));
    if (precision == EDX_C64) EDX_TRY(x32.alloc(n * sizeof(float2), "krylov x32"));
    EDX_TRY(w.alloc(n * sizeof(double2), "krylov w"));
    const int nbv = vec_blocks(n);
    EDX_TRY(part.alloc(((size_t)cap_cols * GEMV_NBX + nbv + 64) * sizeof(double2), "partials"));
    EDX_TRY(hv.alloc((size_t)cap_cols * sizeof(double2), "coefficients"));
    Lanczos L{h, precision, n, basis.as<double2>(), n, cap_cols, x32.as<float2>(), w.as<double2>(),
              part.as<double2>(), hv.as<double2>(), nbv};
    hipStream_t s = edx::stream();
    EDX_HIP(hipMemcpyAsync(L.basis, d_v0, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    double2 nn;
    EDX_TRY(host_dot(L, L.basis, L.basis, &nn));
    *norm = nn.x;
    EDX_CHECK(nn.x > 0, EDX_ERR_NUMERIC, "zero start vector");
    hipLaunchKernelGGL(k_scale_real, dim3(nbv), dim3(VB), 0, s, n, L.basis, 1.0 / sqrt(nn.x));
    std::vector<double> al, be;
    int iters = 0;
    EDX_TRY(lanczos_steps(L, 0, maxn, al, be, iters));
    *neff = (int)al.size();
    for (int i = 0; i < *neff; ++i) {
        alpha[i] = al[i];
        beta[i] = be[i];
    }
    EDX_HIP(hipStreamSynchronize(s));
    return EDX_OK;
}

int edx_lanczos_refine(const edx_csr *h, int precision, int nev, int maxiter, double tol,
                       int nstart, const void *d_start, int nvec, double *evals, void *d_evecs,
                       int *niter_out) {
    EDX_CHECK(d_start && nstart >= 1, EDX_ERR_ARG, "refinement needs start vectors");
    return lanczos_impl(h, precision, nev, maxiter, tol, 0, 0, nvec, evals, d_evecs, niter_out,
                        (const double2 *)d_start, nstart);
}

}  // extern "C"

// ------------------------------------------------ the no-allocation variant
//
// edx_build_krylov_noalloc runs the recurrence of edx_build_krylov with its
// device memory (vectors, partials, state, history, run flags) carved from
// `scratch`, a file-scope device array of NOALLOC_SCRATCH_BYTES that every GPU
// the library runs on holds until the process ends.  In host code `scratch`
// names a host-side stand-in, not device memory: its device address is
// resolved once per device with hipGetSymbolAddress (with that device
// current) and only that address is handed to kernels and copies.  Calls on
// one device share the array, so calls that overlap on one device race.

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
__device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_ptr;   // device -> device address of scratch

int noalloc_scratch(char **out) {
    const int dev = edx::current_device();
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        int cur = -1;
        EDX_HIP(hipGetDevice(&cur));
        EDX_CHECK(cur == dev, EDX_ERR_ARG, "device %d is not current (%d is)", dev, cur);
        char *ptr = nullptr;
        EDX_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
        EDX_CHECK(ptr, EDX_ERR_HIP, "no device address for the scratch array");
        it = g_scratch_ptr.emplace(dev, ptr).first;
    }
    *out = it->second;
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_build_krylov_noalloc(const edx_csr *h, int precision, const void *d_v0, int maxn,
                             int *neff, double *alpha, double *beta, double *norm) {
    EDX_CHECK(h && d_v0 && neff && alpha && beta && norm && maxn >= 1, EDX_ERR_ARG, "bad argument");
    EDX_CHECK(h->m == h->n, EDX_ERR_ARG, "matrix is %lld x %lld, not square", (long long)h->m,
              (long long)h->n);
    // preparing a form (and tuning its schedule) allocates: it must come first
    EDX_CHECK(precision == EDX_C128 || (precision == EDX_C64D && h->d_diag) ||
                  (precision == EDX_C128I && h->d_off_idx),
              EDX_ERR_ARG,
              "precision %d is not prepared on this matrix (call edx_csr_prepare first; the "
              "complex64-vector form EDX_C64 is not offered here)", precision);
    const int64_t plen = edx::spmv_partials_len(h);
    const size_t need = krylov_scratch_bytes(h->n, plen, maxn);
    EDX_CHECK(need <= NOALLOC_SCRATCH_BYTES, EDX_ERR_CAPACITY,
              "a Krylov run over %lld rows with maxn %d needs %zu bytes of scratch; the static "
              "buffer holds %zu", (long long)h->n, maxn, need, NOALLOC_SCRATCH_BYTES);
    char *base = nullptr;
    EDX_TRY(noalloc_scratch(&base));
    const KrylovScratch k = krylov_scratch_carve(base, h->n, plen, maxn);
    auto prod = [&](const double2 *x, const SpmvEpi &e, int *nb, const int *) {
        return edx::spmv_launch(h, precision, x, e, 0, nb);
    };
    EDX_TRY(krylov_loop_in(h->n, prod, k, (const double2 *)d_v0, maxn, neff, alpha, beta, norm));
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    return EDX_OK;
}

}  // extern "C"
