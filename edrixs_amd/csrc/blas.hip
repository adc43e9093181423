// Vector kernels for the Krylov solvers (part of edx_kernels.hip).
//
// Every reduction goes through per-block partial sums reduced by one block in
// a fixed order, so results do not depend on scheduling.  Scalars produced by
// one kernel and consumed by the next stay in device memory; kernels can skip
// themselves on a device flag, so a solver iteration is a fixed launch
// sequence with no host round trip.

namespace {

constexpr int VB = 256;   // threads per block for vector kernels

int vec_blocks(int64_t n) {
    int64_t g = (n + VB - 1) / VB;
    int64_t cap = (int64_t)edx::num_cus() * 4;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// sum of partials[0, nb) by one block; returns the total in thread 0
__device__ double2 reduce_partials(const double2 *partials, int nb) {
    __shared__ double sx[VB], sy[VB];
    double2 s = make_double2(0, 0);
    for (int i = threadIdx.x; i < nb; i += VB) s = cadd(s, partials[i]);
    sx[threadIdx.x] = s.x;
    sy[threadIdx.x] = s.y;
    __syncthreads();
    for (int off = VB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sx[threadIdx.x] += sx[threadIdx.x + off];
            sy[threadIdx.x] += sy[threadIdx.x + off];
        }
        __syncthreads();
    }
    return make_double2(sx[0], sy[0]);
}

// partials[b] = sum over block b's indices of conj(a) * b
__global__ __launch_bounds__(VB) void k_dot(int64_t n, const double2 *a, const double2 *b,
                                            double2 *partials, const int *run) {
    if (run && *run == 0) return;
    double2 s = make_double2(0, 0);
    for (int64_t i = blockIdx.x * (int64_t)VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
        s = cadd(s, cconjmul(a[i], b[i]));
    s = block_sum_256(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(VB) void k_reduce_to(const double2 *partials, int nb, double2 *out,
                                                  const int *run) {
    if (run && *run == 0) return;
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) *out = s;
}

// y = y / s  (s real, read from device)
__global__ __launch_bounds__(VB) void k_div_real(int64_t n, double2 *y, const double *s,
                                                 const int *run) {
    if (run && *run == 0) return;
    const double d = *s;
    for (int64_t i = blockIdx.x * (int64_t)VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
        y[i] = make_double2(y[i].x / d, y[i].y / d);
}

// y = y - a * x (a real, device) and partials of |y|^2
__global__ __launch_bounds__(VB) void k_axpy_norm(int64_t n, double2 *y, const double2 *x,
                                                  const double *a, double2 *partials,
                                                  const int *run) {
    if (run && *run == 0) return;
    const double al = *a;
    double2 s = make_double2(0, 0);
    for (int64_t i = blockIdx.x * (int64_t)VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
        double2 v = make_double2(y[i].x - al * x[i].x, y[i].y - al * x[i].y);
        y[i] = v;
        s.x += v.x * v.x + v.y * v.y;
    }
    s = block_sum_256(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ------------------------------------------------ Krylov (build_krylov_mp)

struct KrylovDev {
    double2 coef[2];   // SpMV epilogue: y = coef0*A x + coef1*z
    double2 sum;       // scratch for a reduced scalar
    double norm;       // <v0|v0>
    double sqrt_norm;
    double scur;       // scale of the current vector when it is kept unnormalized
    double acoef;      // alpha_i / scur
    int neff;
};

__global__ void k_krylov_init(KrylovDev *st, int *run) {
    st->coef[0] = make_double2(1.0, 0.0);
    st->coef[1] = make_double2(-0.0, 0.0);
    st->scur = 1.0;
    st->acoef = 0.0;
    st->neff = 0;
    run[0] = 1;
}

__global__ __launch_bounds__(VB) void k_krylov_norm(const double2 *partials, int nb, KrylovDev *st) {
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) {
        st->norm = s.x;
        st->sqrt_norm = sqrt(s.x);
    }
}

// alpha_i = Re <v_i | w>   (src/lanczos.f90:277-281)
__global__ __launch_bounds__(VB) void k_krylov_alpha(const double2 *partials, int nb,
                                                     KrylovDev *st, double *alpha, int i,
                                                     const int *run) {
    if (!run[i - 1]) return;
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) {
        alpha[i - 1] = s.x;
        st->neff = i;
    }
}

// beta_i = sqrt(<w|w>); stop when beta_i < 1e-10 (src/lanczos.f90:285-293)
__global__ __launch_bounds__(VB) void k_krylov_beta(const double2 *partials, int nb,
                                                    KrylovDev *st, double *beta, int i, int *run) {
    if (!run[i - 1]) {
        if (threadIdx.x == 0) run[i] = 0;
        return;
    }
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) {
        const double b = sqrt(s.x);
        beta[i] = b;
        st->coef[1] = make_double2(-b, -0.0);
        run[i] = fabs(b) < 1e-10 ? 0 : 1;
    }
}

// -------------------------------------------------------- MINRES (linsys)

struct MinresDev {
    double beta0, beta1;
    double2 sigma0, sigma1, gamma0, gamma1, eta, res;
    double2 alpha, rho1, rho2, rho3, inv_rho1, xcoef;
    double tol;
    int maxit;
    int iters;
};

__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    const double den = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

// principal complex square root
__device__ __forceinline__ double2 csqrt_dev(double2 z) {
    const double r = hypot(z.x, z.y);
    if (r == 0.0) return make_double2(0.0, 0.0);
    double t = sqrt(0.5 * (r + fabs(z.x)));
    if (z.x >= 0.0) return make_double2(t, z.y / (2.0 * t));
    return make_double2(fabs(z.y) / (2.0 * t), z.y >= 0 ? t : -t);
}

__global__ __launch_bounds__(VB) void k_minres_init(const double2 *partials, int nb, MinresDev *st,
                                                    int *run, double tol, int maxit) {
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) {
        st->beta0 = sqrt(s.x);
        st->beta1 = 0.0;
        st->sigma0 = st->sigma1 = make_double2(0, 0);
        st->gamma0 = st->gamma1 = make_double2(1, 0);
        st->eta = make_double2(st->beta0, 0);
        st->res = make_double2(st->beta0, 0);
        st->tol = tol;
        st->maxit = maxit;
        st->iters = 0;
        run[0] = (hypot(st->res.x, st->res.y) > tol && 0 < maxit) ? 1 : 0;   // src/linsys.f90:76
    }
}

__global__ void k_minres_divbeta(int64_t n, double2 *v, const MinresDev *st, const int *run, int i) {
    if (!run[i]) return;
    const double d = st->beta0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        v[k] = make_double2(v[k].x / d, v[k].y / d);
}

__global__ __launch_bounds__(VB) void k_minres_alpha(const double2 *partials, int nb, MinresDev *st,
                                                     const int *run, int i) {
    if (!run[i]) return;
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) st->alpha = s;
}

// v_prev = tmp - alpha v_curr - beta0 v_prev ; partial |v_prev|^2  (src/linsys.f90:87)
__global__ __launch_bounds__(VB) void k_minres_vupdate(int64_t n, const double2 *tmp,
                                                       const double2 *vc, double2 *vp,
                                                       const MinresDev *st, double2 *partials,
                                                       const int *run, int i) {
    if (!run[i]) return;
    const double2 al = st->alpha;
    const double b0 = st->beta0;
    double2 s = make_double2(0, 0);
    for (int64_t k = blockIdx.x * (int64_t)VB + threadIdx.x; k < n; k += (int64_t)gridDim.x * VB) {
        double2 v = csub(csub(tmp[k], cmul(al, vc[k])), cscale(vp[k], b0));
        vp[k] = v;
        s.x += v.x * v.x + v.y * v.y;
    }
    s = block_sum_256(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Givens update of src/linsys.f90:88-110 (complex-symmetric rotation)
__global__ __launch_bounds__(VB) void k_minres_scalars(const double2 *partials, int nb,
                                                       MinresDev *st, int *run, int i) {
    if (!run[i]) {
        if (threadIdx.x == 0) run[i + 1] = 0;
        return;
    }
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) {
        const double beta1 = sqrt(s.x);
        const double beta0 = st->beta0;
        const double2 alpha = st->alpha;
        const double2 delta = csub(cmul(st->gamma0, alpha), cscale(cmul(st->gamma1, st->sigma0), beta0));
        const double2 rho1 = csqrt_dev(cadd(cmul(delta, delta), make_double2(beta1 * beta1, 0.0)));
        const double2 rho2 = cadd(cmul(st->sigma0, alpha), cscale(cmul(st->gamma1, st->gamma0), beta0));
        const double2 rho3 = cscale(st->sigma1, beta0);
        st->gamma1 = st->gamma0;
        st->sigma1 = st->sigma0;
        st->gamma0 = cdiv(delta, rho1);
        st->sigma0 = cdiv(make_double2(beta1, 0.0), rho1);
        st->rho1 = rho1;
        st->rho2 = rho2;
        st->rho3 = rho3;
        st->inv_rho1 = cdiv(make_double2(1.0, 0.0), rho1);
        st->xcoef = cmul(st->gamma0, st->eta);
        const double as0 = hypot(st->sigma0.x, st->sigma0.y);
        st->res = cscale(st->res, as0);
        st->eta = make_double2(-(st->sigma0.x * st->eta.x - st->sigma0.y * st->eta.y),
                               -(st->sigma0.x * st->eta.y + st->sigma0.y * st->eta.x));
        st->beta1 = beta1;
        st->beta0 = beta1;
        st->iters = i + 1;
        run[i + 1] = (hypot(st->res.x, st->res.y) > st->tol && i + 1 < st->maxit) ? 1 : 0;
    }
}

// w_prev = (v_curr - rho3 w_prev - rho2 w_curr) / rho1 ; x += gamma0*eta*w_prev
__global__ __launch_bounds__(VB) void k_minres_wx(int64_t n, const double2 *vc, double2 *wp,
                                                  const double2 *wc, double2 *x,
                                                  const MinresDev *st, const int *run, int i) {
    if (!run[i]) return;
    const double2 r2 = st->rho2, r3 = st->rho3, ir1 = st->inv_rho1, xc = st->xcoef;
    for (int64_t k = blockIdx.x * (int64_t)VB + threadIdx.x; k < n; k += (int64_t)gridDim.x * VB) {
        double2 w = cmul(csub(csub(vc[k], cmul(r3, wp[k])), cmul(r2, wc[k])), ir1);
        wp[k] = w;
        x[k] = cadd(x[k], cmul(xc, w));
    }
}

// ------------------------------------------- Lanczos (full reorthogonalization)

// partials[c * nbx + bx] = sum_{r in block bx} conj(V[r, c]) * w[r], c in [c0, c0+8)
__global__ __launch_bounds__(VB) void k_gemv_ct(int64_t n, const double2 *V, int64_t ld, int k,
                                                const double2 *w, double2 *partials) {
    const int c0 = blockIdx.y * 8;
    double2 acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = make_double2(0, 0);
    for (int64_t r = blockIdx.x * (int64_t)VB + threadIdx.x; r < n; r += (int64_t)gridDim.x * VB) {
        const double2 wr = w[r];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (c0 + c < k) acc[c] = cadd(acc[c], cconjmul(V[(int64_t)(c0 + c) * ld + r], wr));
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        double2 s = block_sum_256(acc[c]);
        if (threadIdx.x == 0 && c0 + c < k) partials[(int64_t)(c0 + c) * gridDim.x + blockIdx.x] = s;
    }
}

// h[c] = sum_b partials[c * nbx + b]
__global__ void k_reduce_cols(const double2 *partials, int nbx, int k, double2 *h) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    double2 s = make_double2(0, 0);
    for (int b = 0; b < nbx; ++b) s = cadd(s, partials[(int64_t)c * nbx + b]);
    h[c] = s;
}

// w[r] -= sum_c V[r, c] h[c]
__global__ __launch_bounds__(VB) void k_gemv_n_sub(int64_t n, const double2 *V, int64_t ld, int k,
                                                   const double2 *h, double2 *w) {
    __shared__ double2 hs[1024];
    for (int c = threadIdx.x; c < k && c < 1024; c += VB) hs[c] = h[c];
    __syncthreads();
    for (int64_t r = blockIdx.x * (int64_t)VB + threadIdx.x; r < n; r += (int64_t)gridDim.x * VB) {
        double2 s = make_double2(0, 0);
        for (int c = 0; c < k; ++c) s = cadd(s, cmul(V[(int64_t)c * ld + r], c < 1024 ? hs[c] : h[c]));
        w[r] = csub(w[r], s);
    }
}

// X[:, j] = sum_c V[:, c] S[c + j*k]  (S real, column-major k x nvec)
__global__ __launch_bounds__(VB) void k_combine(int64_t n, const double2 *V, int64_t ld, int k,
                                                const double *S, int nvec, double2 *X, int64_t ldx) {
    for (int64_t r = blockIdx.x * (int64_t)VB + threadIdx.x; r < n; r += (int64_t)gridDim.x * VB) {
        for (int j = 0; j < nvec; ++j) {
            double2 s = make_double2(0, 0);
            for (int c = 0; c < k; ++c) s = cadd(s, cscale(V[(int64_t)c * ld + r], S[(int64_t)j * k + c]));
            X[(int64_t)j * ldx + r] = s;
        }
    }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// real uniform start vector in [-0.5, 0.5) as in src/lanczos.f90:60-61, seeded
__global__ void k_random_real(int64_t n, uint64_t seed, double2 *v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h = splitmix64(seed * 0x100000001b3ull + (uint64_t)i);
        v[i] = make_double2((double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5, 0.0);
    }
}

__global__ void k_scale_real(int64_t n, double2 *v, double s) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        v[i] = cscale(v[i], s);
}

// w = w - a*x - b*y  (a, b real host scalars)
__global__ void k_axpby_sub(int64_t n, double2 *w, const double2 *x, double a, const double2 *y,
                            double b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = w[i];
        v = csub(v, cscale(x[i], a));
        if (y) v = csub(v, cscale(y[i], b));
        w[i] = v;
    }
}

__global__ void k_copy_div(int64_t n, const double2 *in, double s, double2 *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = make_double2(in[i].x / s, in[i].y / s);
}

__global__ void k_sub_vec(int64_t n, double2 *a, const double2 *b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = csub(a[i], b[i]);
}

}  // namespace
