// Matrix-free Hamiltonian apply (part of edx_kernels.hip).
//
// For spaces too large to keep as CSR (the core-hole space at norb=26, N=13:
// 57,946,200 rows), y = c0 H x + c1 z is computed on the fly.  A 64-lane wave
// owns a row: its lanes test the terms in parallel and apply the same operator
// steps as the CSR build (op_ann / op_cre: occupancy test and fermion sign;
// row_word / col_of: the core-hole layout of src/ham.f90:419-446 and the
// combinatorial rank), so every surviving term lands on the column and with
// the sign the CSR build gives it; the lanes accumulate val * sign * x[col]
// over the surviving terms and the wave reduces.  The omega diagonal of ham_sign = -1 is added last as
// in src/ham.f90:156-176.  Only the summation order of duplicate terms
// differs from a CSR product, so results agree to rounding.
//
// No matrix is read: per row the traffic is the gathers of x and the epilogue
// vectors, and the work is the term tests (integer ALU).

namespace {

constexpr int MF_CAP = 128;   // surviving terms of one row compacted in LDS

__device__ __forceinline__ double2 wave_sum64(double2 v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    return v;
}

// A term survives on word w exactly when its annihilated orbitals are
// occupied and its created ones are empty once those are removed:
//   (w & occ) == occ  and  (w & ~occ & emp) == 0,
// the conditions op_ann / op_cre test one operator at a time (a term that
// annihilates or creates the same orbital twice never survives: occ = ~0).
struct MfMask {
    uint64_t occ, emp;
};

__global__ void k_mf_masks(const uint32_t *ops, int T, int nhop, MfMask *masks) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const uint32_t op = ops[t];
    const int a = op & 63, b = (op >> 8) & 63, c = (op >> 16) & 63, d = (op >> 24) & 63;
    MfMask k;
    if (t < nhop) {
        k.occ = 1ull << a;
        k.emp = 1ull << b;
    } else if (a == b || c == d) {
        k.occ = ~0ull;
        k.emp = 0;
    } else {
        k.occ = (1ull << a) | (1ull << b);
        k.emp = (1ull << c) | (1ull << d);
    }
    masks[t] = k;
}

// Pass 1 of a row: lanes test the terms with the masks and ballot-compact the
// survivors numbered [base, base + MF_CAP) into LDS; returns the row's
// survivor count.
__device__ __forceinline__ int mf_compact(const BuildParams &p, const MfMask *__restrict__ masks,
                                          uint64_t w, int base, uint32_t *s_t) {
    const int lane = threadIdx.x;
    int n = 0;
    for (int t0 = 0; t0 < p.T; t0 += 64) {
        const int t = t0 + lane;
        bool ok = false;
        if (t < p.T) {
            const MfMask k = masks[t];
            ok = (w & k.occ) == k.occ && (w & ~k.occ & k.emp) == 0;
        }
        const uint64_t mask = __ballot(ok);
        const int pos = n + lanes_below(mask) - base;
        n += __popcll(mask);
        if (ok && pos >= 0 && pos < MF_CAP) s_t[pos] = (uint32_t)t;
    }
    return n;
}

// Pass 2: lanes take the compacted survivors: the operator steps of the CSR
// build (new word and fermion sign), the column rank and the x gather.
template <int CK>
__device__ __forceinline__ double2 mf_gather(const BuildParams &p, const double2 *__restrict__ x,
                                             uint64_t w, int count, const uint32_t *s_t) {
    double2 acc = make_double2(0, 0);
    for (int q = threadIdx.x; q < count; q += 64) {
        const int t = (int)s_t[q];
        const uint32_t op = p.ops[t];
        uint64_t m = w;
        int par = 0;
        if (t < p.nhop) {
            op_ann(m, op & 63, par);
            op_cre(m, (op >> 8) & 63, par);
        } else {
            op_ann(m, op & 63, par);
            op_ann(m, (op >> 8) & 63, par);
            op_cre(m, (op >> 16) & 63, par);
            op_cre(m, (op >> 24) & 63, par);
        }
        const int64_t c = col_of<CK>(p, m);
        if (c < 0) continue;
        double2 v = p.tv[t];
        if ((par ^ p.negate) & 1) v = make_double2(-v.x, -v.y);
        acc = cadd(acc, cmul(v, x[c]));
    }
    return acc;
}

// One 64-lane workgroup per row: compact the surviving terms, then rank and
// gather them in parallel, so the costly rank runs once per surviving term
// instead of once per wave-wide batch of terms.  Rows with more than MF_CAP
// survivors take further compaction passes.
template <int RK, int CK>
__global__ __launch_bounds__(64) void k_mf(BuildParams p, const MfMask *__restrict__ masks,
                                           const double2 *__restrict__ x, SpmvEpi e) {
    if (e.run && *e.run == 0) return;
    __shared__ uint32_t s_t[MF_CAP];
    const int lane = threadIdx.x;
    const double2 c0 = e.c0 ? *e.c0 : make_double2(1.0, 0.0);
    const double2 c1 = e.c1 ? *e.c1 : make_double2(1.0, 0.0);
    double2 dot = make_double2(0, 0);
    for (int64_t r = blockIdx.x; r < p.M; r += gridDim.x) {
        const uint64_t w = row_word<RK>(p, r);
        const int n = mf_compact(p, masks, w, 0, s_t);
        __syncthreads();
        double2 acc = mf_gather<CK>(p, x, w, n < MF_CAP ? n : MF_CAP, s_t);
        for (int base = MF_CAP; base < n; base += MF_CAP) {
            __syncthreads();
            mf_compact(p, masks, w, base, s_t);
            __syncthreads();
            acc = cadd(acc, mf_gather<CK>(p, x, w, n - base < MF_CAP ? n - base : MF_CAP, s_t));
        }
        __syncthreads();
        acc = wave_sum64(acc);
        if (lane == 0) {
            double2 ax = acc;
            if (p.omega_on) ax = cadd(ax, cmul(p.omega, x[r]));
            double2 out = cmul(c0, ax);
            if (e.z) out = cadd(out, cmul(c1, e.z[r]));
            e.y[r] = out;
            if (e.dvec) dot = cadd(dot, cconjmul(e.dvec[r], out));
        }
    }
    if (e.dvec && lane == 0) e.partials[blockIdx.x] = dot;
}

}  // namespace

struct edx_mfop {
    BuildParams p{};
    TermTable tt;
    MfMask *d_masks = nullptr;
    int corehole = 0;   // row/column layout: core-hole space (H_n) or plain (H_i)
    ~edx_mfop() { edx::dfree(d_masks); }
};

namespace {

int mf_blocks(const edx_mfop *op) {
    int64_t b = op->p.M;
    const int64_t cap = (int64_t)edx::num_cus() * 256;
    if (b > cap) b = cap;
    return (int)std::max<int64_t>(b, 1);
}

int64_t mf_partials_len(const edx_mfop *op) { return mf_blocks(op) + FOLD_N + 64; }

int mf_launch(const edx_mfop *op, const double2 *x, const SpmvEpi &e, int *nblocks_out) {
    const int nb = mf_blocks(op);
    if (nblocks_out) *nblocks_out = nb;
    if (op->p.M == 0) return EDX_OK;
    if (op->corehole)
        hipLaunchKernelGGL((k_mf<ROW_COREHOLE, COL_COREHOLE>), dim3(nb), dim3(64), 0,
                           edx::stream(), op->p, op->d_masks, x, e);
    else
        hipLaunchKernelGGL((k_mf<ROW_PLAIN, COL_PLAIN>), dim3(nb), dim3(64), 0, edx::stream(),
                           op->p, op->d_masks, x, e);
    EDX_LAUNCH_CHECK();
    if (e.dvec && nb > FOLD_N) {
        hipLaunchKernelGGL(k_fold_partials, dim3(FOLD_N), dim3(256), 0, edx::stream(), e.partials,
                           nb, e.run);
        hipLaunchKernelGGL(k_move_folded, dim3(1), dim3(FOLD_N), 0, edx::stream(), e.partials, nb,
                           e.run);
        EDX_LAUNCH_CHECK();
        if (nblocks_out) *nblocks_out = FOLD_N;
    }
    return EDX_OK;
}

int mf_make(const edx_tensor2 *hop, int64_t nhopp, const edx_tensor4 *coul, int64_t ncoul,
            double omega_re, double omega_im, double ham_sign, edx_mfop **out, edx_mfop **op) {
    EDX_CHECK(out, EDX_ERR_ARG, "null output");
    *out = nullptr;
    EDX_TRY(check_ham_sign(ham_sign));
    EDX_TRY(edx::ensure_binomials());
    edx_mfop *o = new edx_mfop;
    int rc = upload_terms(hop, nhopp, coul, ncoul, o->tt);
    if (rc) {
        delete o;
        return rc;
    }
    o->p = base_params(o->tt);
    const int T = o->tt.T;
    o->d_masks = (MfMask *)edx::dmalloc((size_t)(T ? T : 1) * sizeof(MfMask), "term masks");
    if (!o->d_masks) {
        delete o;
        return EDX_ERR_NOMEM;
    }
    if (T) {
        hipLaunchKernelGGL(k_mf_masks, dim3((T + 255) / 256), dim3(256), 0, edx::stream(),
                           o->tt.d_ops, T, o->tt.nhop, o->d_masks);
        if (hipGetLastError() != hipSuccess) {
            delete o;
            return edx::set_error(EDX_ERR_HIP, "term mask launch failed");
        }
    }
    o->p.omega_on = ham_sign < 0;
    o->p.omega = make_double2(omega_re, omega_im);
    o->p.negate = ham_sign < 0;
    *op = o;
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_mf_ham_i(const edx_basis *fock, const edx_tensor2 *hop, int64_t nhopp,
                 const edx_tensor4 *coul, int64_t ncoul, double omega_re, double omega_im,
                 double ham_sign, edx_mfop **out) {
    EDX_CHECK(fock, EDX_ERR_ARG, "null basis");
    edx_mfop *o = nullptr;
    EDX_TRY(mf_make(hop, nhopp, coul, ncoul, omega_re, omega_im, ham_sign, out, &o));
    o->p.M = fock->n;
    o->p.left = fock->d_words;
    o->p.nleft = fock->n;
    o->p.right = view_of(fock);
    o->corehole = 0;
    *out = o;
    return EDX_OK;
}

int edx_mf_ham_n(const edx_basis *fock_n, int num_val_orbs, int num_core_orbs,
                 const edx_tensor2 *hop, int64_t nhopp, const edx_tensor4 *coul, int64_t ncoul,
                 double omega_re, double omega_im, double ham_sign, edx_mfop **out) {
    EDX_CHECK(fock_n, EDX_ERR_ARG, "null basis");
    EDX_TRY(check_core(num_val_orbs, num_core_orbs));
    edx_mfop *o = nullptr;
    EDX_TRY(mf_make(hop, nhopp, coul, ncoul, omega_re, omega_im, ham_sign, out, &o));
    o->p.M = fock_n->n * num_core_orbs;
    o->p.left = fock_n->d_words;
    o->p.nleft = fock_n->n;
    o->p.right = view_of(fock_n);
    o->p.nv = num_val_orbs;
    o->p.nc = num_core_orbs;
    o->corehole = 1;
    *out = o;
    return EDX_OK;
}

int edx_mf_size(const edx_mfop *op, int64_t *n) {
    EDX_CHECK(op && n, EDX_ERR_ARG, "null argument");
    *n = op->p.M;
    return EDX_OK;
}

int edx_mf_apply(const edx_mfop *op, const void *d_x, void *d_y) {
    EDX_CHECK(op && d_x && d_y, EDX_ERR_ARG, "null argument");
    SpmvEpi e{};
    e.y = (double2 *)d_y;
    e.z = (const double2 *)d_y;   // y += H x
    hipEvent_t pe = edx::prof_start(EDX_C128, EDX_KERNEL_MF, edx::mf_bytes(op->p.M, op->p.M));
    EDX_TRY(mf_launch(op, (const double2 *)d_x, e, nullptr));
    edx::prof_stop(pe);
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    return EDX_OK;
}

int edx_build_krylov_mf(const edx_mfop *op, const void *d_v0, int maxn, int *neff, double *alpha,
                        double *beta, double *norm) {
    EDX_CHECK(op && d_v0 && neff && alpha && beta && norm && maxn >= 1, EDX_ERR_ARG,
              "bad argument");
    auto prod = [&](const double2 *x, const SpmvEpi &e, int *nb, const int *run) {
        (void)run;
        hipEvent_t pe = edx::prof_start(EDX_C128, EDX_KERNEL_MF, edx::mf_bytes(op->p.M, op->p.M));
        int rc = mf_launch(op, x, e, nb);
        edx::prof_stop(pe);
        return rc;
    };
    int rc = krylov_loop(op->p.M, prod, mf_partials_len(op), (const double2 *)d_v0, maxn, neff,
                         alpha, beta, norm);
    if (rc) return rc;
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    return EDX_OK;
}

void edx_mf_free(edx_mfop *op) { delete op; }

}  // extern "C"
