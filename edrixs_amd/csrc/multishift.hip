// Multi-shift Galerkin (FOM) solver for the RIXS intermediate states (part of
// edx_kernels.hip).
//
//   edx_shifted_fom_project  <- the per-incident-energy solve of the RIXS driver
//                               (src/rixs_driver.f90:127-150: build omega - H_n,
//                               pminres_csr) and the emission operator applied to
//                               its solution (src/rixs_driver.f90:163-185)
//   edx_shifted_fom_project_batch  the same for several right-hand sides (the
//                               driver's incident polarizations) in one
//                               interleaved recurrence over the multi-RHS SpMM
//
// The Krylov space K_m(H, b) is the same for every shift sigma, so one Lanczos
// recurrence on H (the build_krylov_mp recurrence, src/lanczos.f90:210-303)
// serves every incident energy at once.  With H V_m = V_m T_m + beta_m v_{m+1}
// e_m^T the Galerkin iterate of (sigma - H) x = b is
//     x_m = V_m y_m,   (sigma - T_m) y_m = |b| e_1,
// whose residual is exactly beta_m |e_m^T y_m| times v_{m+1}.  Writing
// sigma - T_m = L_m U_m (L unit lower bidiagonal, U upper bidiagonal with
// pivots eta_j and superdiagonal -beta_j) gives the short recurrences
//     eta_1 = sigma - alpha_1,  eta_j = sigma - alpha_j - beta_{j-1}^2 / eta_{j-1}
//     zeta_1 = |b|,             zeta_j = beta_{j-1} zeta_{j-1} / eta_{j-1}
//     p_j = (v_j + beta_{j-1} p_{j-1}) / eta_j,   x_j = x_{j-1} + zeta_j p_j
//     |r_j| = beta_j |zeta_j / eta_j|
// (the direct Lanczos form, Saad, Iterative Methods for Sparse Linear Systems,
// 2nd ed., sec. 6.7.1).  Im eta_j has the sign of Im sigma and
// |Im eta_j| >= |Im sigma|, so a shift with a nonzero imaginary part never
// meets a zero pivot.  Only projections P x_j are kept: q_j = P p_j obeys the
// same recurrence with P v_j in place of v_j, so each shift carries two vectors
// of P's row space (for the emission operator, the final-state space) and the
// n-space solution is never stored.

namespace {

constexpr int MS_MAXK = 256;   // shifts per call (one block holds their scalars)

struct MsShift {
    double2 sigma;
    double2 eta;       // pivot eta_i of the current step
    double2 inv_eta;   // 1 / eta_i
    double2 zeta;      // zeta_i
    double res;        // residual estimate of the last finished iterate (absolute)
    int active;        // still iterating
    int iters;         // iterate the stored projection belongs to
};

struct MsCtl {
    double tol_abs;
    int nactive;
};

__device__ __forceinline__ double2 crecip(double2 a) {
    const double d = a.x * a.x + a.y * a.y;
    return make_double2(a.x / d, -a.y / d);
}

__device__ __forceinline__ double cabs2(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }

// Step i (1-based) of the shift recurrences, after alpha_i is known.  A shift
// whose iterate i-1 already meets the tolerance is frozen there.  Block r
// serves right-hand side r of R: alpha_i, beta_{i-1} and the run flag of step
// i-1 at [(i - 1) R + r], its K shifts at sh[r K ..], its control at ctl[r].
__global__ __launch_bounds__(256) void k_ms_scalars(int i, const double *alpha, const double *beta,
                                                    int R, const double *sqrt_norm, MsShift *sh,
                                                    int K, MsCtl *ctl, const int *run) {
    const int rr = blockIdx.x;
    const int64_t o = (int64_t)(i - 1) * R + rr;
    if (!run[o]) return;
    sh += (int64_t)rr * K;
    ctl += rr;
    __shared__ int cnt[256];
    int c = 0;
    const double tol = ctl->tol_abs;
    const double nb = sqrt_norm[rr];
    for (int k = threadIdx.x; k < K; k += 256) {
        MsShift s = sh[k];
        const double a = alpha[o];
        if (i == 1) {
            s.eta = make_double2(s.sigma.x - a, s.sigma.y);
            s.zeta = make_double2(nb, 0.0);
            s.active = 1;
            s.iters = 0;
            s.res = nb;
        } else if (s.active) {
            const double bp = beta[o];
            const double2 ratio = cmul(s.zeta, s.inv_eta);   // zeta_{i-1} / eta_{i-1}
            s.res = bp * cabs2(ratio);
            if (s.res <= tol) {
                s.active = 0;
            } else {
                s.zeta = cscale(ratio, bp);
                const double2 t = cscale(s.inv_eta, bp * bp);
                s.eta = make_double2(s.sigma.x - a - t.x, s.sigma.y - t.y);
            }
        }
        if (s.active) {
            s.inv_eta = crecip(s.eta);
            s.iters = i;
            ++c;
        }
        sh[k] = s;
    }
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) cnt[threadIdx.x] += cnt[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) ctl->nactive = cnt[0];
}

// A projection operator: a CSR with complex128 values, and where it has one
// (csr_add_packed_all) its packed words (column | value number << 24) with
// the table of distinct values, read instead of the columns and values.
// prow == nullptr: the identity.
struct MsP {
    const int64_t *prow;
    const int32_t *pcol;
    const double2 *pval;
    const uint32_t *ppk;
    const double2 *ptab;
};

MsP ms_p(const edx_csr *P) {
    if (!P) return MsP{};
    return MsP{P->d_rowptr, P->d_cols, (const double2 *)P->d_vals, P->d_apk, P->d_atab};
}

// (P v)[r], four entries at a time (their columns and values, then four
// gathers in flight), summed in entry order (the packed words give the same
// values, so the same sums); element c of v is v[c * vs + vo] (column vo of an
// interleaved block of vs vectors).
__device__ __forceinline__ double2 ms_proj_row(int64_t r, const MsP &P, const double2 *v, int vs,
                                               int vo) {
    if (!P.prow) return v[r * vs + vo];
    double2 w = make_double2(0, 0);
    int64_t j = P.prow[r];
    const int64_t je = P.prow[r + 1];
    if (P.ppk) {
        constexpr uint32_t mask = (1u << PK_COL_BITS) - 1;
        for (; j + 4 <= je; j += 4) {
            uint32_t wd[4];
            double2 x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) wd[u] = P.ppk[j + u];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = v[(int64_t)(wd[u] & mask) * vs + vo];
#pragma unroll
            for (int u = 0; u < 4; ++u) w = cadd(w, cmul(P.ptab[wd[u] >> PK_COL_BITS], x[u]));
        }
        for (; j < je; ++j) {
            const uint32_t wd = P.ppk[j];
            w = cadd(w, cmul(P.ptab[wd >> PK_COL_BITS], v[(int64_t)(wd & mask) * vs + vo]));
        }
        return w;
    }
    const int32_t *pcol = P.pcol;
    const double2 *pval = P.pval;
    for (; j + 4 <= je; j += 4) {
        int c[4];
        double2 a[4], x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = pcol[j + u];
            a[u] = pval[j + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = v[(int64_t)c[u] * vs + vo];
#pragma unroll
        for (int u = 0; u < 4; ++u) w = cadd(w, cmul(a[u], x[u]));
    }
    for (; j < je; ++j) w = cadd(w, cmul(pval[j], v[(int64_t)pcol[j] * vs + vo]));
    return w;
}

// Projected updates go in windows of W steps, so Q and out are read and
// written once per window: each of the first W-1 steps keeps w = s P v_i in a
// slot with its shift scalars (k_ms_project), the last applies every step of
// the window (k_ms_update), per shift in step order, as a per-step update
// would: q_k = (w + beta_{i-1} q_k) / eta_k, out_k += zeta_k q_k for the
// shifts active at that step.  P is a CSR with FP64 values (nullptr rowptr:
// the identity); Q and out are mP x K column-major.
struct MsSlot {
    double2 inv_eta, zeta;
    int active;      // the shift took this step (0 where the step did not run)
};

__global__ __launch_bounds__(256) void k_ms_project(int64_t mP, MsP P,
                                                    const double2 *v, int vs, int vo,
                                                    const double *vscale, double2 *wout,
                                                    MsSlot *slot,
                                                    const MsShift *sh, int K, const int *run,
                                                    const MsCtl *ctl) {
    const int on = run[0] && ctl->nactive > 0;
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < K; k += 256)
            slot[k] = MsSlot{sh[k].inv_eta, sh[k].zeta, on ? sh[k].active : 0};
    if (!on) return;
    const double sc = *vscale;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < mP;
         r += (int64_t)gridDim.x * blockDim.x) {
        const double2 w = ms_proj_row(r, P, v, vs, vo);
        wout[r] = make_double2(w.x * sc, w.y * sc);
    }
}

// The projections of one step for up to MS_PGROUP (vector column, projection)
// entries in one launch: a row's gathers for the entries (the right-hand sides
// of an interleaved block share cache lines) come from one thread in turn
// (measured faster than workgroups of different entries side by side).
constexpr int MS_PGROUP = 8;
struct MsProjSet {
    MsP P[MS_PGROUP];
    int64_t mp[MS_PGROUP];
    double2 *wout[MS_PGROUP];       // this step's slot of w = s P v
    MsSlot *slot[MS_PGROUP];        // this step's slot of shift scalars [K]
    const MsShift *sh[MS_PGROUP];
    const MsCtl *ctl[MS_PGROUP];
    const double *vscale[MS_PGROUP];
    const int *run[MS_PGROUP];      // the entry's run flag of this step
    int vo[MS_PGROUP];              // the entry's column of v
    int n;
    int64_t mmax;
};

__global__ __launch_bounds__(256) void k_ms_project_set(MsProjSet ps, const double2 *v, int vs,
                                                        int K) {
    __shared__ int s_on[MS_PGROUP];
    __shared__ double s_sc[MS_PGROUP];
    if (threadIdx.x < ps.n) {
        const int e = threadIdx.x;
        const int on = ps.run[e][0] && ps.ctl[e]->nactive > 0;
        s_on[e] = on;
        s_sc[e] = on ? *ps.vscale[e] : 0.0;
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int e = 0; e < ps.n; ++e)
            for (int k = threadIdx.x; k < K; k += 256)
                ps.slot[e][k] = MsSlot{ps.sh[e][k].inv_eta, ps.sh[e][k].zeta,
                                       s_on[e] ? ps.sh[e][k].active : 0};
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < ps.mmax;
         r += (int64_t)gridDim.x * blockDim.x)
        for (int e = 0; e < ps.n; ++e) {
            if (!s_on[e] || r >= ps.mp[e]) continue;
            const double2 w = ms_proj_row(r, ps.P[e], v, vs, ps.vo[e]);
            ps.wout[e][r] = make_double2(w.x * s_sc[e], w.y * s_sc[e]);
        }
}

// The window's update: the NS saved steps (slots 0..NS-1, their w in
// wslots[j * mP ..]) and then the current step where with_curr and it runs;
// beta[j bstride] = beta_{i_j - 1} of the window's step j (j = NS: the current
// step); v, vs, vo as in ms_proj_row.
// Dynamic LDS: (NS + 1) K entries of inv_eta, zeta and active.
template <int NS>
__global__ __launch_bounds__(256) void k_ms_update(int64_t mP, MsP P,
                                                   const double2 *v, int vs, int vo,
                                                   const double *vscale,
                                                   const double2 *wslots, const MsSlot *slots,
                                                   const double *beta, int bstride,
                                                   const MsShift *sh, int K, double2 *Q,
                                                   double2 *out, const int *run, const MsCtl *ctl,
                                                   int with_curr) {
    extern __shared__ double2 s_lds[];
    double2 *s_ie = s_lds;                        // [(NS + 1) K]
    double2 *s_z = s_ie + (NS + 1) * K;           // [(NS + 1) K]
    int *s_a = (int *)(s_z + (NS + 1) * K);       // [(NS + 1) K]
    int *s_any = s_a + (NS + 1) * K;              // [K]: some step of the window moves shift k
    __shared__ int s_step[NS + 1];                // some shift takes step j
    __shared__ double s_b[NS + 1];
    const bool cur = with_curr && run[0] && ctl->nactive > 0;
    if (threadIdx.x <= NS) s_step[threadIdx.x] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < (NS + 1) * K; t += 256) {
        const int j = t / K, k = t - j * K;
        MsSlot q;
        if (j < NS) q = slots[t];
        else q = MsSlot{sh[k].inv_eta, sh[k].zeta, cur ? sh[k].active : 0};
        s_ie[t] = q.inv_eta;
        s_z[t] = q.zeta;
        s_a[t] = q.active;
        if (q.active) atomicOr(&s_step[j], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += 256) {
        int a = 0;
        for (int j = 0; j <= NS; ++j) a |= s_a[j * K + k];
        s_any[k] = a;
    }
    if (threadIdx.x <= NS)
        s_b[threadIdx.x] = s_step[threadIdx.x] ? beta[(int64_t)threadIdx.x * bstride] : 0.0;
    __syncthreads();
    int anyw = 0;
#pragma unroll
    for (int j = 0; j <= NS; ++j) anyw |= s_step[j];
    if (!anyw) return;
    const double sc = s_step[NS] ? *vscale : 0.0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < mP;
         r += (int64_t)gridDim.x * blockDim.x) {
        double2 w[NS + 1];
#pragma unroll
        for (int j = 0; j < NS; ++j)
            w[j] = s_step[j] ? wslots[(int64_t)j * mP + r] : make_double2(0, 0);
        w[NS] = make_double2(0, 0);
        if (s_step[NS]) {
            const double2 p = ms_proj_row(r, P, v, vs, vo);
            w[NS] = make_double2(p.x * sc, p.y * sc);
        }
        // four shifts at a time: their Q and out loads in flight together
        for (int k0 = 0; k0 < K; k0 += 4) {
            double2 q[4], acc[4];
            bool act[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = k0 + u;
                act[u] = k < K && s_any[k];
                if (act[u]) {
                    q[u] = Q[(int64_t)k * mP + r];
                    acc[u] = out[(int64_t)k * mP + r];
                }
            }
#pragma unroll
            for (int j = 0; j <= NS; ++j) {
                const double bj = s_b[j];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = j * K + k0 + u;
                    if (!act[u] || !s_a[t]) continue;
                    q[u] = cmul(make_double2(w[j].x + bj * q[u].x, w[j].y + bj * q[u].y), s_ie[t]);
                    acc[u] = cadd(acc[u], cmul(s_z[t], q[u]));
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!act[u]) continue;
                Q[(int64_t)(k0 + u) * mP + r] = q[u];
                out[(int64_t)(k0 + u) * mP + r] = acc[u];
            }
        }
    }
}

// The Lanczos vectors are kept unnormalized: u_i with v_i = u_i / s_i and
// s_i = beta_i (s_1 = 1).  alpha_i = <u_i | y> / s_i for the product
// y = (1/s_i) H u_i - (beta_i / s_{i-1}) u_{i-1}, and u_{i+1} = y - (alpha_i / s_i) u_i.
__global__ __launch_bounds__(VB) void k_ms_alpha(const double2 *partials, int nb, KrylovDev *st,
                                                 double *alpha, int i, const int *run) {
    if (!run[i - 1]) return;
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) {
        const double inv = st->coef[0].x;
        const double a = s.x * inv;
        alpha[i - 1] = a;
        st->acoef = a * inv;
        st->neff = i;
    }
}

// beta_i as in build_krylov_mp; the recurrence also stops once no shift is active
__global__ __launch_bounds__(VB) void k_ms_beta(const double2 *partials, int nb, KrylovDev *st,
                                                double *beta, int i, int *run, const MsCtl *ctl) {
    if (!run[i - 1]) {
        if (threadIdx.x == 0) run[i] = 0;
        return;
    }
    double2 s = reduce_partials(partials, nb);
    if (threadIdx.x == 0) {
        const double b = sqrt(s.x);
        beta[i] = b;
        st->coef[1] = make_double2(-b / st->scur, -0.0);
        st->scur = b;
        st->coef[0] = make_double2(b > 0.0 ? 1.0 / b : 0.0, 0.0);
        run[i] = (fabs(b) < 1e-10 || ctl->nactive == 0) ? 0 : 1;
    }
}

// shifts still active when the recurrence ends keep iterate neff (block r:
// right-hand side r of R, beta_m at beta[m R + r])
__global__ void k_ms_finish(const int *neff, const double *beta, int R, MsShift *sh, int K) {
    const int rr = blockIdx.x;
    const int m = neff[rr];
    sh += (int64_t)rr * K;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        MsShift s = sh[k];
        if (s.active && m > 0) {
            s.res = beta[(int64_t)m * R + rr] * cabs2(cmul(s.zeta, s.inv_eta));
            s.iters = m;
        }
        sh[k] = s;
    }
}

// Steps per windowed update: 8 (EDX_MS_WINDOW=1..8 overrides; 1 updates every
// step), fewer where the window's shift scalars (and one more step's) would
// not fit 24 KB of LDS.
int ms_window(int K) {
    int w = 8;
    if (const char *e = getenv("EDX_MS_WINDOW")) w = std::min(8, std::max(1, atoi(e)));
    while (w > 1 &&
           (size_t)(w + 1) * K * (2 * sizeof(double2) + sizeof(int)) + K * sizeof(int) > 24576)
        --w;
    return w;
}

int ms_update_launch(int ns, int64_t mp, const MsP &P, const double2 *v, int vs, int vo,
                     const double *vscale, const double2 *wslots,
                     const MsSlot *slots, const double *beta, const MsShift *sh, int K, double2 *Q,
                     double2 *out, const int *run, const MsCtl *ctl, int with_curr,
                     int bstride = 1) {
    const size_t lds = (size_t)(ns + 1) * K * (2 * sizeof(double2) + sizeof(int)) + K * sizeof(int);
    const dim3 g(vec_blocks(mp)), b(256);
    hipStream_t s = edx::stream();
#define EDX_MSU(NS)                                                                               \
    case NS:                                                                                      \
        hipLaunchKernelGGL(k_ms_update<NS>, g, b, lds, s, mp, P, v, vs, vo, vscale,               \
                           wslots, slots, beta, bstride, sh, K, Q, out, run, ctl, with_curr);     \
        break;
    switch (ns) {
        EDX_MSU(0)
        EDX_MSU(1)
        EDX_MSU(2)
        EDX_MSU(3)
        EDX_MSU(4)
        EDX_MSU(5)
        EDX_MSU(6)
        EDX_MSU(7)
        EDX_MSU(8)
        default: return edx::set_error(EDX_ERR_ARG, "window of %d saved steps", ns);
    }
#undef EDX_MSU
    return EDX_OK;
}

int shifted_fom_run(const edx_csr *h, int precision, const double2 *b, int K, const double *sigma,
                    int nproj, const edx_csr *const *proj, double2 *const *out, double tol,
                    int maxit, int *iters, double *final_res) {
    const int64_t n = h->n;
    hipStream_t s = edx::stream();
    for (int p = 0; p < nproj; ++p) {
        const int64_t mp = proj[p] ? proj[p]->m : n;
        if (mp && K) EDX_HIP(hipMemsetAsync(out[p], 0, (size_t)mp * K * sizeof(double2), s));
    }
    DevBuf vb, x32, part, st, ab, runf, shb, ctlb, qb, wb, slb;
    EDX_TRY(vb.alloc(2 * (size_t)n * sizeof(double2), "fom lanczos vectors"));
    if (precision == EDX_C64) EDX_TRY(x32.alloc((size_t)n * sizeof(float2), "fom x32"));
    const int nbv = vec_blocks(n);
    EDX_TRY(part.alloc((size_t)std::max<int64_t>(edx::spmv_partials_len(h), nbv) * sizeof(double2),
                       "partials"));
    EDX_TRY(st.alloc(sizeof(KrylovDev), "fom krylov state"));
    EDX_TRY(ab.alloc((2 * (size_t)maxit + 2) * sizeof(double), "alpha beta"));
    EDX_TRY(runf.alloc(((size_t)maxit + 2) * sizeof(int), "run flags"));
    EDX_TRY(shb.alloc((size_t)K * sizeof(MsShift), "shift state"));
    EDX_TRY(ctlb.alloc(sizeof(MsCtl), "shift control"));
    size_t qtot = 0;
    std::vector<size_t> qoff(nproj);
    for (int p = 0; p < nproj; ++p) {
        qoff[p] = qtot;
        qtot += (size_t)(proj[p] ? proj[p]->m : n) * K;
    }
    EDX_TRY(qb.alloc((qtot ? qtot : 1) * sizeof(double2), "projected directions"));
    // windowed updates (k_ms_project / k_ms_update): win steps per pass over Q
    // and out; wb keeps the first win-1 steps' w = s P v per projection
    const int win = ms_window(K);
    size_t wtot = 0;
    std::vector<size_t> woff(nproj);
    for (int p = 0; p < nproj; ++p) {
        woff[p] = wtot;
        wtot += (size_t)(proj[p] ? proj[p]->m : n) * (win - 1);
    }
    if (win > 1) {
        EDX_TRY(wb.alloc((wtot ? wtot : 1) * sizeof(double2), "windowed projections"));
        EDX_TRY(slb.alloc((size_t)(win - 1) * K * sizeof(MsSlot), "windowed shift scalars"));
    }
    double2 *v[2] = {vb.as<double2>(), vb.as<double2>() + n};
    double *d_alpha = ab.as<double>(), *d_beta = ab.as<double>() + maxit + 1;
    int *run = runf.as<int>();
    KrylovDev *S = st.as<KrylovDev>();
    MsShift *sh = shb.as<MsShift>();
    MsCtl *ctl = ctlb.as<MsCtl>();

    std::vector<MsShift> hs(K);
    for (int k = 0; k < K; ++k) {
        hs[k] = MsShift{};
        hs[k].sigma = make_double2(sigma[2 * k], sigma[2 * k + 1]);
    }
    EDX_HIP(hipMemcpyAsync(sh, hs.data(), K * sizeof(MsShift), hipMemcpyHostToDevice, s));
    EDX_HIP(hipMemsetAsync(qb.p, 0, (qtot ? qtot : 1) * sizeof(double2), s));
    EDX_HIP(hipMemsetAsync(vb.p, 0, 2 * (size_t)n * sizeof(double2), s));
    EDX_HIP(hipMemsetAsync(ab.p, 0, (2 * (size_t)maxit + 2) * sizeof(double), s));
    EDX_HIP(hipMemsetAsync(run, 0, ((size_t)maxit + 2) * sizeof(int), s));
    EDX_HIP(hipMemcpyAsync(v[0], b, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_krylov_init, dim3(1), dim3(1), 0, s, S, run);
    hipLaunchKernelGGL(k_dot, dim3(nbv), dim3(VB), 0, s, n, v[0], v[0], part.as<double2>(), nullptr);
    hipLaunchKernelGGL(k_krylov_norm, dim3(1), dim3(VB), 0, s, part.as<double2>(), nbv, S);
    EDX_LAUNCH_CHECK();
    KrylovDev hk;
    EDX_HIP(edx::copy_sync(&hk, S, sizeof(KrylovDev), hipMemcpyDeviceToHost));
    if (!(hk.sqrt_norm > 0.0)) {   // b = 0: every solution is 0
        for (int k = 0; k < K; ++k) {
            iters[k] = 0;
            final_res[k] = 0.0;
        }
        return EDX_OK;
    }
    MsCtl hc{tol * hk.sqrt_norm, K};
    EDX_HIP(hipMemcpyAsync(ctl, &hc, sizeof(MsCtl), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_div_real, dim3(nbv), dim3(VB), 0, s, n, v[0], &S->sqrt_norm, nullptr);
    EDX_LAUNCH_CHECK();

    int curr = 0, prev = 1, checked = 0, last = 0;
    for (int i = 1; i <= maxit; ++i) {
        SpmvEpi e{};
        e.y = v[prev];
        e.z = v[prev];
        e.c0 = &S->coef[0];
        e.c1 = &S->coef[1];
        e.dvec = v[curr];
        e.partials = part.as<double2>();
        e.run = run + (i - 1);
        int nb = 0;
        EDX_TRY(apply(h, precision, v[curr], x32.as<float2>(), e, &nb, run + (i - 1)));
        hipLaunchKernelGGL(k_ms_alpha, dim3(1), dim3(VB), 0, s, part.as<double2>(), nb, S,
                           d_alpha, i, run);
        hipLaunchKernelGGL(k_ms_scalars, dim3(1), dim3(256), 0, s, i, d_alpha, d_beta, 1,
                           &S->sqrt_norm, sh, K, ctl, run);
        const int j = (i - 1) % win;   // the step's place in its window
        for (int p = 0; p < nproj; ++p) {
            const edx_csr *P = proj[p];
            const int64_t mp = P ? P->m : n;
            const MsP mP = ms_p(P);
            if (j < win - 1)
                hipLaunchKernelGGL(k_ms_project, dim3(vec_blocks(mp)), dim3(256), 0, s, mp, mP,
                                   v[curr], 1, 0, &S->coef[0].x,
                                   wb.as<double2>() + woff[p] + (size_t)j * mp,
                                   slb.as<MsSlot>() + (size_t)j * K, sh, K, run + (i - 1), ctl);
            else
                EDX_TRY(ms_update_launch(win - 1, mp, mP, v[curr], 1, 0, &S->coef[0].x,
                                         wb.as<double2>() + woff[p], slb.as<MsSlot>(),
                                         d_beta + (i - win), sh, K, qb.as<double2>() + qoff[p],
                                         out[p], run + (i - 1), ctl, 1));
        }
        last = i;
        hipLaunchKernelGGL(k_axpy_norm, dim3(nbv), dim3(VB), 0, s, n, v[prev], v[curr],
                           &S->acoef, part.as<double2>(), run + (i - 1));
        hipLaunchKernelGGL(k_ms_beta, dim3(1), dim3(VB), 0, s, part.as<double2>(), nbv, S, d_beta,
                           i, run, ctl);
        EDX_LAUNCH_CHECK();
        std::swap(curr, prev);
        if (i - checked >= 16 && i < maxit) {
            int r = 1;
            EDX_HIP(edx::copy_sync(&r, run + i, sizeof(int), hipMemcpyDeviceToHost));
            checked = i;
            if (!r) break;
        }
    }
    const int nrem = last % win;   // the last window's saved steps, not yet applied
    if (nrem) {
        for (int p = 0; p < nproj; ++p) {
            const edx_csr *P = proj[p];
            const int64_t mp = P ? P->m : n;
            EDX_TRY(ms_update_launch(nrem, mp, ms_p(P), v[curr], 1, 0, &S->coef[0].x, wb.as<double2>() + woff[p], slb.as<MsSlot>(),
                                     d_beta + (last - nrem), sh, K, qb.as<double2>() + qoff[p],
                                     out[p], run + (last - 1), ctl, 0));
        }
    }
    hipLaunchKernelGGL(k_ms_finish, dim3(1), dim3(256), 0, s, &S->neff, d_beta, 1, sh, K);
    EDX_LAUNCH_CHECK();
    EDX_HIP(edx::copy_sync(hs.data(), sh, K * sizeof(MsShift), hipMemcpyDeviceToHost));
    for (int k = 0; k < K; ++k) {
        iters[k] = hs[k].iters;
        final_res[k] = hs[k].res / hk.sqrt_norm;
    }
    return EDX_OK;
}


// Batched form: R right-hand sides in one interleaved Lanczos recurrence (the
// product is the multi-RHS SpMM; scalars, run flags and shift states are per
// RHS), each with the same K shifts and its own projection proj[r] (nullptr:
// the identity) into out[r].  A right-hand side also stops once none of its
// shifts is active.
__global__ void k_msb_init(int R, int K, const double *norm, double *sqrt_norm, MsCtl *ctl,
                           double tol) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const double nb = sqrt(norm[r]);
    sqrt_norm[r] = nb;
    ctl[r] = MsCtl{tol * nb, K};
}

__global__ void k_msb_stop(int i, int R, int *run, const MsCtl *ctl) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R && ctl[r].nactive == 0) run[(int64_t)i * R + r] = 0;
}

int shifted_fom_batch_run(const edx_csr *h, int precision, int R, const double2 *Bcm, int K,
                          const double *sigma, const edx_csr *const *proj, double2 *const *out,
                          double tol, int maxit, int *iters, double *final_res) {
    const int64_t n = h->n;
    hipStream_t s = edx::stream();
    for (int r = 0; r < R; ++r) {
        const int64_t mp = proj[r] ? proj[r]->m : n;
        if (mp) EDX_HIP(hipMemsetAsync(out[r], 0, (size_t)mp * K * sizeof(double2), s));
    }
    const size_t nr = (size_t)n * R;
    Dbuf vb, x32, part, sc, flags, shb, ctlb, qb, wb, slb, nrm;
    EDX_TRY(vb.alloc(2 * nr * sizeof(double2), "batched fom lanczos vectors"));
    if (precision == EDX_C64) EDX_TRY(x32.alloc(nr * sizeof(float2), "batched fom x32"));
    const int nbs = bspmm_blocks(n), nbe = elem_grid_x(n);
    EDX_TRY(part.alloc(std::max<size_t>((size_t)std::max(nbs, nbe) * R + 64,
                                        (size_t)edx::spmv_partials_len(h)) * sizeof(double2),
                       "partials"));
    // alpha [maxit R], beta [(maxit + 1) R], the per-RHS tail; run [(maxit + 1) R], neff [R]
    const size_t nsc = (size_t)maxit * R + (size_t)(maxit + 1) * R + (size_t)BK_TAIL * R;
    EDX_TRY(sc.alloc(nsc * sizeof(double), "batched fom scalars"));
    EDX_TRY(flags.alloc(((size_t)(maxit + 1) * R + R) * sizeof(int), "batched fom flags"));
    EDX_TRY(shb.alloc((size_t)R * K * sizeof(MsShift), "shift states"));
    EDX_TRY(ctlb.alloc((size_t)R * sizeof(MsCtl), "shift controls"));
    EDX_TRY(nrm.alloc((size_t)R * sizeof(double), "rhs norms"));
    std::vector<size_t> qoff(R), woff(R);
    size_t qtot = 0, wtot = 0;
    const int win = ms_window(K);
    for (int r = 0; r < R; ++r) {
        const size_t mp = (size_t)(proj[r] ? proj[r]->m : n);
        qoff[r] = qtot;
        qtot += mp * K;
        woff[r] = wtot;
        wtot += mp * win;
    }
    // every step's projections go to slots (one fused launch per step for up
    // to MS_PGROUP right-hand sides), the window's last step applies them all
    EDX_TRY(qb.alloc((qtot ? qtot : 1) * sizeof(double2), "projected directions"));
    EDX_TRY(wb.alloc((wtot ? wtot : 1) * sizeof(double2), "windowed projections"));
    EDX_TRY(slb.alloc((size_t)R * win * K * sizeof(MsSlot), "windowed shift scalars"));
    BKrylov st;
    st.alpha = sc.as<double>();
    st.beta = st.alpha + (size_t)maxit * R;
    st.norm = st.beta + (size_t)(maxit + 1) * R;
    st.c1 = (double2 *)(st.norm + R);    // 16-byte aligned: 2 maxit R + R doubles precede it
    st.sums = st.c1 + R;
    st.c0 = st.sums + R;
    st.acoef = (double *)(st.c0 + R);
    st.scur = st.acoef + R;
    st.run = flags.as<int>();
    st.neff = st.run + (size_t)(maxit + 1) * R;
    MsShift *sh = shb.as<MsShift>();
    MsCtl *ctl = ctlb.as<MsCtl>();
    double *sqn = nrm.as<double>();
    std::vector<MsShift> hs((size_t)R * K);
    for (int r = 0; r < R; ++r)
        for (int k = 0; k < K; ++k) {
            hs[(size_t)r * K + k] = MsShift{};
            hs[(size_t)r * K + k].sigma = make_double2(sigma[2 * k], sigma[2 * k + 1]);
        }
    EDX_HIP(hipMemcpyAsync(sh, hs.data(), hs.size() * sizeof(MsShift), hipMemcpyHostToDevice, s));
    EDX_HIP(hipMemsetAsync(sc.p, 0, nsc * sizeof(double), s));
    EDX_HIP(hipMemsetAsync(flags.p, 0, ((size_t)(maxit + 1) * R + R) * sizeof(int), s));
    EDX_HIP(hipMemsetAsync(qb.p, 0, (qtot ? qtot : 1) * sizeof(double2), s));
    EDX_HIP(hipMemsetAsync(vb.p, 0, 2 * nr * sizeof(double2), s));
    double2 *v[2] = {vb.as<double2>(), vb.as<double2>() + nr};
    const int kb = (R + 63) / 64;
    ELAUNCH(k_cm_to_il, n, R, n, R, Bcm, v[0]);
    ELAUNCH(k_bdot, n, R, n, R, v[0], v[0], part.as<double2>());
    hipLaunchKernelGGL(k_breduce, dim3(R), dim3(256), 0, s, part.as<double2>(), nbe, R, st.sums);
    hipLaunchKernelGGL(k_bk_init, dim3(kb), dim3(64), 0, s, st, R, st.sums);
    hipLaunchKernelGGL(k_msb_init, dim3(kb), dim3(64), 0, s, R, K, st.norm, sqn, ctl, tol);
    ELAUNCH(k_bdiv_sqrt_norm, n, R, n, R, v[0], st.norm);
    EDX_LAUNCH_CHECK();

    int curr = 0, prev = 1, checked = 0, last = 0;
    for (int i = 1; i <= maxit; ++i) {
        if (precision == EDX_C64) ELAUNCH(k_il_to_c64, n, R, n, R, v[curr], x32.as<float2>());
        BEpi e{};   // y = (1/s_i) H u_i - (beta_i/s_{i-1}) u_{i-1}; partial <u_i|y>
        e.y = v[prev];
        e.z = v[prev];
        e.c0 = st.c0;
        e.c1 = st.c1;
        e.dvec = v[curr];
        e.partials = part.as<double2>();
        int nb = 0;
        EDX_TRY(edx::bspmm(h, precision, R, v[curr], x32.as<float2>(), e, &nb));
        hipLaunchKernelGGL(k_breduce, dim3(R), dim3(256), 0, s, part.as<double2>(), nb, R, st.sums);
        hipLaunchKernelGGL(k_bk_alpha, dim3(kb), dim3(64), 0, s, st, R, i, st.sums);
        hipLaunchKernelGGL(k_ms_scalars, dim3(R), dim3(256), 0, s, i, st.alpha, st.beta, R, sqn, sh,
                           K, ctl, st.run);
        const int j = (i - 1) % win;
        for (int r0 = 0; r0 < R; r0 += MS_PGROUP) {
            MsProjSet ps{};
            ps.n = std::min(MS_PGROUP, R - r0);
            for (int e = 0; e < ps.n; ++e) {
                const int r = r0 + e;
                const edx_csr *P = proj[r];
                ps.P[e] = ms_p(P);
                ps.mp[e] = P ? P->m : n;
                ps.mmax = std::max(ps.mmax, ps.mp[e]);
                ps.wout[e] = wb.as<double2>() + woff[r] + (size_t)j * ps.mp[e];
                ps.slot[e] = slb.as<MsSlot>() + ((size_t)r * win + j) * K;
                ps.sh[e] = sh + (size_t)r * K;
                ps.ctl[e] = ctl + r;
                ps.vscale[e] = &st.c0[r].x;
                ps.run[e] = st.run + (size_t)(i - 1) * R + r;
                ps.vo[e] = r;
            }
            hipLaunchKernelGGL(k_ms_project_set, dim3(vec_blocks(ps.mmax)), dim3(256), 0, s, ps,
                               v[curr], R, K);
        }
        if (j == win - 1)
            for (int r = 0; r < R; ++r) {
                const int64_t mp = proj[r] ? proj[r]->m : n;
                EDX_TRY(ms_update_launch(win, mp, MsP{}, nullptr, R, r,
                                         nullptr, wb.as<double2>() + woff[r],
                                         slb.as<MsSlot>() + (size_t)r * win * K,
                                         st.beta + (size_t)(i - win) * R + r, sh + (size_t)r * K,
                                         K, qb.as<double2>() + qoff[r], out[r],
                                         st.run + (size_t)(i - 1) * R + r, ctl + r, 0, R));
            }
        last = i;
        // u_{i+1} = y - (alpha_i/s_i) u_i, kept unnormalized (s_{i+1} = beta_{i+1})
        ELAUNCH(k_baxpy_norm, n, R, n, R, v[prev], v[curr], st.acoef, part.as<double2>());
        hipLaunchKernelGGL(k_breduce, dim3(R), dim3(256), 0, s, part.as<double2>(), nbe, R, st.sums);
        hipLaunchKernelGGL(k_bk_beta, dim3(kb), dim3(64), 0, s, st, R, i, st.sums);
        hipLaunchKernelGGL(k_msb_stop, dim3(kb), dim3(64), 0, s, i, R, st.run, ctl);
        EDX_LAUNCH_CHECK();
        std::swap(curr, prev);
        if (i - checked >= 16 && i < maxit) {
            std::vector<int> rf(R);
            EDX_HIP(edx::copy_sync(rf.data(), st.run + (size_t)i * R, R * sizeof(int),
                                   hipMemcpyDeviceToHost));
            checked = i;
            bool any = false;
            for (int r = 0; r < R; ++r) any |= rf[r] != 0;
            if (!any) break;
        }
    }
    const int nrem = last % win;   // the last window's saved steps, not yet applied
    for (int r = 0; r < R && nrem; ++r) {
        const int64_t mp = proj[r] ? proj[r]->m : n;
        EDX_TRY(ms_update_launch(nrem, mp, MsP{}, nullptr, R, r, nullptr,
                                 wb.as<double2>() + woff[r], slb.as<MsSlot>() + (size_t)r * win * K,
                                 st.beta + (size_t)(last - nrem) * R + r, sh + (size_t)r * K, K,
                                 qb.as<double2>() + qoff[r], out[r],
                                 st.run + (size_t)(last - 1) * R + r, ctl + r, 0, R));
    }
    hipLaunchKernelGGL(k_ms_finish, dim3(R), dim3(256), 0, s, st.neff, st.beta, R, sh, K);
    EDX_LAUNCH_CHECK();
    EDX_HIP(edx::copy_sync(hs.data(), sh, hs.size() * sizeof(MsShift), hipMemcpyDeviceToHost));
    std::vector<double> hn(R);
    EDX_HIP(edx::copy_sync(hn.data(), sqn, R * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < R; ++r)
        for (int k = 0; k < K; ++k) {
            const MsShift &q = hs[(size_t)r * K + k];
            iters[(size_t)r * K + k] = hn[r] > 0.0 ? q.iters : 0;
            final_res[(size_t)r * K + k] = hn[r] > 0.0 ? q.res / hn[r] : 0.0;
        }
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_shifted_fom_project(const edx_csr *h, int precision, const void *d_b, int nshift,
                            const double *sigma, int nproj, const edx_csr *const *proj,
                            void *const *d_out, double tol, int maxit, int *iters,
                            double *final_res) {
    EDX_TRY(check_square(h, precision));
    EDX_CHECK(h->n > 0, EDX_ERR_ARG, "empty matrix");
    EDX_CHECK(nshift >= 1 && nshift <= MS_MAXK, EDX_ERR_ARG, "nshift %d not in [1, %d]", nshift,
              MS_MAXK);
    EDX_CHECK(nproj >= 1 && proj && d_out, EDX_ERR_ARG, "need at least one projection");
    EDX_CHECK(d_b && sigma && iters && final_res, EDX_ERR_ARG, "null argument");
    EDX_CHECK(maxit >= 1, EDX_ERR_ARG, "maxit %d < 1", maxit);
    EDX_CHECK(tol >= 0.0, EDX_ERR_ARG, "negative tolerance");
    for (int k = 0; k < nshift; ++k)
        EDX_CHECK(sigma[2 * k + 1] != 0.0 && std::isfinite(sigma[2 * k]) &&
                      std::isfinite(sigma[2 * k + 1]),
                  EDX_ERR_ARG, "shift %d = (%g, %g): needs a finite nonzero imaginary part", k,
                  sigma[2 * k], sigma[2 * k + 1]);
    for (int p = 0; p < nproj; ++p) {
        EDX_CHECK(d_out[p], EDX_ERR_ARG, "null output %d", p);
        if (proj[p])
            EDX_CHECK(proj[p]->n == h->n, EDX_ERR_ARG,
                      "projection %d has %lld columns, matrix has %lld rows", p,
                      (long long)proj[p]->n, (long long)h->n);
    }
    for (int p = 0; p < nproj; ++p)
        if (proj[p]) EDX_TRY(edx::csr_add_packed_all(const_cast<edx_csr *>(proj[p])));
    std::vector<double2 *> outs(nproj);
    for (int p = 0; p < nproj; ++p) outs[p] = (double2 *)d_out[p];
    return shifted_fom_run(h, precision, (const double2 *)d_b, nshift, sigma, nproj, proj,
                           outs.data(), tol, maxit, iters, final_res);
}

int edx_shifted_fom_project_batch(const edx_csr *h, int precision, int nrhs, const void *d_B,
                                  int nshift, const double *sigma, const edx_csr *const *proj,
                                  void *const *d_out, double tol, int maxit, int *iters,
                                  double *final_res) {
    EDX_TRY(check_square(h, precision));
    EDX_CHECK(h->n > 0, EDX_ERR_ARG, "empty matrix");
    EDX_CHECK(nrhs >= 1 && nrhs <= 64, EDX_ERR_ARG, "nrhs %d not in [1, 64]", nrhs);
    EDX_CHECK(nshift >= 1 && nshift <= MS_MAXK, EDX_ERR_ARG, "nshift %d not in [1, %d]", nshift,
              MS_MAXK);
    EDX_CHECK(proj && d_out && d_B && sigma && iters && final_res, EDX_ERR_ARG, "null argument");
    EDX_CHECK(maxit >= 1, EDX_ERR_ARG, "maxit %d < 1", maxit);
    EDX_CHECK(tol >= 0.0, EDX_ERR_ARG, "negative tolerance");
    for (int k = 0; k < nshift; ++k)
        EDX_CHECK(sigma[2 * k + 1] != 0.0 && std::isfinite(sigma[2 * k]) &&
                      std::isfinite(sigma[2 * k + 1]),
                  EDX_ERR_ARG, "shift %d = (%g, %g): needs a finite nonzero imaginary part", k,
                  sigma[2 * k], sigma[2 * k + 1]);
    for (int r = 0; r < nrhs; ++r) {
        EDX_CHECK(d_out[r], EDX_ERR_ARG, "null output %d", r);
        if (proj[r])
            EDX_CHECK(proj[r]->n == h->n, EDX_ERR_ARG,
                      "projection %d has %lld columns, matrix has %lld rows", r,
                      (long long)proj[r]->n, (long long)h->n);
    }
    for (int r = 0; r < nrhs; ++r)
        if (proj[r]) EDX_TRY(edx::csr_add_packed_all(const_cast<edx_csr *>(proj[r])));
    std::vector<double2 *> outs(nrhs);
    for (int r = 0; r < nrhs; ++r) outs[r] = (double2 *)d_out[r];
    return shifted_fom_batch_run(h, precision, nrhs, (const double2 *)d_B, nshift, sigma, proj,
                                 outs.data(), tol, maxit, iters, final_res);
}

}  // extern "C"
