// Grouped Krylov recurrences across matrices (part of edx_kernels.hip).
//
// A serving loop holds many small, independent Krylov recurrences at once:
// the XAS components and RIXS final states of every admitted request, each
// over its own request's matrix, started and finished at different times.
// Launched one by one, each of their products is a few microseconds of work
// behind a launch; here every step of every live member is ONE SpMV launch,
// one vector launch and three tiny reductions, whatever the number of members.
//
// Each member is a build_krylov_mp recurrence (src/lanczos.f90:266-298) on one
// vector.  Its rows map to a contiguous range of the SpMV grid (32 rows per
// block, 8 lanes per row; a member of G_PIPE_MIN_ROWS rows or more gets a
// fixed sub-grid that walks its rows in the pipelined schedule instead) and
// its elements to a contiguous range of the vector grid; a block finds its
// member by binary search over the range starts.  All
// reductions run over the member's own blocks in a fixed order, so a member's
// numbers are the same whichever other members share the launches: a request
// gets the same spectrum alone and in any mix.
//
// Members live in device memory with their (alpha, beta) history; the host
// adds and removes members between advances and reads the coefficients.  The
// step's launches read the step counter and vector parity from device state,
// so GRAPH_STEPS steps are captured once per membership as a HIP graph and
// replayed (EDX_NO_GRAPH=1 launches them one by one instead).

namespace {

constexpr int GG = 8;                    // lanes per row
constexpr int G_ROWS = 256 / GG;         // rows per SpMV block
constexpr int G_VEC = 1024;              // vector elements per vector block
constexpr int64_t G_PIPE_MIN_ROWS = 65536;   // members this large take the pipelined SpMV
constexpr int GP_G = 4, GP_U = 8;            // its lanes per row and prefetched nonzeros per lane

struct GState {
    double2 c1;      // -beta of the last step (addend scale of the next product)
    double norm;     // <v0|v0>
    int run;         // recurrence still going
    int steps;       // steps done; parity selects the current vector
};

struct GMember {
    const int64_t *rowptr;
    const int32_t *cols;
    const void *vals;        // double2 (complex128 form) or float2 (split form off-diagonal)
    const uint16_t *idx;     // indexed form: value numbers ...
    const double2 *tab;      // ... into this table
    const uint32_t *pk;      // packed indexed form (column | value number << 24), or null
    int ntab;                // table entries staged in LDS by the pipelined product (0: none)
    const double2 *diag;     // split and indexed forms: diagonal
    int64_t m;
    double2 *v0, *v1;
    double *alpha, *beta;    // [cap] each
    GState *st;
    int64_t blk0, nblk;      // SpMV blocks
    int64_t vblk0, nvblk;    // vector blocks
    int pipe;                // SpMV blocks form a pipelined sub-grid (blk0, nblk multiples of 8)
    // slot-factored form (indexed precision, interleaved core-hole matrices):
    // its kernel arguments (H_v, typed blocks, remainder, tables); fac = core
    // slots or 0
    FacArgs fa;
    int fac;
};

__device__ __forceinline__ int g_find(const GMember *mem, int nmem, int64_t bid, bool vec) {
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t start = vec ? mem[mid].vblk0 : mem[mid].blk0;
        if (start <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class VP>
__device__ __forceinline__ VP g_vals(const GMember &M);
template <>
__device__ __forceinline__ const double2 *g_vals<const double2 *>(const GMember &M) {
    return (const double2 *)M.vals;
}
template <>
__device__ __forceinline__ const float2 *g_vals<const float2 *>(const GMember &M) {
    return (const float2 *)M.vals;
}
template <>
__device__ __forceinline__ IdxPtr g_vals<IdxPtr>(const GMember &M) {
    return IdxPtr{M.idx, M.tab};
}

// A large member's product y = H x + c1 y in the pipelined schedule of
// k_spmv_pipe: the member's nblk workgroups (a multiple of 8, starting at a
// multiple of 8) split its rows into XCD-contiguous eighths and walk them
// grid-stride; each lane loads its next row's GP_U column/value numbers and
// epilogue operands before it gathers x for the current row.  Returns this
// thread's share of <x|y>.
template <class VP, bool SPLIT>
__device__ double2 g_spmv_pipe(const GMember &M, VP vals, const double2 *s_tab, double2 c1,
                               const double2 *__restrict__ x, double2 *y, int64_t l) {
    using PV = PipeVal<VP>;
    using Hd = typename PV::type;
    auto value_of = [&](Hd hv) {
        if constexpr (PV::indexed) {
            if (M.ntab) return s_tab[hv];
        }
        return PV::value(vals, hv);
    };
    const int lig = threadIdx.x & (GP_G - 1);
    constexpr int64_t gpb = 256 / GP_G;
    const int64_t per_x = M.nblk / 8, xcd = l % 8, w = l / 8;
    const int64_t rows_x = (M.m + 7) / 8;
    const int64_t rbeg = xcd * rows_x, rend = min(M.m, rbeg + rows_x);
    const int64_t stride = per_x * gpb;
    double2 dot = make_double2(0, 0);
    int64_t r = rbeg + w * gpb + threadIdx.x / GP_G;
    int c[GP_U];
    Hd h[GP_U];
    int64_t b = 0, end = 0;
    double2 pd = make_double2(0, 0), px = pd, py = pd;
    if (r < rend) {
        b = M.rowptr[r];
        end = M.rowptr[r + 1];
        if (lig == 0) {
            if (SPLIT) pd = M.diag[r];
            px = x[r];
            py = y[r];
        }
    }
#pragma unroll
    for (int u = 0; u < GP_U; ++u) {
        const int64_t j = b + lig + u * GP_G;
        PV::fetch(M.cols, vals, j, j < end, c[u], h[u]);
    }
    while (__any(r < rend)) {
        const int64_t rn = r + stride;
        int cn[GP_U];
        Hd hn[GP_U];
        int64_t bn = 0, endn = 0;
        double2 nd = make_double2(0, 0), nx = nd, ny = nd;
        if (rn < rend) {
            bn = M.rowptr[rn];
            endn = M.rowptr[rn + 1];
            if (lig == 0) {
                if (SPLIT) nd = M.diag[rn];
                nx = x[rn];
                ny = y[rn];
            }
        }
#pragma unroll
        for (int u = 0; u < GP_U; ++u) {
            const int64_t j = bn + lig + u * GP_G;
            PV::fetch(M.cols, vals, j, j < endn, cn[u], hn[u]);
        }
        double sr = 0, si = 0;
        if (r < rend) {
#pragma unroll
            for (int u = 0; u < GP_U; ++u) {
                if (c[u] >= 0) {
                    const auto a = value_of(h[u]);
                    const double2 xv = x[c[u]];
                    sr += (double)a.x * xv.x - (double)a.y * xv.y;
                    si += (double)a.x * xv.y + (double)a.y * xv.x;
                }
            }
            for (int64_t j = b + lig + GP_U * GP_G; j < end; j += GP_G) {
                int cj;
                Hd hj;
                PV::fetch(M.cols, vals, j, true, cj, hj);
                const auto a = value_of(hj);
                const double2 xv = x[cj];
                sr += (double)a.x * xv.x - (double)a.y * xv.y;
                si += (double)a.x * xv.y + (double)a.y * xv.x;
            }
        }
#pragma unroll
        for (int off = GP_G / 2; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, GP_G);
            si += __shfl_xor(si, off, GP_G);
        }
        if (r < rend && lig == 0) {
            double2 ax = make_double2(sr, si);
            if constexpr (SPLIT) ax = cadd(ax, cmul(pd, px));
            const double2 out = cadd(ax, cmul(c1, py));
            y[r] = out;
            dot = cadd(dot, cconjmul(px, out));
        }
        r = rn;
        b = bn;
        end = endn;
        pd = nd;
        px = nx;
        py = ny;
#pragma unroll
        for (int u = 0; u < GP_U; ++u) {
            c[u] = cn[u];
            h[u] = hn[u];
        }
    }
    return dot;
}

// A factored member's product y = H x + c1 y over its 64 / NC valence rows per
// wave (the product of k_spmv_fac): block l of the member takes valence rows
// 4 (64 / NC) l ... + 4 (64 / NC) - 1.  Returns this thread's share of <x|y>.
template <int NC>
__device__ double2 g_spmv_fac(const GMember &M, const double2 *s_tab, const uint64_t *s_typ, double2 c1,
                              const double2 *__restrict__ x, double2 *y, int64_t l) {
    constexpr int VPW = 64 / NC;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int vl = lane / NC, s = lane - vl * NC;
    const int64_t mv = M.m / NC;
    const int64_t v = (l * 4 + wave) * VPW + vl;
    double2 dot = make_double2(0, 0);
    if (vl < VPW && v < mv) {
        const int64_t r = v * NC + s;
        const double2 dg = M.diag[r], xr = x[r], yr = y[r];
        const double2 ax = cadd(fac_offdiag_sum<NC, 8>(M.fa, x, s_tab, s_typ, v, s, r), cmul(dg, xr));
        const double2 out = cadd(ax, cmul(c1, yr));
        y[r] = out;
        dot = cconjmul(xr, out);
    }
    return dot;
}

template <class VP, bool SPLIT>
__global__ __launch_bounds__(256) void k_gspmv(const GMember *mem, int nmem, double2 *partials) {
    const int64_t bid = blockIdx.x;
    const GMember M = mem[g_find(mem, nmem, bid, false)];
    const GState st = *M.st;
    // stopped member, or a padding block after it (uniform over the block)
    if (!st.run || bid - M.blk0 >= M.nblk) {
        if (threadIdx.x == 0) partials[bid] = make_double2(0, 0);
        return;
    }
    const double2 *x = (st.steps & 1) ? M.v1 : M.v0;
    double2 *y = (st.steps & 1) ? M.v0 : M.v1;
    if constexpr (std::is_same<VP, IdxPtr>::value) {
        if (M.fac == 6) {   // the member's value table and block types in LDS, then the factored product
            extern __shared__ double2 s_tab[];
            const uint64_t *s_typ = fac_stage_lds(M.fa, s_tab);
            double2 dot = g_spmv_fac<6>(M, s_tab, s_typ, st.c1, x, y, bid - M.blk0);
            dot = block_sum_256(dot);
            if (threadIdx.x == 0) partials[bid] = dot;
            return;
        }
    }
    if (M.pipe) {
        // indexed forms: the member's value table in LDS, the packed words where they exist
        extern __shared__ double2 s_tab[];
        if constexpr (PipeVal<VP>::indexed) {
            if (M.ntab) {
                for (int i = threadIdx.x; i < M.ntab; i += 256) s_tab[i] = M.tab[i];
                __syncthreads();
            }
        }
        const int64_t l = bid - M.blk0;
        double2 dot;
        if constexpr (std::is_same<VP, IdxPtr>::value) {
            if (M.pk)
                dot = g_spmv_pipe<PkPtr, SPLIT>(M, PkPtr{M.pk, M.tab}, s_tab, st.c1, x, y, l);
            else
                dot = g_spmv_pipe<VP, SPLIT>(M, g_vals<VP>(M), s_tab, st.c1, x, y, l);
        } else {
            dot = g_spmv_pipe<VP, SPLIT>(M, g_vals<VP>(M), s_tab, st.c1, x, y, l);
        }
        dot = block_sum_256(dot);
        if (threadIdx.x == 0) partials[bid] = dot;
        return;
    }
    const VP vals = g_vals<VP>(M);
    const int lig = threadIdx.x & (GG - 1);
    const int64_t r = (bid - M.blk0) * G_ROWS + threadIdx.x / GG;
    double2 dot = make_double2(0, 0);
    if (r < M.m) {
        double sr = 0, si = 0;
        for (int64_t j = M.rowptr[r] + lig; j < M.rowptr[r + 1]; j += GG) {
            const auto a = nt_load2(vals + j);
            const double2 xv = x[nt_load(M.cols + j)];
            sr += (double)a.x * xv.x - (double)a.y * xv.y;
            si += (double)a.x * xv.y + (double)a.y * xv.x;
        }
#pragma unroll
        for (int off = GG / 2; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, GG);
            si += __shfl_xor(si, off, GG);
        }
        if (lig == 0) {
            double2 ax = make_double2(sr, si);
            if constexpr (SPLIT) ax = cadd(ax, cmul(M.diag[r], x[r]));
            const double2 out = cadd(ax, cmul(st.c1, y[r]));
            y[r] = out;
            dot = cconjmul(x[r], out);
        }
    }
    dot = block_sum_256(dot);
    if (threadIdx.x == 0) partials[bid] = dot;
}

// alpha_i = Re <x|y> over the member's SpMV blocks (src/lanczos.f90:277-281)
__global__ __launch_bounds__(256) void k_galpha(const GMember *mem, const double2 *partials) {
    const GMember M = mem[blockIdx.x];
    if (!M.st->run) return;
    double2 s = make_double2(0, 0);
    for (int64_t b = threadIdx.x; b < M.nblk; b += 256) s = cadd(s, partials[M.blk0 + b]);
    s = block_sum_256(s);
    if (threadIdx.x == 0) M.alpha[M.st->steps] = s.x;
}

// y -= alpha x; partial |y|^2 per vector block
__global__ __launch_bounds__(256) void k_gaxpy(const GMember *mem, int nmem, double2 *vpartials) {
    const int64_t bid = blockIdx.x;
    const GMember M = mem[g_find(mem, nmem, bid, true)];
    const GState st = *M.st;
    if (!st.run) {
        if (threadIdx.x == 0) vpartials[bid] = make_double2(0, 0);
        return;
    }
    const double2 *x = (st.steps & 1) ? M.v1 : M.v0;
    double2 *y = (st.steps & 1) ? M.v0 : M.v1;
    const double al = M.alpha[st.steps];
    const int64_t base = (bid - M.vblk0) * G_VEC;
    double2 s = make_double2(0, 0);
    for (int t = threadIdx.x; t < G_VEC; t += 256) {
        const int64_t i = base + t;
        if (i < M.m) {
            const double2 v = make_double2(y[i].x - al * x[i].x, y[i].y - al * x[i].y);
            y[i] = v;
            s.x += v.x * v.x + v.y * v.y;
        }
    }
    s = block_sum_256(s);
    if (threadIdx.x == 0) vpartials[bid] = s;
}

// beta_i = |y|; breakdown below 1e-10 (src/lanczos.f90:285-293) or a full
// history ends the member; the step counter advances, flipping the vectors
__global__ __launch_bounds__(256) void k_gbeta(const GMember *mem, const double2 *vpartials, int cap) {
    const GMember M = mem[blockIdx.x];
    if (!M.st->run) return;
    double2 s = make_double2(0, 0);
    for (int64_t b = threadIdx.x; b < M.nvblk; b += 256) s = cadd(s, vpartials[M.vblk0 + b]);
    s = block_sum_256(s);
    if (threadIdx.x == 0) {
        GState st = *M.st;
        const double b = sqrt(s.x);
        M.beta[st.steps] = b;
        st.c1 = make_double2(-b, -0.0);
        st.steps += 1;
        st.run = (fabs(b) >= 1e-10 && st.steps < cap) ? 1 : 0;
        *M.st = st;
    }
}

// the new current vector /= beta (members still running)
__global__ __launch_bounds__(256) void k_gdiv(const GMember *mem, int nmem) {
    const int64_t bid = blockIdx.x;
    const GMember M = mem[g_find(mem, nmem, bid, true)];
    const GState st = *M.st;
    if (!st.run) return;
    double2 *v = (st.steps & 1) ? M.v1 : M.v0;
    const double b = M.beta[st.steps - 1];
    const int64_t base = (bid - M.vblk0) * G_VEC;
    for (int t = threadIdx.x; t < G_VEC; t += 256) {
        const int64_t i = base + t;
        if (i < M.m) v[i] = make_double2(v[i].x / b, v[i].y / b);
    }
}

// member start: norm = <v0|v0> from dot partials, v0 /= sqrt(norm)
__global__ __launch_bounds__(256) void k_ginit(const double2 *partials, int nb, GState *st) {
    double2 s = make_double2(0, 0);
    for (int b = threadIdx.x; b < nb; b += 256) s = cadd(s, partials[b]);
    s = block_sum_256(s);
    if (threadIdx.x == 0) {
        GState g;
        g.c1 = make_double2(-0.0, -0.0);
        g.norm = s.x;
        g.run = s.x > 0.0 ? 1 : 0;
        g.steps = 0;
        *st = g;
    }
}

__global__ void k_gnormalize(int64_t n, double2 *v, const GState *st) {
    const double nr = st->norm;
    if (!(nr > 0.0)) return;
    const double d = sqrt(nr);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        v[i] = make_double2(v[i].x / d, v[i].y / d);
}

struct GroupMem {
    int id;
    const edx_csr *h;
    double2 *v;        // 2 n
    double *ab;        // alpha [cap], beta [cap]
    GState *st;
};

}  // namespace

struct edx_group {
    int precision = 0;
    int cap = 0;
    int next_id = 1;
    std::vector<GroupMem> mems;
    bool dirty = true;
    // GRAPH_STEPS steps captured once per membership and replayed on the
    // null stream (the step's launches take no step-dependent arguments)
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    GMember *d_mem = nullptr;
    size_t d_mem_cap = 0;
    double2 *partials = nullptr;
    size_t partials_cap = 0;
    double2 *vpartials = nullptr;
    size_t vpartials_cap = 0;
    int64_t nblk = 0, nvblk = 0;
    int64_t pipe_nblk = 0;   // SpMV sub-grid of a large member (0: not yet known)
    int lds_tab = 0;         // LDS entries of the SpMV launch: the largest staged member table
    ~edx_group() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        for (auto &m : mems) {
            edx::dfree(m.v);
            edx::dfree(m.ab);
            edx::dfree(m.st);
        }
        edx::dfree(d_mem);
        edx::dfree(partials);
        edx::dfree(vpartials);
    }
};

namespace {

int grow(void **p, size_t *cap, size_t need, const char *what) {
    if (need <= *cap) return EDX_OK;
    size_t nc = std::max(need, *cap * 2);
    void *q = edx::dmalloc(nc, what);
    if (!q) return EDX_ERR_NOMEM;
    edx::dfree(*p);
    *p = q;
    *cap = nc;
    return EDX_OK;
}

constexpr int GRAPH_STEPS = 16;

// the launches of one step of every member, on stream s
int group_step(edx_group *g, hipStream_t s) {
    const int nmem = (int)g->mems.size();
    if (g->precision == EDX_C64D)
        hipLaunchKernelGGL((k_gspmv<const float2 *, true>), dim3((unsigned)g->nblk), dim3(256), 0,
                           s, g->d_mem, nmem, g->partials);
    else if (g->precision == EDX_C128I)
        hipLaunchKernelGGL((k_gspmv<IdxPtr, true>), dim3((unsigned)g->nblk), dim3(256),
                           g->lds_tab * sizeof(double2), s, g->d_mem, nmem, g->partials);
    else
        hipLaunchKernelGGL((k_gspmv<const double2 *, false>), dim3((unsigned)g->nblk), dim3(256),
                           0, s, g->d_mem, nmem, g->partials);
    hipLaunchKernelGGL(k_galpha, dim3(nmem), dim3(256), 0, s, g->d_mem, g->partials);
    hipLaunchKernelGGL(k_gaxpy, dim3((unsigned)g->nvblk), dim3(256), 0, s, g->d_mem, nmem,
                       g->vpartials);
    hipLaunchKernelGGL(k_gbeta, dim3(nmem), dim3(256), 0, s, g->d_mem, g->vpartials, g->cap);
    hipLaunchKernelGGL(k_gdiv, dim3((unsigned)g->nvblk), dim3(256), 0, s, g->d_mem, nmem);
    EDX_LAUNCH_CHECK();
    return EDX_OK;
}

// capture GRAPH_STEPS steps for the current membership
int group_capture(edx_group *g) {
    if (g->exec) return EDX_OK;
    if (!g->cap_stream) EDX_HIP(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking));
    EDX_HIP(hipStreamBeginCapture(g->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = EDX_OK;
    for (int i = 0; i < GRAPH_STEPS && rc == EDX_OK; ++i) rc = group_step(g, g->cap_stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->cap_stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return edx::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g->exec = nullptr;
        return edx::hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    return EDX_OK;
}

// SpMV sub-grid of a large member: up to 4 workgroups per CU, no more than
// fit at once (the kernel's register use sets that per form), rounded to 8
int64_t pipe_subgrid(int precision) {
    int per_cu = 0;
    hipError_t e;
    if (precision == EDX_C64D)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gspmv<const float2 *, true>, 256, 0);
    else if (precision == EDX_C128I)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gspmv<IdxPtr, true>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gspmv<const double2 *, false>, 256, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    return ((int64_t)edx::num_cus() * std::min(per_cu, 4) + 7) / 8 * 8;
}

int group_layout(edx_group *g) {
    if (!g->dirty) return EDX_OK;
    if (g->exec) {
        EDX_HIP(hipGraphExecDestroy(g->exec));
        g->exec = nullptr;
    }
    std::vector<GMember> hm(g->mems.size());
    int64_t blk = 0, vblk = 0;
    int lds_tab = 0;
    for (size_t j = 0; j < g->mems.size(); ++j) {
        const GroupMem &m = g->mems[j];
        const edx_csr *h = m.h;
        GMember &d = hm[j];
        const bool split = g->precision == EDX_C64D || g->precision == EDX_C128I;
        d.rowptr = split ? h->d_off_rowptr : h->d_rowptr;
        d.cols = split ? h->d_off_cols : h->d_cols;
        d.vals = g->precision == EDX_C64D ? (const void *)h->d_off_vals32 : (const void *)h->d_vals;
        d.idx = h->d_off_idx;
        d.tab = h->d_vtab;
        d.pk = g->precision == EDX_C128I ? h->d_off_pk : nullptr;
        d.diag = split ? h->d_diag : nullptr;
        d.m = h->m;
        d.v0 = m.v;
        d.v1 = m.v + h->n;
        d.alpha = m.ab;
        d.beta = m.ab + g->cap;
        d.st = m.st;
        // the slot-factored product where the matrix has the form (its table
        // always fits LDS: at most 256 values)
        d.fac = g->precision == EDX_C128I && h->d_fr_rowptr && h->nc_slots == 6 ? 6 : 0;
        if (d.fac) d.fa = edx::fac_args(h);
        d.pipe = !d.fac && h->m >= G_PIPE_MIN_ROWS;
        d.ntab = (d.fac || (d.pipe && g->precision == EDX_C128I && h->nvtab <= LDS_TAB_MAX &&
                            lds_tab_enabled()))
                     ? (int)h->nvtab
                     : 0;
        // LDS of the SpMV launch in double2 units: a factored member's block
        // types follow its value table
        lds_tab = std::max(lds_tab, d.ntab + (d.fac ? (d.fa.ntyp + 1) / 2 : 0));
        if (d.fac) {    // 4 (64 / 6) valence rows per block
            d.nblk = (h->m / 6 + 4 * (64 / 6) - 1) / (4 * (64 / 6));
        } else if (d.pipe) {   // a resident sub-grid starting at a multiple of 8
            if (!g->pipe_nblk) g->pipe_nblk = pipe_subgrid(g->precision);
            blk = (blk + 7) / 8 * 8;
            d.nblk = g->pipe_nblk;
        } else {
            d.nblk = (h->m + G_ROWS - 1) / G_ROWS;
        }
        d.nvblk = (h->m + G_VEC - 1) / G_VEC;
        if (d.nblk < 1) d.nblk = 1;
        if (d.nvblk < 1) d.nvblk = 1;
        d.blk0 = blk;
        d.vblk0 = vblk;
        blk += d.nblk;
        vblk += d.nvblk;
    }
    g->nblk = blk;
    g->nvblk = vblk;
    g->lds_tab = lds_tab;
    EDX_TRY(grow((void **)&g->d_mem, &g->d_mem_cap, std::max<size_t>(hm.size(), 1) * sizeof(GMember),
                 "group members"));
    EDX_TRY(grow((void **)&g->partials, &g->partials_cap, (size_t)std::max<int64_t>(blk, 1) * sizeof(double2),
                 "group partials"));
    EDX_TRY(grow((void **)&g->vpartials, &g->vpartials_cap,
                 (size_t)std::max<int64_t>(vblk, 1) * sizeof(double2), "group partials"));
    if (!hm.empty())
        EDX_HIP(edx::copy_sync(g->d_mem, hm.data(), hm.size() * sizeof(GMember), hipMemcpyHostToDevice));
    g->dirty = false;
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_group_create(int precision, int max_steps, edx_group **out) {
    EDX_CHECK(out, EDX_ERR_ARG, "null output handle");
    *out = nullptr;
    EDX_CHECK(precision == EDX_C128 || precision == EDX_C64D || precision == EDX_C128I,
              EDX_ERR_ARG,
              "grouped recurrences run in complex128, the split or the indexed form (got %d)",
              precision);
    EDX_CHECK(max_steps >= 1, EDX_ERR_ARG, "max_steps %d < 1", max_steps);
    edx_group *g = new edx_group;
    g->precision = precision;
    g->cap = max_steps;
    *out = g;
    return EDX_OK;
}

int edx_group_add(edx_group *g, const edx_csr *h, const void *d_v0, int *member) {
    EDX_CHECK(g && h && d_v0 && member, EDX_ERR_ARG, "null argument");
    EDX_CHECK(h->m == h->n && h->m > 0, EDX_ERR_ARG, "member matrix must be square and non-empty");
    EDX_TRY(edx::ensure_precision(const_cast<edx_csr *>(h), g->precision));
    GroupMem m{};
    m.id = g->next_id;
    m.h = h;
    m.v = (double2 *)edx::dmalloc((size_t)2 * h->n * sizeof(double2), "group vectors");
    m.ab = (double *)edx::dmalloc((size_t)2 * g->cap * sizeof(double), "group coefficients");
    m.st = (GState *)edx::dmalloc(sizeof(GState), "group state");
    if (!m.v || !m.ab || !m.st) {
        edx::dfree(m.v);
        edx::dfree(m.ab);
        edx::dfree(m.st);
        return EDX_ERR_NOMEM;
    }
    hipStream_t s = edx::stream();
    const int64_t n = h->n;
    const int nbv = vec_blocks(n);
    double2 *part = (double2 *)edx::dmalloc((size_t)nbv * sizeof(double2), "group init partials");
    if (!part) {
        edx::dfree(m.v);
        edx::dfree(m.ab);
        edx::dfree(m.st);
        return EDX_ERR_NOMEM;
    }
    int rc = EDX_OK;
    if (hipMemcpyAsync(m.v, d_v0, n * sizeof(double2), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemsetAsync(m.v + n, 0, n * sizeof(double2), s) != hipSuccess ||
        hipMemsetAsync(m.ab, 0, (size_t)2 * g->cap * sizeof(double), s) != hipSuccess) {
        rc = edx::set_error(EDX_ERR_HIP, "group member upload failed");
    } else {
        hipLaunchKernelGGL(k_dot, dim3(nbv), dim3(VB), 0, s, n, m.v, m.v, part, nullptr);
        hipLaunchKernelGGL(k_ginit, dim3(1), dim3(256), 0, s, part, nbv, m.st);
        hipLaunchKernelGGL(k_gnormalize, dim3(nbv), dim3(256), 0, s, n, m.v, m.st);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            rc = edx::set_error(EDX_ERR_HIP, "group member start failed");
    }
    edx::dfree(part);
    if (rc) {
        edx::dfree(m.v);
        edx::dfree(m.ab);
        edx::dfree(m.st);
        return rc;
    }
    g->mems.push_back(m);
    g->next_id++;
    g->dirty = true;
    *member = m.id;
    return EDX_OK;
}

int edx_group_remove(edx_group *g, int member) {
    EDX_CHECK(g, EDX_ERR_ARG, "null group");
    for (size_t j = 0; j < g->mems.size(); ++j) {
        if (g->mems[j].id != member) continue;
        EDX_HIP(hipStreamSynchronize(edx::stream()));
        edx::dfree(g->mems[j].v);
        edx::dfree(g->mems[j].ab);
        edx::dfree(g->mems[j].st);
        g->mems.erase(g->mems.begin() + j);
        g->dirty = true;
        return EDX_OK;
    }
    return edx::set_error(EDX_ERR_ARG, "no member %d in the group", member);
}

int edx_group_advance(edx_group *g, int nsteps, int *live) {
    EDX_CHECK(g && nsteps >= 0, EDX_ERR_ARG, "bad argument");
    EDX_TRY(group_layout(g));
    const int nmem = (int)g->mems.size();
    hipStream_t s = edx::stream();
    int nlive = nmem;
    const bool graph = graphs_enabled() && nmem > 0 && nsteps >= GRAPH_STEPS;
    if (graph) EDX_TRY(group_capture(g));
    for (int i = 0; i < nsteps && nmem > 0 && nlive > 0;) {
        int done;
        if (graph && nsteps - i >= GRAPH_STEPS) {
            EDX_HIP(hipGraphLaunch(g->exec, s));
            done = GRAPH_STEPS;
        } else {
            EDX_TRY(group_step(g, s));
            done = 1;
        }
        const int before = i;
        i += done;
        if (i / 32 != before / 32 || i == nsteps) {   // stop launching once every member has ended
            nlive = 0;
            for (const auto &m : g->mems) {
                GState st;
                EDX_HIP(edx::copy_sync(&st, m.st, sizeof(GState), hipMemcpyDeviceToHost));
                nlive += st.run;
            }
        }
    }
    if (live) *live = nlive;
    return EDX_OK;
}

int edx_group_read(const edx_group *g, int member, int *steps, double *alpha, double *beta,
                   double *norm, int *running) {
    EDX_CHECK(g, EDX_ERR_ARG, "null group");
    for (const auto &m : g->mems) {
        if (m.id != member) continue;
        GState st;
        EDX_HIP(edx::copy_sync(&st, m.st, sizeof(GState), hipMemcpyDeviceToHost));
        if (steps) *steps = st.steps;
        if (norm) *norm = st.norm;
        if (running) *running = st.run;
        if (alpha && st.steps)
            EDX_HIP(edx::copy_sync(alpha, m.ab, st.steps * sizeof(double), hipMemcpyDeviceToHost));
        if (beta && st.steps)
            EDX_HIP(edx::copy_sync(beta, m.ab + g->cap, st.steps * sizeof(double), hipMemcpyDeviceToHost));
        return EDX_OK;
    }
    return edx::set_error(EDX_ERR_ARG, "no member %d in the group", member);
}

void edx_group_free(edx_group *g) {
    if (g) (void)hipStreamSynchronize(edx::stream());
    delete g;
}

}  // extern "C"
