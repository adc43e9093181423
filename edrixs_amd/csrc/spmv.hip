// Complex CSR SpMV / SpMM (part of edx_kernels.hip).
//
// Reference: matvec_csr, src/spmv.f90:73-103, y(i) += sum_j aa(j) x(jaa(j)).
// Here a group of G lanes (G = 4..64, chosen from the mean row length) owns a
// row: lane l streams nonzeros l, l+G, ... of the row with non-temporal loads
// (values and column indices are read once per product, x is gathered through
// the caches), the G partial sums are combined with cross-lane shuffles, and
// the group leader applies a fused epilogue
//     y[r] = c0 * (A x)[r] + c1 * z[r]           (c0, c1 in device memory)
// optionally accumulating conj(d[r]) * y[r] into per-block partial sums for the
// next Krylov scalar.  Reductions are per block in a fixed order, so results
// are identical run to run.

namespace {

template <class T>
__device__ __forceinline__ T nt_load(const T *p) {
    return __builtin_nontemporal_load(p);
}

__device__ __forceinline__ double2 nt_load2(const double2 *p) {
    const double *q = reinterpret_cast<const double *>(p);
    return make_double2(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
}

__device__ __forceinline__ void nt_store2(double2 *p, double2 v) {
    double *q = reinterpret_cast<double *>(p);
    __builtin_nontemporal_store(v.x, q);
    __builtin_nontemporal_store(v.y, q + 1);
}
__device__ __forceinline__ float2 nt_load2(const float2 *p) {
    const uint64_t v = __builtin_nontemporal_load(reinterpret_cast<const uint64_t *>(p));
    return make_float2(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
}

// Value source of the indexed form: nonzero j holds the 16-bit number of its
// value in a table of the matrix's distinct off-diagonal values (exact
// complex128).  Behaves like a value pointer in the kernels: vals + j, nt_load2.
struct IdxPtr {
    const uint16_t *idx;
    const double2 *tab;
    __device__ __forceinline__ IdxPtr operator+(int64_t j) const { return IdxPtr{idx + j, tab}; }
};

__device__ __forceinline__ double2 nt_load2(IdxPtr p) {
    return p.tab[__builtin_nontemporal_load(p.idx)];
}

// The packed indexed form: one 32-bit word per nonzero, its column in the low
// PK_COL_BITS bits and its value number above them (matrices with fewer than
// 2^24 columns and at most 256 distinct off-diagonal values).
constexpr int PK_COL_BITS = 24;
struct PkPtr {
    const uint32_t *pk;
    const double2 *tab;
};

template <class A>
struct AccT;
template <>
struct AccT<double2> {
    using type = double;
};
template <>
struct AccT<float2> {
    using type = float;
};

struct SpmvEpi {
    double2 *y;
    const double2 *z;       // may alias y; nullptr -> no addend
    const double2 *c0;      // scales A x; nullptr -> 1
    const double2 *c1;      // scales z; nullptr -> 1
    const double2 *dvec;    // nullptr -> no dot
    double2 *partials;      // one per block when dvec != nullptr
    const int *run;         // nullptr or flag: skip the launch when *run == 0
    int xmap;               // gridDim.x % 8 == 0: XCD-contiguous block order (set by the launcher)
    int lds_tab;            // indexed form: table entries staged in LDS (0: read from global)
};

// largest value table of the indexed form that the kernels stage in LDS
constexpr int LDS_TAB_MAX = 2048;

inline bool lds_tab_enabled() {
    static const bool on = [] {
        const char *v = std::getenv("EDX_NO_LDS_TAB");
        return !(v && *v && *v != '0');
    }();
    return on;
}

// block-level sum of a complex value (256 threads), result valid in thread 0
__device__ __forceinline__ double2 block_sum_256(double2 v) {
    __shared__ double sx[4], sy[4];
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sx[wave] = v.x;
        sy[wave] = v.y;
    }
    __syncthreads();
    double2 r = make_double2(0, 0);
    if (threadIdx.x == 0) {
        r.x = (sx[0] + sx[1]) + (sx[2] + sx[3]);
        r.y = (sy[0] + sy[1]) + (sy[2] + sy[3]);
    }
    __syncthreads();
    return r;
}

// SPLIT: the split form (complex64 off-diagonal values in the CSR arrays, the
// exact FP64 diagonal in diag, complex128 x, FP64 accumulation); the leader
// adds diag[r] * x[r] to the off-diagonal sum.
template <int G, class VP, class XT, class ACC, bool SPLIT>
__global__ __launch_bounds__(256) void k_spmv(int64_t m, const int64_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ cols, VP vals,
                                              const XT *__restrict__ x,
                                              const double2 *__restrict__ diag, SpmvEpi e) {
    if (e.run && *e.run == 0) return;
    using acc_t = ACC;
    const int lig = threadIdx.x & (G - 1);
    // xmap: the workgroups the hardware places on one XCD (blockIdx.x % 8) take
    // one contiguous eighth of the row blocks, so neighbouring rows, which
    // gather nearby x, share that XCD's L2
    const int64_t bid = e.xmap ? (int64_t)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8
                               : (int64_t)blockIdx.x;
    const int64_t group = (bid * 256 + threadIdx.x) / G;
    const int64_t ngroups = (int64_t)gridDim.x * (256 / G);
    const double2 c0 = e.c0 ? *e.c0 : make_double2(1.0, 0.0);
    const double2 c1 = e.c1 ? *e.c1 : make_double2(1.0, 0.0);
    double2 dot = make_double2(0.0, 0.0);
    for (int64_t r = group; r < m; r += ngroups) {
        const int64_t b = rowptr[r], end = rowptr[r + 1];
        acc_t sr = 0, si = 0;
        int64_t j = b + lig;
        for (; j + G < end; j += 2 * G) {   // two independent loads in flight per lane
            const auto a0 = nt_load2(vals + j);
            const auto a1 = nt_load2(vals + j + G);
            const int32_t k0 = nt_load(cols + j);
            const int32_t k1 = nt_load(cols + j + G);
            const XT x0 = x[k0];
            const XT x1 = x[k1];
            sr += (acc_t)a0.x * (acc_t)x0.x - (acc_t)a0.y * (acc_t)x0.y;
            si += (acc_t)a0.x * (acc_t)x0.y + (acc_t)a0.y * (acc_t)x0.x;
            sr += (acc_t)a1.x * (acc_t)x1.x - (acc_t)a1.y * (acc_t)x1.y;
            si += (acc_t)a1.x * (acc_t)x1.y + (acc_t)a1.y * (acc_t)x1.x;
        }
        if (j < end) {
            const auto a0 = nt_load2(vals + j);
            const XT x0 = x[nt_load(cols + j)];
            sr += (acc_t)a0.x * (acc_t)x0.x - (acc_t)a0.y * (acc_t)x0.y;
            si += (acc_t)a0.x * (acc_t)x0.y + (acc_t)a0.y * (acc_t)x0.x;
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, G);
            si += __shfl_xor(si, off, G);
        }
        if (lig == 0) {
            double2 ax = make_double2((double)sr, (double)si);
            if constexpr (SPLIT) ax = cadd(ax, cmul(diag[r], x[r]));
            double2 out = cmul(c0, ax);
            if (e.z) out = cadd(out, cmul(c1, e.z[r]));
            e.y[r] = out;
            if (e.dvec) dot = cadd(dot, cconjmul(e.dvec[r], out));
        }
    }
    if (e.dvec) {
        double2 s = block_sum_256(dot);
        if (threadIdx.x == 0) e.partials[blockIdx.x] = s;
    }
}

// What a lane of the pipelined SpMV keeps per prefetched nonzero: the value
// itself, or for the indexed forms its value number (looked up when used).
// fetch: the column and the entry of nonzero j when ok (else column -1).
// indexed: the value comes from a table of distinct values (the kernels may
// stage it in LDS).
template <class VP>
struct PipeVal {
    using type = typename std::remove_cv<typename std::remove_pointer<VP>::type>::type;
    static constexpr bool indexed = false;
    __device__ static type load(VP v, int64_t j) { return v[j]; }
    __device__ static type zero() { return type{}; }
    __device__ static void fetch(const int32_t *cols, VP v, int64_t j, bool ok, int &c, type &h) {
        c = ok ? cols[j] : -1;
        h = ok ? v[j] : zero();
    }
    __device__ static type value(VP, type h) { return h; }
};
template <>
struct PipeVal<IdxPtr> {
    using type = uint16_t;
    static constexpr bool indexed = true;
    __device__ static type load(IdxPtr v, int64_t j) { return v.idx[j]; }
    __device__ static type zero() { return 0; }
    __device__ static void fetch(const int32_t *cols, IdxPtr v, int64_t j, bool ok, int &c, type &h) {
        c = ok ? cols[j] : -1;
        h = ok ? v.idx[j] : (uint16_t)0;
    }
    __device__ static double2 value(IdxPtr v, type h) { return v.tab[h]; }
};
template <>
struct PipeVal<PkPtr> {
    using type = uint32_t;
    static constexpr bool indexed = true;
    __device__ static type zero() { return 0; }
    __device__ static void fetch(const int32_t *, PkPtr v, int64_t j, bool ok, int &c, type &h) {
        const uint32_t w = ok ? v.pk[j] : 0u;
        c = ok ? (int)(w & ((1u << PK_COL_BITS) - 1)) : -1;
        h = w >> PK_COL_BITS;
    }
    __device__ static double2 value(PkPtr v, type h) { return v.tab[h]; }
};

// Pipelined single-vector SpMV, the same product and epilogue as k_spmv: a
// capped grid (gridDim.x a multiple of 8) walks the rows grid-stride, the
// workgroups of one XCD (blockIdx.x % 8) over one contiguous eighth of them,
// G lanes per row; each lane holds the column and value numbers of up to U
// nonzeros of its current row and loads those of its next row before it
// gathers x for the current one, so the index loads of one row overlap the
// gathers of the previous.  Lane l of a row holds its nonzeros l, l + G, ...,
// l + (U-1) G (index loads through the cache); nonzeros beyond U * G in a row
// are read in place.
template <int G, int U, class VP, class XT, class ACC, bool SPLIT, int WPE = 1>
__global__ __launch_bounds__(256, WPE) void k_spmv_pipe(int64_t m, const int64_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ cols, VP vals,
                                                   const XT *__restrict__ x,
                                                   const double2 *__restrict__ diag, SpmvEpi e) {
    if (e.run && *e.run == 0) return;
    using PV = PipeVal<VP>;
    using H = typename PV::type;
    using acc_t = ACC;
    // indexed form: the value table in LDS, so value lookups stay off the vector cache
    extern __shared__ double2 s_tab[];
    constexpr bool IDX = PV::indexed;
    if constexpr (IDX) {
        if (e.lds_tab) {
            for (int i = threadIdx.x; i < e.lds_tab; i += 256) s_tab[i] = vals.tab[i];
            __syncthreads();
        }
    }
    auto value_of = [&](H hv) {
        if constexpr (IDX) {
            if (e.lds_tab) return s_tab[hv];
        }
        return PV::value(vals, hv);
    };
    const int lig = threadIdx.x & (G - 1);
    constexpr int64_t gpb = 256 / G;   // row groups per workgroup
    const int64_t per_x = gridDim.x / 8, xcd = blockIdx.x % 8, w = blockIdx.x / 8;
    const int64_t rows_x = (m + 7) / 8;
    const int64_t rbeg = xcd * rows_x, rend = min(m, rbeg + rows_x);
    const int64_t stride = per_x * gpb;
    const double2 c0 = e.c0 ? *e.c0 : make_double2(1.0, 0.0);
    const double2 c1 = e.c1 ? *e.c1 : make_double2(1.0, 0.0);
    double2 dot = make_double2(0.0, 0.0);
    int64_t r = rbeg + w * gpb + threadIdx.x / G;
    int c[U];
    H h[U];
    int64_t b = 0, end = 0;
    // the epilogue's operands of the row, loaded with its indices by lane 0:
    // diag[r] * x[r] (split forms), the addend z[r], the dot vector d[r]
    double2 pd = make_double2(0, 0), pz = pd, pv = pd;
    XT px{};
    auto epi_load = [&](int64_t row, double2 &d_, XT &x_, double2 &z_, double2 &v_) {
        if (lig != 0 || row >= rend) return;
        if constexpr (SPLIT) {
            d_ = diag[row];
            x_ = x[row];
        }
        if (e.z) z_ = e.z[row];
        if (e.dvec) v_ = e.dvec[row];
    };
    if (r < rend) {
        b = rowptr[r];
        end = rowptr[r + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t j = b + lig + u * G;
        PV::fetch(cols, vals, j, j < end, c[u], h[u]);
    }
    epi_load(r, pd, px, pz, pv);
    while (__any(r < rend)) {
        const int64_t rn = r + stride;
        int cn[U];
        H hn[U];
        int64_t bn = 0, endn = 0;
        if (rn < rend) {
            bn = rowptr[rn];
            endn = rowptr[rn + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = bn + lig + u * G;
            PV::fetch(cols, vals, j, j < endn, cn[u], hn[u]);
        }
        double2 nd = make_double2(0, 0), nz = nd, nv = nd;
        XT nx{};
        epi_load(rn, nd, nx, nz, nv);
        acc_t sr = 0, si = 0;
        if (r < rend) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (c[u] >= 0) {
                    const auto a = value_of(h[u]);
                    const XT xv = x[c[u]];
                    sr += (acc_t)a.x * (acc_t)xv.x - (acc_t)a.y * (acc_t)xv.y;
                    si += (acc_t)a.x * (acc_t)xv.y + (acc_t)a.y * (acc_t)xv.x;
                }
            }
            for (int64_t j = b + lig + U * G; j < end; j += G) {
                int cj;
                H hj;
                PV::fetch(cols, vals, j, true, cj, hj);
                const auto a = value_of(hj);
                const XT xv = x[cj];
                sr += (acc_t)a.x * (acc_t)xv.x - (acc_t)a.y * (acc_t)xv.y;
                si += (acc_t)a.x * (acc_t)xv.y + (acc_t)a.y * (acc_t)xv.x;
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, G);
            si += __shfl_xor(si, off, G);
        }
        if (r < rend && lig == 0) {
            double2 ax = make_double2((double)sr, (double)si);
            if constexpr (SPLIT) ax = cadd(ax, cmul(pd, px));
            double2 out = cmul(c0, ax);
            if (e.z) out = cadd(out, cmul(c1, pz));
            e.y[r] = out;
            if (e.dvec) dot = cadd(dot, cconjmul(pv, out));
        }
        r = rn;
        b = bn;
        end = endn;
        pd = nd;
        px = nx;
        pz = nz;
        pv = nv;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = cn[u];
            h[u] = hn[u];
        }
    }
    if (e.dvec) {
        double2 s = block_sum_256(dot);
        if (threadIdx.x == 0) e.partials[blockIdx.x] = s;
    }
}

// Multi-RHS: Y[:, k] = c0 * A X[:, k] + c1 * Z[:, k], column-major blocks of
// nrhs columns.  Each lane group handles one row and NR right-hand sides at a
// time, so every value / column index read serves NR products.
template <int G, int NR, class VT, class XT>
__global__ __launch_bounds__(256) void k_spmm(int64_t m, int64_t n, const int64_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ cols,
                                              const VT *__restrict__ vals,
                                              const XT *__restrict__ X, int nrhs,
                                              double2 *__restrict__ Y,
                                              const double2 *__restrict__ Z) {
    using acc_t = typename AccT<VT>::type;
    const int lig = threadIdx.x & (G - 1);
    const int64_t group = ((int64_t)blockIdx.x * 256 + threadIdx.x) / G;
    const int64_t ngroups = (int64_t)gridDim.x * (256 / G);
    const int k0 = blockIdx.y * NR;
    for (int64_t r = group; r < m; r += ngroups) {
        const int64_t b = rowptr[r], end = rowptr[r + 1];
        acc_t sr[NR], si[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) sr[q] = si[q] = 0;
        for (int64_t j = b + lig; j < end; j += G) {
            const VT a = nt_load2(vals + j);
            const int64_t c = nt_load(cols + j);
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                if (k0 + q < nrhs) {
                    const XT xv = X[(int64_t)(k0 + q) * n + c];
                    sr[q] += (acc_t)a.x * (acc_t)xv.x - (acc_t)a.y * (acc_t)xv.y;
                    si[q] += (acc_t)a.x * (acc_t)xv.y + (acc_t)a.y * (acc_t)xv.x;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) {
#pragma unroll
            for (int off = G / 2; off > 0; off >>= 1) {
                sr[q] += __shfl_xor(sr[q], off, G);
                si[q] += __shfl_xor(si[q], off, G);
            }
        }
        if (lig == 0) {
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                if (k0 + q < nrhs) {
                    const int64_t o = (int64_t)(k0 + q) * m + r;
                    double2 out = make_double2((double)sr[q], (double)si[q]);
                    if (Z) out = cadd(out, Z[o]);
                    Y[o] = out;
                }
            }
        }
    }
}

__global__ void k_to_c64_vec(int64_t n, const double2 *in, float2 *out, const int *run) {
    if (run && *run == 0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

// Per-block dot partials of a launch with more than FOLD_N blocks are folded
// to FOLD_N sums (block j adds the j-th contiguous chunk in a fixed order) so
// the one-block scalar kernels that consume them stay short.
constexpr int FOLD_N = 256;

// Matrices with fewer rows keep the launch-shape rules: their products are
// too short to time and too cheap to matter.
constexpr int64_t TUNE_MIN_ROWS = 1 << 17;

__global__ __launch_bounds__(256) void k_fold_partials(double2 *partials, int nb, const int *run) {
    if (run && *run == 0) return;
    const int chunk = (nb + FOLD_N - 1) / FOLD_N;
    const int lo = blockIdx.x * chunk;
    const int hi = min(nb, lo + chunk);
    double2 s = make_double2(0, 0);
    for (int i = lo + threadIdx.x; i < hi; i += 256) s = cadd(s, partials[i]);
    s = block_sum_256(s);
    if (threadIdx.x == 0) partials[nb + blockIdx.x] = s;
}

__global__ void k_move_folded(double2 *partials, int nb, const int *run) {
    if (run && *run == 0) return;
    if (threadIdx.x < FOLD_N) partials[threadIdx.x] = partials[nb + threadIdx.x];
}

int lanes_for_avg(double avg) {
    if (avg < 12) return 4;
    if (avg < 40) return 8;
    if (avg < 160) return 16;
    if (avg < 480) return 32;
    return 64;
}

int lanes_for(const edx_csr *a) {
    return lanes_for_avg(a->m ? (double)a->nnz / (double)a->m : 0.0);
}

int lanes_for_split(const edx_csr *a) {
    return lanes_for_avg(a->m ? (double)a->off_nnz / (double)a->m : 0.0);
}

}  // namespace

namespace edx {

// ---------------------------------------------------------------- profiling
// Launches are bracketed by a ring of PROF_CAP event pairs.  Before a slot is
// reused, its launch (PROF_CAP launches back, long finished: the host never
// runs that far ahead of the device) is folded into per-(precision, kernel,
// right-hand sides) totals, so every launch of a profiled region counts
// however long it runs.
struct ProfRec {
    int kind;      // precision
    int kernel;    // EDX_KERNEL_*
    int nrhs;      // right-hand sides of the launch
    double bytes;  // the bytes the launch's storage form reads and writes
    double flops;
    double abytes; // algorithmic bytes of the product as a complex128 CSR (csr_alg_bytes)
    bool stopped;  // the stop event was recorded (the launch was made)
};
struct ProfSum {
    int64_t n = 0;
    double ms = 0.0, bytes = 0.0, flops = 0.0, abytes = 0.0;
};
constexpr int PROF_KINDS = 8, PROF_KERNELS = 16;
constexpr int PROF_RHS = 17;   // buckets 1..16 right-hand sides, 0: more than 16
inline int prof_rhs_bucket(int nrhs) { return nrhs >= 1 && nrhs < PROF_RHS ? nrhs : 0; }
static bool g_prof = false;
static std::vector<hipEvent_t> g_ev;      // pairs: 2k start, 2k+1 stop
static std::vector<ProfRec> g_rec;        // PROF_CAP slots
static uint64_t g_prof_n = 0;             // launches recorded since edx_profile(1)
static ProfSum g_sum[PROF_KINDS][PROF_KERNELS][PROF_RHS];   // launches folded out of the ring
constexpr size_t PROF_CAP = 16384;

static int prof_fold(size_t slot, ProfSum &acc) {
    const ProfRec &r = g_rec[slot];
    if (!r.stopped) return EDX_OK;
    float t = 0.f;
    EDX_HIP(hipEventSynchronize(g_ev[2 * slot + 1]));
    EDX_HIP(hipEventElapsedTime(&t, g_ev[2 * slot], g_ev[2 * slot + 1]));
    acc.n += 1;
    acc.ms += t;
    acc.bytes += r.bytes;
    acc.flops += r.flops;
    acc.abytes += r.abytes;
    return EDX_OK;
}

static hipEvent_t prof_start(int kind, int kernel, double bytes, double flops = 0.0,
                             int nrhs = 1, double abytes = 0.0) {
    if (!g_prof || g_ev.size() < 2 * PROF_CAP) return nullptr;
    if (kind < 0 || kind >= PROF_KINDS || kernel < 0 || kernel >= PROF_KERNELS) return nullptr;
    const size_t slot = (size_t)(g_prof_n % PROF_CAP);
    if (g_prof_n >= PROF_CAP) {
        const ProfRec &old = g_rec[slot];
        if (prof_fold(slot, g_sum[old.kind][old.kernel][prof_rhs_bucket(old.nrhs)]) != EDX_OK)
            return nullptr;
    }
    g_rec[slot] = ProfRec{kind, kernel, nrhs, bytes, flops, abytes, false};
    ++g_prof_n;
    hipEvent_t e = g_ev[2 * slot];
    (void)hipEventRecord(e, stream());
    return e;
}

static void prof_stop(hipEvent_t started) {
    if (!started) return;
    const size_t slot = (size_t)((g_prof_n - 1) % PROF_CAP);
    g_rec[slot].stopped = hipEventRecord(g_ev[2 * slot + 1], stream()) == hipSuccess;
}

// Totals of one precision, one kernel (-1: all kernels) and one count of
// right-hand sides (-1: all; above 16: every launch of more than 16): the
// folded launches plus those still in the ring.
static int prof_total(int precision, int kernel, int nrhs, ProfSum *out) {
    ProfSum acc;
    if (precision >= 0 && precision < PROF_KINDS)
        for (int k = 0; k < PROF_KERNELS; ++k)
            for (int b = 0; b < PROF_RHS; ++b)
                if ((kernel < 0 || kernel == k) && (nrhs < 0 || prof_rhs_bucket(nrhs) == b)) {
                    const ProfSum &s = g_sum[precision][k][b];
                    acc.n += s.n;
                    acc.ms += s.ms;
                    acc.bytes += s.bytes;
                    acc.flops += s.flops;
                    acc.abytes += s.abytes;
                }
    const uint64_t first = g_prof_n > PROF_CAP ? g_prof_n - PROF_CAP : 0;
    for (uint64_t i = first; i < g_prof_n; ++i) {
        const size_t slot = (size_t)(i % PROF_CAP);
        if (g_rec[slot].kind != precision) continue;
        if (kernel >= 0 && g_rec[slot].kernel != kernel) continue;
        if (nrhs >= 0 && prof_rhs_bucket(g_rec[slot].nrhs) != prof_rhs_bucket(nrhs)) continue;
        EDX_TRY(prof_fold(slot, acc));
    }
    *out = acc;
    return EDX_OK;
}

// Algorithmic bytes of a product with K right-hand sides counted as the
// complex128 CSR of the same operator, whatever form the launch reads: 16-byte
// value and 4-byte column per nonzero (diagonal included), 8-byte row
// pointers, and per right-hand side x once and y written, read as well where
// the epilogue accumulates into it (nnz (16 + 4) + 8 (m + 1) + K (16 n +
// 16 m (1 + acc))).
double csr_alg_bytes(const edx_csr *a, int K, bool acc) {
    return (double)a->nnz * 20.0 + (double)(a->m + 1) * 8.0 +
           K * ((double)a->n * 16.0 + (double)a->m * 16.0 * (acc ? 2.0 : 1.0));
}

// Algorithmic HBM bytes of one product with K right-hand sides: every stored
// value, column index and row pointer once, x once, y written once, plus the
// addend z and the dot vector d when the epilogue reads them.
double spmv_bytes_k(const edx_csr *a, int precision, int K, bool z, bool dvec) {
    const double yv = (double)a->m * 16.0 * (1.0 + (z ? 1.0 : 0.0) + (dvec ? 1.0 : 0.0));
    if (precision == EDX_C64D)
        return (double)a->off_nnz * 12.0 + (double)(a->m + 1) * 8.0 + (double)a->m * 16.0 +
               K * ((double)a->n * 16.0 + yv);
    if (precision == EDX_C128I)   // int32 column + uint16 value number per nonzero
        return (double)a->off_nnz * 6.0 + (double)(a->m + 1) * 8.0 + (double)a->m * 16.0 +
               (double)a->nvtab * 16.0 + K * ((double)a->n * 16.0 + yv);
    const double sv = precision == EDX_C64 ? 8.0 : 16.0;
    return (double)a->nnz * (sv + 4.0) + (double)(a->m + 1) * 8.0 + K * ((double)a->n * sv + yv);
}

// Algorithmic bytes of a matrix-free apply: x once, y written, the addend read.
double mf_bytes(int64_t m, int64_t n) { return (double)n * 16.0 + (double)m * 32.0; }

double spmv_bytes(const edx_csr *a, int precision, int nrhs, bool z, bool dvec) {
    return spmv_bytes_k(a, precision, nrhs, z, dvec);
}

// One lane group per row and no grid-stride cap: blocks are dispatched in row
// order, so the rows in flight form one sliding window and their gathers of x
// share cache lines (a capped grid-stride grid lets the blocks drift apart and
// was 30-40 % slower on H_n at norb=24).
int spmv_blocks(const edx_csr *a, int G) {
    int64_t blocks = (a->m * G + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks >= 64) blocks = (blocks + 7) / 8 * 8;   // XCD-contiguous order needs a multiple of 8
    return (int)blocks;
}

// Pipelined schedules (k_spmv_pipe), selected through the lanes setting:
// code 100 + 10 G + B is G lanes per row on a grid of B workgroups per CU.
// 145 (packed indexed form only): the kernel built for five resident
// workgroups per CU, which its 96 registers allow without spills.
constexpr int PIPE_MAX_BPC = 5;
bool is_pipe_code(int code) { return code == 143 || code == 144 || code == 145 || code == 184; }
int pipe_blocks(int bpc) { return (num_cus() * bpc + 7) / 8 * 8; }

// Schedule FAC_CODE: the slot-factored product of a core-hole matrix in the
// interleaved layout (factor.hip), indexed form only.
constexpr int FAC_CODE = 206;
int fac_launch(const edx_csr *a, const double2 *x, const SpmvEpi &e, int *nblocks_out);
int fac_blocks(const edx_csr *a);
double fac_bytes(const edx_csr *a, bool z, bool dvec);

// Whether schedule code G can run on this matrix in this precision.
bool schedule_available(const edx_csr *a, int precision, int G) {
    if (G == 145) return precision == EDX_C128I && a->d_off_pk;
    if (G == FAC_CODE) return precision == EDX_C128I && a->d_fr_rowptr;
    return true;
}

// Length of a partials buffer that serves every single-vector SpMV launch of
// the matrix: one entry per block of the widest launch plus the folded sums.
int64_t spmv_partials_len(const edx_csr *a) {
    int64_t nb = std::max(spmv_blocks(a, 64), pipe_blocks(PIPE_MAX_BPC));
    if (a->nc_slots > 0 && a->m == a->n) nb = std::max<int64_t>(nb, fac_blocks(a));
    return nb + FOLD_N + 64;
}

// Launch y = c0*A x + c1*z (+ dot partials).  x is complex128 for C128 and
// C64D and complex64 (already converted) for C64.  Returns the block count used.

// XCD-contiguous block order of the SpMV/SpMM launches (EDX_NO_XCDMAP=1 turns
// it off)
bool xcd_map_enabled() {
    static const bool on = [] {
        const char *off = getenv("EDX_NO_XCDMAP");
        return !(off && off[0] == '1');
    }();
    return on;
}

int spmv_launch(const edx_csr *a, int precision, const void *x, const SpmvEpi &e_in, int lanes,
                int *nblocks_out) {
    if (precision == EDX_C64D && !a->d_diag)
        return set_error(EDX_ERR_ARG, "split (c64d) form not prepared");
    if (precision == EDX_C128I && !a->d_off_idx)
        return set_error(EDX_ERR_ARG, "indexed (c128i) form not prepared");
    int G = lanes;
    if (!G) {
        if (precision == EDX_C64) G = a->spmv_lanes32 ? a->spmv_lanes32 : lanes_for(a);
        else if (precision == EDX_C64D) G = a->spmv_lanes64d ? a->spmv_lanes64d : lanes_for_split(a);
        else if (precision == EDX_C128I) G = a->spmv_lanes128i ? a->spmv_lanes128i : lanes_for_split(a);
        else G = a->spmv_lanes ? a->spmv_lanes : lanes_for(a);
    }
    if (G == FAC_CODE) {
        if (!schedule_available(a, precision, G))
            return set_error(EDX_ERR_ARG, "SpMV schedule %d needs the slot-factored form", G);
        int nb = 0;
        hipEvent_t pe = prof_start(precision, EDX_KERNEL_SPMV_FAC,
                                   fac_bytes(a, e_in.z != nullptr, e_in.dvec != nullptr), 0.0, 1,
                                   csr_alg_bytes(a, 1, e_in.z != nullptr));
        EDX_TRY(fac_launch(a, (const double2 *)x, e_in, &nb));
        prof_stop(pe);
        if (nblocks_out) *nblocks_out = nb;
        if (e_in.dvec && nb > FOLD_N) {
            hipLaunchKernelGGL(k_fold_partials, dim3(FOLD_N), dim3(256), 0, stream(), e_in.partials,
                               nb, e_in.run);
            hipLaunchKernelGGL(k_move_folded, dim3(1), dim3(FOLD_N), 0, stream(), e_in.partials, nb,
                               e_in.run);
            EDX_LAUNCH_CHECK();
            if (nblocks_out) *nblocks_out = FOLD_N;
        }
        return EDX_OK;
    }
    const bool pipe = is_pipe_code(G);
    int nb = pipe ? pipe_blocks(G % 10) : spmv_blocks(a, G);
    if (nblocks_out) *nblocks_out = nb;
    if (a->m == 0) return EDX_OK;
    SpmvEpi e = e_in;
    e.xmap = nb % 8 == 0 && xcd_map_enabled();
    e.lds_tab = precision == EDX_C128I && pipe && a->nvtab <= LDS_TAB_MAX && lds_tab_enabled()
                    ? (int)a->nvtab
                    : 0;
    // the packed indexed form reads 4 bytes of index per nonzero instead of 6
    const bool packed = precision == EDX_C128I && pipe && a->d_off_pk;
    hipEvent_t pe = prof_start(precision, EDX_KERNEL_SPMV,
                               spmv_bytes(a, precision, 1, e.z != nullptr, e.dvec != nullptr) -
                                   (packed ? 2.0 * (double)a->off_nnz : 0.0),
                               0.0, 1, csr_alg_bytes(a, 1, e.z != nullptr));
    if (pipe) {
#define EDX_PIPE_CASE(CODE, GG, UU)                                                               \
    case CODE:                                                                                    \
        if (precision == EDX_C128)                                                                \
            hipLaunchKernelGGL((k_spmv_pipe<GG, UU, const double2 *, double2, double, false>),     \
                               dim3(nb), dim3(256), 0, stream(), a->m, a->d_rowptr, a->d_cols,     \
                               (const double2 *)a->d_vals, (const double2 *)x, nullptr, e);        \
        else if (precision == EDX_C64D)                                                           \
            hipLaunchKernelGGL((k_spmv_pipe<GG, UU, const float2 *, double2, double, true>),       \
                               dim3(nb), dim3(256), 0, stream(), a->m, a->d_off_rowptr,            \
                               a->d_off_cols, (const float2 *)a->d_off_vals32, (const double2 *)x, \
                               a->d_diag, e);                                                      \
        else if (precision == EDX_C128I && a->d_off_pk)                                           \
            hipLaunchKernelGGL((k_spmv_pipe<GG, UU, PkPtr, double2, double, true>), dim3(nb),      \
                               dim3(256), e.lds_tab * sizeof(double2), stream(), a->m,             \
                               a->d_off_rowptr, nullptr, PkPtr{a->d_off_pk, a->d_vtab},            \
                               (const double2 *)x, a->d_diag, e);                                  \
        else if (precision == EDX_C128I)                                                          \
            hipLaunchKernelGGL((k_spmv_pipe<GG, UU, IdxPtr, double2, double, true>), dim3(nb),     \
                               dim3(256), e.lds_tab * sizeof(double2), stream(), a->m,             \
                               a->d_off_rowptr, a->d_off_cols, IdxPtr{a->d_off_idx, a->d_vtab},    \
                               (const double2 *)x, a->d_diag, e);                                  \
        else                                                                                      \
            hipLaunchKernelGGL((k_spmv_pipe<GG, UU, const float2 *, float2, float, false>),        \
                               dim3(nb), dim3(256), 0, stream(), a->m, a->d_rowptr, a->d_cols,     \
                               (const float2 *)a->d_vals32, (const float2 *)x, nullptr, e);        \
        break;
        if (G == 145) {
            if (!(precision == EDX_C128I && a->d_off_pk))
                return set_error(EDX_ERR_ARG, "SpMV schedule 145 needs the packed indexed form");
            hipLaunchKernelGGL((k_spmv_pipe<4, 8, PkPtr, double2, double, true, 5>), dim3(nb),
                               dim3(256), e.lds_tab * sizeof(double2), stream(), a->m,
                               a->d_off_rowptr, nullptr, PkPtr{a->d_off_pk, a->d_vtab},
                               (const double2 *)x, a->d_diag, e);
        } else switch (G) {
            EDX_PIPE_CASE(143, 4, 8)
            EDX_PIPE_CASE(144, 4, 8)
            EDX_PIPE_CASE(184, 8, 4)
            default: return set_error(EDX_ERR_ARG, "unsupported SpMV schedule %d", G);
        }
#undef EDX_PIPE_CASE
        EDX_LAUNCH_CHECK();
        prof_stop(pe);
        if (e.dvec && nb > FOLD_N) {
            hipLaunchKernelGGL(k_fold_partials, dim3(FOLD_N), dim3(256), 0, stream(), e.partials,
                               nb, e.run);
            hipLaunchKernelGGL(k_move_folded, dim3(1), dim3(FOLD_N), 0, stream(), e.partials, nb,
                               e.run);
            EDX_LAUNCH_CHECK();
            if (nblocks_out) *nblocks_out = FOLD_N;
        }
        return EDX_OK;
    }
#define EDX_SPMV_CASE(GG)                                                                     \
    case GG:                                                                                  \
        if (precision == EDX_C128)                                                            \
            hipLaunchKernelGGL((k_spmv<GG, const double2 *, double2, double, false>), dim3(nb),   \
                               dim3(256), 0, stream(), a->m, a->d_rowptr, a->d_cols,           \
                               (const double2 *)a->d_vals, (const double2 *)x, nullptr, e);    \
        else if (precision == EDX_C64D)                                                       \
            hipLaunchKernelGGL((k_spmv<GG, const float2 *, double2, double, true>), dim3(nb),     \
                               dim3(256), 0, stream(), a->m, a->d_off_rowptr, a->d_off_cols,   \
                               (const float2 *)a->d_off_vals32, (const double2 *)x, a->d_diag, e); \
        else if (precision == EDX_C128I)                                                      \
            hipLaunchKernelGGL((k_spmv<GG, IdxPtr, double2, double, true>), dim3(nb), dim3(256),  \
                               0, stream(), a->m, a->d_off_rowptr, a->d_off_cols,              \
                               IdxPtr{a->d_off_idx, a->d_vtab}, (const double2 *)x, a->d_diag, e); \
        else                                                                                  \
            hipLaunchKernelGGL((k_spmv<GG, const float2 *, float2, float, false>), dim3(nb),      \
                               dim3(256), 0, stream(), a->m, a->d_rowptr, a->d_cols,           \
                               (const float2 *)a->d_vals32, (const float2 *)x, nullptr, e);    \
        break;
    switch (G) {
        EDX_SPMV_CASE(4)
        EDX_SPMV_CASE(8)
        EDX_SPMV_CASE(16)
        EDX_SPMV_CASE(32)
        EDX_SPMV_CASE(64)
        default: return set_error(EDX_ERR_ARG, "unsupported lanes per row %d", G);
    }
#undef EDX_SPMV_CASE
    EDX_LAUNCH_CHECK();
    prof_stop(pe);
    if (e.dvec && nb > FOLD_N) {
        hipLaunchKernelGGL(k_fold_partials, dim3(FOLD_N), dim3(256), 0, stream(), e.partials, nb,
                           e.run);
        hipLaunchKernelGGL(k_move_folded, dim3(1), dim3(FOLD_N), 0, stream(), e.partials, nb, e.run);
        EDX_LAUNCH_CHECK();
        if (nblocks_out) *nblocks_out = FOLD_N;
    }
    return EDX_OK;
}

int to_c64(int64_t n, const double2 *in, float2 *out, const int *run) {
    int64_t g = (n + 255) / 256;
    if (g > (int64_t)num_cus() * 16) g = (int64_t)num_cus() * 16;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_to_c64_vec, dim3((unsigned)g), dim3(256), 0, stream(), n, in, out, run);
    EDX_LAUNCH_CHECK();
    return EDX_OK;
}

int spmm_launch(const edx_csr *a, int precision, int nrhs, const void *X, double2 *Y,
                const double2 *Z) {
    if (a->m == 0 || nrhs == 0) return EDX_OK;
    const int G = lanes_for(a) > 16 ? 16 : lanes_for(a);
    constexpr int NR = 8;
    int64_t blocks = (a->m * G + 255) / 256;
    int64_t cap = (int64_t)num_cus() * 8;
    if (blocks > cap) blocks = cap;
    dim3 grid((unsigned)blocks, (unsigned)((nrhs + NR - 1) / NR));
    hipEvent_t pe = prof_start(precision, EDX_KERNEL_SPMM_CM,
                               spmv_bytes(a, precision, nrhs, Z != nullptr, false), 0.0, nrhs,
                               csr_alg_bytes(a, nrhs, Z != nullptr));
#define EDX_SPMM_CASE(GG)                                                                       \
    case GG:                                                                                    \
        if (precision == EDX_C128)                                                              \
            hipLaunchKernelGGL((k_spmm<GG, NR, double2, double2>), grid, dim3(256), 0, stream(),   \
                               a->m, a->n, a->d_rowptr, a->d_cols, a->d_vals, (const double2 *)X, \
                               nrhs, Y, Z);                                                     \
        else                                                                                    \
            hipLaunchKernelGGL((k_spmm<GG, NR, float2, float2>), grid, dim3(256), 0, stream(),     \
                               a->m, a->n, a->d_rowptr, a->d_cols, a->d_vals32, (const float2 *)X, \
                               nrhs, Y, Z);                                                     \
        break;
    switch (G) {
        EDX_SPMM_CASE(4)
        EDX_SPMM_CASE(8)
        EDX_SPMM_CASE(16)
        default: return set_error(EDX_ERR_ARG, "unsupported lanes %d", G);
    }
#undef EDX_SPMM_CASE
    EDX_LAUNCH_CHECK();
    prof_stop(pe);
    return EDX_OK;
}

}  // namespace edx

extern "C" {

int edx_profile(int enable) {
    if (enable) {
        if (edx::g_ev.size() < 2 * edx::PROF_CAP) {
            for (size_t i = edx::g_ev.size(); i < 2 * edx::PROF_CAP; ++i) {
                hipEvent_t e;
                EDX_HIP(hipEventCreate(&e));
                edx::g_ev.push_back(e);
            }
            edx::g_rec.resize(edx::PROF_CAP);
        }
        edx::g_prof_n = 0;
        for (auto &plane : edx::g_sum)
            for (auto &row : plane)
                for (auto &s : row) s = edx::ProfSum{};
    }
    edx::g_prof = enable != 0;
    return EDX_OK;
}

int edx_profile_read(int precision, int kernel, int64_t *launches, double *total_ms,
                     double *total_bytes) {
    return edx_profile_read_rhs(precision, kernel, -1, launches, total_ms, total_bytes);
}

int edx_profile_read_rhs(int precision, int kernel, int nrhs, int64_t *launches,
                         double *total_ms, double *total_bytes) {
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    edx::ProfSum s;
    EDX_TRY(edx::prof_total(precision, kernel, nrhs, &s));
    if (launches) *launches = s.n;
    if (total_ms) *total_ms = s.ms;
    if (total_bytes) *total_bytes = s.bytes;
    return EDX_OK;
}

int edx_profile_read_alg(int precision, int kernel, int nrhs, double *alg_bytes) {
    EDX_CHECK(alg_bytes, EDX_ERR_ARG, "null argument");
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    edx::ProfSum s;
    EDX_TRY(edx::prof_total(precision, kernel, nrhs, &s));
    *alg_bytes = s.abytes;
    return EDX_OK;
}

int edx_profile_read_flops(int precision, int kernel, double *total_flops) {
    EDX_CHECK(total_flops, EDX_ERR_ARG, "null argument");
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    edx::ProfSum s;
    EDX_TRY(edx::prof_total(precision, kernel, -1, &s));
    *total_flops = s.flops;
    return EDX_OK;
}

int edx_spmv(const edx_csr *a, int precision, const void *d_x, void *d_y) {
    EDX_CHECK(a && d_x && d_y, EDX_ERR_ARG, "null argument");
    EDX_CHECK(precision == EDX_C128 || (precision == EDX_C64 && a->d_vals32) ||
                  (precision == EDX_C64D && a->d_diag) || (precision == EDX_C128I && a->d_off_idx),
              EDX_ERR_ARG, "precision %d not prepared on this matrix", precision);
    SpmvEpi e{};
    e.y = (double2 *)d_y;
    e.z = (const double2 *)d_y;   // y += A x
    const void *x = d_x;
    float2 *x32 = nullptr;
    if (precision == EDX_C64) {
        x32 = (float2 *)edx::dmalloc((a->n ? a->n : 1) * sizeof(float2), "x c64");
        EDX_CHECK(x32, EDX_ERR_NOMEM, "%s", edx_last_error());
        int rc = edx::to_c64(a->n, (const double2 *)d_x, x32, nullptr);
        if (rc) {
            edx::dfree(x32);
            return rc;
        }
        x = x32;
    }
    int rc = edx::spmv_launch(a, precision, x, e, 0, nullptr);
    hipError_t s = hipStreamSynchronize(edx::stream());
    edx::dfree(x32);
    if (rc) return rc;
    EDX_HIP(s);
    return EDX_OK;
}

int edx_spmm(const edx_csr *a, int precision, int nrhs, const void *d_X, void *d_Y) {
    EDX_CHECK(a && d_X && d_Y && nrhs >= 0, EDX_ERR_ARG, "bad argument");
    EDX_CHECK(precision == EDX_C128 || (precision == EDX_C64 && a->d_vals32), EDX_ERR_ARG,
              "precision %d not available on this matrix", precision);
    const void *X = d_X;
    float2 *x32 = nullptr;
    if (precision == EDX_C64) {
        int64_t len = a->n * (int64_t)nrhs;
        x32 = (float2 *)edx::dmalloc((len ? len : 1) * sizeof(float2), "X c64");
        EDX_CHECK(x32, EDX_ERR_NOMEM, "%s", edx_last_error());
        int rc = edx::to_c64(len, (const double2 *)d_X, x32, nullptr);
        if (rc) {
            edx::dfree(x32);
            return rc;
        }
        X = x32;
    }
    int rc = edx::spmm_launch(a, precision, nrhs, X, (double2 *)d_Y, (const double2 *)d_Y);
    hipError_t s = hipStreamSynchronize(edx::stream());
    edx::dfree(x32);
    if (rc) return rc;
    EDX_HIP(s);
    return EDX_OK;
}

}  // extern "C"
