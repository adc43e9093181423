// Multi-RHS solvers over one CSR (part of edx_kernels.hip).
//
// K right-hand sides are stored interleaved, X[r*K + k], so one pass over the
// matrix (values, column indices, row pointers) serves all K products and every
// gathered x line carries K useful values.  The SpMM epilogue is fused and
// per-RHS:
//     y_k[r] = c0_k (H x_k)[r] + c1_k z_k[r] + c2_k x_k[r]
// (the c2 term applies the shift of sigma_k I - H without a shifted matrix),
// with optional per-RHS dot partials sum_r conj(d_k[r]) y_k[r].
//
// Two SpMM schedules:
//   * K < 8: lanes split a row's nonzeros (G per row, from the mean row
//     length), up to 4 RHS per pass;
//   * K >= 8: 16-lane groups own rows and lanes own right-hand sides (lane l
//     handles RHS l, l+16, ...), so the row's values and column indices are
//     loaded once for all K and the gathers X[c*K + k] are contiguous.
// Element-wise kernels use blocks of 64 row-slots x min(K,16) RHS-slots
// (RHS fastest, so consecutive threads touch consecutive elements); per-RHS
// reductions run over a fixed row partition, independent of K.
//
// On top of it:
//   * batched Krylov (build_krylov_mp, src/lanczos.f90:210-303, per RHS), used
//     for the three XAS components and the RIXS final-state spectra;
//   * batched MINRES for (sigma_k I - H_n) x_k = b_k (pminres_csr,
//     src/linsys.f90:3-128, per RHS) -- the RIXS intermediate solves for many
//     incident energies and polarizations in one recurrence.
//
// Value formats (edx_precision): C128 exact complex128; C64 complex64 values,
// complex64 x, FP32 accumulation; C64D complex64 off-diagonal values with the
// exact complex128 diagonal kept apart, complex128 x and FP64 accumulation.

namespace {

constexpr int BKT = 4;      // right-hand sides per pass of the row-split SpMM
constexpr int RL_MAXK = 128;  // right-hand sides per pass of the RHS-parallel SpMM

struct BEpi {
    double2 *y;            // interleaved m x K
    const double2 *z;      // interleaved or nullptr
    const double2 *c0;     // per RHS [K] or nullptr (1)
    const double2 *c1;     // per RHS [K] or nullptr (1)
    const double2 *c2;     // per RHS [K] or nullptr (no x term)
    const double2 *dvec;   // interleaved or nullptr
    double2 *partials;     // [nblocks * K]
};

__device__ __forceinline__ double2 epilogue(const BEpi &e, int k, int64_t o, double2 hx,
                                            const double2 *X64) {
    double2 out = e.c0 ? cmul(e.c0[k], hx) : hx;
    if (e.z) out = cadd(out, e.c1 ? cmul(e.c1[k], e.z[o]) : e.z[o]);
    if (e.c2) out = cadd(out, cmul(e.c2[k], X64[o]));
    return out;
}

// The same with the addend read non-temporally (each element is read once).
__device__ __forceinline__ double2 epilogue_nt(const BEpi &e, int k, int64_t o, double2 hx,
                                               const double2 *X64) {
    double2 out = e.c0 ? cmul(e.c0[k], hx) : hx;
    if (e.z) {
        const double2 z = nt_load2(e.z + o);
        out = cadd(out, e.c1 ? cmul(e.c1[k], z) : z);
    }
    if (e.c2) out = cadd(out, cmul(e.c2[k], X64[o]));
    return out;
}

__device__ __forceinline__ uint16_t shfl16(uint16_t v, int src) {
    return (uint16_t)__shfl((int)v, src, 16);
}
__device__ __forceinline__ uint32_t shfl16(uint32_t v, int src) {
    return (uint32_t)__shfl((int)v, src, 16);
}
__device__ __forceinline__ float2 shfl16(float2 v, int src) {
    return make_float2(__shfl(v.x, src, 16), __shfl(v.y, src, 16));
}
__device__ __forceinline__ double2 shfl16(double2 v, int src) {
    return make_double2(__shfl(v.x, src, 16), __shfl(v.y, src, 16));
}

// Pipelined RHS-parallel schedule: 16-lane groups own rows and lane l owns
// right-hand side blockIdx.y * 16 + l, as in k_spmm_rl, on a capped grid whose
// workgroups of one XCD (blockIdx.x % 8) walk one contiguous eighth of the
// rows.  Lane l holds the column and value numbers of nonzeros l and l + 16 of
// its group's row (one coalesced load per row instead of a broadcast per
// nonzero), hands them to the group by shuffles while each lane gathers its own
// right-hand side, and loads those of the group's next row first.  Nonzeros
// beyond 32 in a row are read in place.  LTAB (indexed forms): the value table
// (ntab entries) is staged in LDS and read from there.
template <class VP, class XT, class ACC, bool SPLIT, bool LTAB = false>
__global__ __launch_bounds__(256) void k_spmm_rlp(int64_t m, int K, const int64_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ cols, VP vals,
                                                  const double2 *__restrict__ diag,
                                                  const XT *__restrict__ X,
                                                  const double2 *__restrict__ X64, BEpi e,
                                                  int ntab = 0) {
    using PV = PipeVal<VP>;
    using H = typename PV::type;
    __shared__ double2 red[16 * 16];
    extern __shared__ double2 s_tab[];
    if constexpr (LTAB) {
        for (int i = threadIdx.x; i < ntab; i += 256) s_tab[i] = vals.tab[i];
        __syncthreads();
    }
    auto value = [&](H h) {
        if constexpr (LTAB) return s_tab[h];
        else return PV::value(vals, h);
    };
    const int l = threadIdx.x & 15;
    const int grp = threadIdx.x >> 4;
    const int k = blockIdx.y * 16 + l;
    const bool kv = k < K;
    const int64_t per_x = gridDim.x / 8, xcd = blockIdx.x % 8, w = blockIdx.x / 8;
    const int64_t rows_x = (m + 7) / 8;
    const int64_t rbeg = xcd * rows_x, rend = min(m, rbeg + rows_x);
    const int64_t stride = per_x * 16;
    double2 dot = make_double2(0, 0);
    int64_t r = rbeg + w * 16 + grp;
    int64_t b = 0, end = 0;
    if (r < rend) {
        b = rowptr[r];
        end = rowptr[r + 1];
    }
    int ca, cb;
    H ha, hb;
    PV::fetch(cols, vals, b + l, b + l < end, ca, ha);
    PV::fetch(cols, vals, b + 16 + l, b + 16 + l < end, cb, hb);
    while (__any(r < rend)) {
        const int64_t rn = r + stride;
        int64_t bn = 0, endn = 0;
        if (rn < rend) {
            bn = rowptr[rn];
            endn = rowptr[rn + 1];
        }
        int can, cbn;
        H han, hbn;
        PV::fetch(cols, vals, bn + l, bn + l < endn, can, han);
        PV::fetch(cols, vals, bn + 16 + l, bn + 16 + l < endn, cbn, hbn);
        if (r < rend) {
            ACC sr = 0, si = 0;
            const int cnt = (int)(end - b);
            const int nf = cnt < 32 ? cnt : 32;
            // four nonzeros at a time: their columns and values by shuffles,
            // then four gathers in flight before the sums (in nonzero order)
            int t = 0;
            for (; t + 4 <= nf; t += 4) {
                int c[4];
                decltype(value(ha)) a[4];
                XT xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int tt = t + u, src = tt & 15;
                    c[u] = __shfl(tt < 16 ? ca : cb, src, 16);
                    a[u] = value(shfl16(tt < 16 ? ha : hb, src));
                }
                if (kv) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) xv[u] = X[(int64_t)c[u] * K + k];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        sr += (ACC)a[u].x * (ACC)xv[u].x - (ACC)a[u].y * (ACC)xv[u].y;
                        si += (ACC)a[u].x * (ACC)xv[u].y + (ACC)a[u].y * (ACC)xv[u].x;
                    }
                }
            }
            for (; t < nf; ++t) {
                const int src = t & 15;
                const int c = __shfl(t < 16 ? ca : cb, src, 16);
                const auto a = value(shfl16(t < 16 ? ha : hb, src));
                if (kv) {
                    const XT xv = X[(int64_t)c * K + k];
                    sr += (ACC)a.x * (ACC)xv.x - (ACC)a.y * (ACC)xv.y;
                    si += (ACC)a.x * (ACC)xv.y + (ACC)a.y * (ACC)xv.x;
                }
            }
            for (int64_t j = b + 32; j < end; ++j) {
                int c;
                H h;
                PV::fetch(cols, vals, j, true, c, h);
                const auto a = value(h);
                if (kv) {
                    const XT xv = X[(int64_t)c * K + k];
                    sr += (ACC)a.x * (ACC)xv.x - (ACC)a.y * (ACC)xv.y;
                    si += (ACC)a.x * (ACC)xv.y + (ACC)a.y * (ACC)xv.x;
                }
            }
            if (kv) {
                // single-use operands and the output non-temporal, so the
                // reused lines of X stay in L2 (2 % faster, K=16 over H_i)
                const int64_t o = r * K + k;
                double2 hx = make_double2((double)sr, (double)si);
                if (SPLIT) hx = cadd(hx, cmul(nt_load2(diag + r), X64[o]));
                const double2 out = epilogue_nt(e, k, o, hx, X64);
                nt_store2(e.y + o, out);
                if (e.dvec) dot = cadd(dot, cconjmul(nt_load2(e.dvec + o), out));
            }
        }
        r = rn;
        b = bn;
        end = endn;
        ca = can;
        cb = cbn;
        ha = han;
        hb = hbn;
    }
    if (e.dvec) {
        red[grp * 16 + l] = dot;
        __syncthreads();
        if (threadIdx.x < 16) {
            const int kk = blockIdx.y * 16 + threadIdx.x;
            if (kk < K) {
                double2 s = make_double2(0, 0);
                for (int g = 0; g < 16; ++g) s = cadd(s, red[g * 16 + threadIdx.x]);
                e.partials[(int64_t)blockIdx.x * K + kk] = s;
            }
        }
    }
}

// Row-split schedule: G lanes per row, KT right-hand sides per pass (blockIdx.y).
template <int G, int KT, class VP, class XT, class ACC, bool SPLIT>
__global__ __launch_bounds__(256) void k_spmm_il(int64_t m, int K, const int64_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ cols, VP vals,
                                                 const double2 *__restrict__ diag,
                                                 const XT *__restrict__ X,
                                                 const double2 *__restrict__ X64, BEpi e) {
    const int lig = threadIdx.x & (G - 1);
    const int64_t group = ((int64_t)blockIdx.x * 256 + threadIdx.x) / G;
    const int64_t ngroups = (int64_t)gridDim.x * (256 / G);
    const int k0 = blockIdx.y * KT;
    double2 dot[KT];
#pragma unroll
    for (int q = 0; q < KT; ++q) dot[q] = make_double2(0, 0);
    for (int64_t r = group; r < m; r += ngroups) {
        const int64_t b = rowptr[r], end = rowptr[r + 1];
        ACC sr[KT], si[KT];
#pragma unroll
        for (int q = 0; q < KT; ++q) sr[q] = si[q] = (ACC)0;
        for (int64_t j = b + lig; j < end; j += G) {
            const auto a = nt_load2(vals + j);
            const int64_t c = nt_load(cols + j);
            const XT *xr = X + c * K + k0;
#pragma unroll
            for (int q = 0; q < KT; ++q) {
                if (k0 + q < K) {
                    const XT xv = xr[q];
                    sr[q] += (ACC)a.x * (ACC)xv.x - (ACC)a.y * (ACC)xv.y;
                    si[q] += (ACC)a.x * (ACC)xv.y + (ACC)a.y * (ACC)xv.x;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < KT; ++q) {
#pragma unroll
            for (int off = G / 2; off > 0; off >>= 1) {
                sr[q] += __shfl_xor(sr[q], off, G);
                si[q] += __shfl_xor(si[q], off, G);
            }
        }
        if (lig == 0) {
            double2 dg = make_double2(0, 0);
            if (SPLIT) dg = diag[r];
#pragma unroll
            for (int q = 0; q < KT; ++q) {
                const int k = k0 + q;
                if (k >= K) break;
                const int64_t o = r * K + k;
                double2 hx = make_double2((double)sr[q], (double)si[q]);
                if (SPLIT) hx = cadd(hx, cmul(dg, X64[o]));
                const double2 out = epilogue(e, k, o, hx, X64);
                e.y[o] = out;
                if (e.dvec) dot[q] = cadd(dot[q], cconjmul(e.dvec[o], out));
            }
        }
    }
    if (e.dvec) {
#pragma unroll
        for (int q = 0; q < KT; ++q) {
            double2 s = block_sum_256(dot[q]);
            if (threadIdx.x == 0 && k0 + q < K) e.partials[(int64_t)blockIdx.x * K + k0 + q] = s;
        }
    }
}

// RHS-parallel schedule: 16 groups of 16 lanes per block, one row per group,
// lane l owns right-hand sides k0 + l + 16 q (q < NPL); a pass (blockIdx.y)
// covers a tile of 16 NPL right-hand sides, k0 = blockIdx.y * 16 NPL.
// With xmap (gridDim.x a multiple of 8), the workgroups the hardware places on
// one XCD (blockIdx.x % 8) sweep one contiguous eighth of the rows, so rows
// that gather the same X lines share that XCD's L2; otherwise consecutive
// 16-row groups go to consecutive workgroups, i.e. to all eight XCDs.
template <int NPL, class VP, class XT, class ACC, bool SPLIT>
__global__ __launch_bounds__(256) void k_spmm_rl(int64_t m, int K, const int64_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ cols, VP vals,
                                                 const double2 *__restrict__ diag,
                                                 const XT *__restrict__ X,
                                                 const double2 *__restrict__ X64, BEpi e, int xmap) {
    __shared__ double2 red[16 * 16 * NPL];
    const int l = threadIdx.x & 15;
    const int grp = threadIdx.x >> 4;
    const int k0 = blockIdx.y * 16 * NPL;   // this pass's tile of 16 NPL right-hand sides
    int64_t rfirst = (int64_t)blockIdx.x * 16 + grp, rend = m, ngroups = (int64_t)gridDim.x * 16;
    if (xmap) {
        const int64_t per = (m + 7) / 8, xcd = blockIdx.x % 8;
        rfirst = xcd * per + (int64_t)(blockIdx.x / 8) * 16 + grp;
        rend = min(m, (xcd + 1) * per);
        ngroups = (int64_t)(gridDim.x / 8) * 16;
    }
    double2 dot[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) dot[q] = make_double2(0, 0);
    for (int64_t r = rfirst; r < rend; r += ngroups) {
        const int64_t b = rowptr[r], end = rowptr[r + 1];
        ACC sr[NPL], si[NPL];
#pragma unroll
        for (int q = 0; q < NPL; ++q) sr[q] = si[q] = (ACC)0;
        int64_t j = b;
        for (; j + 3 < end; j += 4) {
            decltype(nt_load2(vals)) a[4];
            int64_t c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = nt_load2(vals + j + u);
                c[u] = nt_load(cols + j + u);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int q = 0; q < NPL; ++q) {
                    const int k = k0 + l + 16 * q;
                    if (k < K) {
                        const XT xv = X[c[u] * K + k];
                        sr[q] += (ACC)a[u].x * (ACC)xv.x - (ACC)a[u].y * (ACC)xv.y;
                        si[q] += (ACC)a[u].x * (ACC)xv.y + (ACC)a[u].y * (ACC)xv.x;
                    }
                }
            }
        }
        for (; j < end; ++j) {
            const auto a = nt_load2(vals + j);
            const int64_t c = nt_load(cols + j);
#pragma unroll
            for (int q = 0; q < NPL; ++q) {
                const int k = k0 + l + 16 * q;
                if (k < K) {
                    const XT xv = X[c * K + k];
                    sr[q] += (ACC)a.x * (ACC)xv.x - (ACC)a.y * (ACC)xv.y;
                    si[q] += (ACC)a.x * (ACC)xv.y + (ACC)a.y * (ACC)xv.x;
                }
            }
        }
        double2 dg = make_double2(0, 0);
        if (SPLIT) dg = diag[r];
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
            const int k = k0 + l + 16 * q;
            if (k < K) {
                const int64_t o = r * K + k;
                double2 hx = make_double2((double)sr[q], (double)si[q]);
                if (SPLIT) hx = cadd(hx, cmul(dg, X64[o]));
                const double2 out = epilogue(e, k, o, hx, X64);
                e.y[o] = out;
                if (e.dvec) dot[q] = cadd(dot[q], cconjmul(e.dvec[o], out));
            }
        }
    }
    if (e.dvec) {
#pragma unroll
        for (int q = 0; q < NPL; ++q) red[(grp * NPL + q) * 16 + l] = dot[q];
        __syncthreads();
        for (int t = threadIdx.x; t < 16 * NPL; t += 256) {
            const int q = t / 16, ll = t % 16;
            const int k = k0 + ll + 16 * q;
            if (k >= K) continue;
            double2 s = make_double2(0, 0);
            for (int g = 0; g < 16; ++g) s = cadd(s, red[(g * NPL + q) * 16 + ll]);
            e.partials[(int64_t)blockIdx.x * K + k] = s;
        }
    }
}

int bspmm_blocks(int64_t m) {
    // fixed by the row count only
    int64_t blocks = (m * 16 + 255) / 256;
    int64_t cap = (int64_t)edx::num_cus() * 8;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return (int)blocks;
}

int pick_lanes(const edx_csr *a) {
    if (a->spmm_lanes) return a->spmm_lanes;
    double avg = a->m ? (double)a->nnz / (double)a->m : 0.0;
    if (avg < 6) return 4;
    if (avg < 20) return 8;
    if (avg < 80) return 16;
    return 32;
}

// ------------------------------------------------ element-wise layout helpers
//
// grid (eb, ceil(K/16)), block 64 * KTE threads, KTE = min(K, 16);
// thread -> (row slot s = tid / KTE, rhs k = blockIdx.y*16 + tid % KTE).

constexpr int ESLOTS = 64;

int elem_grid_x(int64_t n) {
    int64_t g = (n + ESLOTS - 1) / ESLOTS;
    int64_t cap = (int64_t)edx::num_cus() * 8;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

struct ELaunch {
    dim3 grid, block;
};

ELaunch elaunch(int64_t n, int K) {
    const int kte = K < 16 ? K : 16;
    return ELaunch{dim3((unsigned)elem_grid_x(n), (unsigned)((K + 15) / 16)), dim3(ESLOTS * kte)};
}

#define ELEM_SETUP                                                     \
    const int kte = blockDim.x / ESLOTS;                               \
    const int kk = threadIdx.x % kte, slot = threadIdx.x / kte;        \
    const int k = blockIdx.y * 16 + kk;                                \
    const bool kv = k < K;                                             \
    const int64_t rstride = (int64_t)gridDim.x * ESLOTS;               \
    const int64_t rstart = (int64_t)blockIdx.x * ESLOTS + slot;

// per-(block, k) sum over the 64 slots in slot order -> partials[blockIdx.x*K + k]
__device__ __forceinline__ void elem_reduce(double2 acc, int K, double2 *partials) {
    __shared__ double2 red[ESLOTS * 16];
    const int kte = blockDim.x / ESLOTS;
    const int kk = threadIdx.x % kte, slot = threadIdx.x / kte;
    const int k = blockIdx.y * 16 + kk;
    red[threadIdx.x] = acc;
    __syncthreads();
    if (slot == 0 && k < K) {
        double2 s = make_double2(0, 0);
        for (int q = 0; q < ESLOTS; ++q) s = cadd(s, red[q * kte + kk]);
        partials[(int64_t)blockIdx.x * K + k] = s;
    }
}

// partials of sum_r conj(a) b per RHS
__global__ void k_bdot(int64_t n, int K, const double2 *a, const double2 *b, double2 *partials) {
    ELEM_SETUP
    double2 s = make_double2(0, 0);
    if (kv)
        for (int64_t r = rstart; r < n; r += rstride) s = cadd(s, cconjmul(a[r * K + k], b[r * K + k]));
    elem_reduce(s, K, partials);
}

// out[k] = sum_b partials[b*K + k]  (one block per RHS, fixed order)
__global__ __launch_bounds__(256) void k_breduce(const double2 *partials, int nb, int K, double2 *out) {
    const int k = blockIdx.x;
    __shared__ double sx[256], sy[256];
    double2 s = make_double2(0, 0);
    for (int b = threadIdx.x; b < nb; b += 256) s = cadd(s, partials[(int64_t)b * K + k]);
    sx[threadIdx.x] = s.x;
    sy[threadIdx.x] = s.y;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sx[threadIdx.x] += sx[threadIdx.x + off];
            sy[threadIdx.x] += sy[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = make_double2(sx[0], sy[0]);
}

__global__ void k_il_to_c64(int64_t n, int K, const double2 *in, float2 *out) {
    ELEM_SETUP
    if (!kv) return;
    for (int64_t r = rstart; r < n; r += rstride)
        out[r * K + k] = make_float2((float)in[r * K + k].x, (float)in[r * K + k].y);
}

// column-major (K columns of length n) <-> interleaved (n rows of K)
__global__ void k_cm_to_il(int64_t n, int K, const double2 *cm, double2 *il) {
    ELEM_SETUP
    if (!kv) return;
    for (int64_t r = rstart; r < n; r += rstride) il[r * K + k] = cm[(int64_t)k * n + r];
}

__global__ void k_il_to_cm(int64_t n, int K, const double2 *il, double2 *cm) {
    ELEM_SETUP
    if (!kv) return;
    for (int64_t r = rstart; r < n; r += rstride) cm[(int64_t)k * n + r] = il[r * K + k];
}

// interleaved block of K (n rows) -> interleaved block of the Kn columns
// keep[0..Kn): launched with elaunch(n, Kn)
__global__ void k_il_keep(int64_t n, int K, int Kn, const int *keep, const double2 *in,
                          double2 *out) {
    const int kte = blockDim.x / ESLOTS;
    const int kk = threadIdx.x % kte, slot = threadIdx.x / kte;
    const int k = blockIdx.y * 16 + kk;
    if (k >= Kn) return;
    const int src = keep[k];
    const int64_t rstride = (int64_t)gridDim.x * ESLOTS;
    for (int64_t r = (int64_t)blockIdx.x * ESLOTS + slot; r < n; r += rstride)
        out[r * Kn + k] = in[r * K + src];
}

// a -= b
__global__ void k_bsub(int64_t n, int K, double2 *a, const double2 *b) {
    ELEM_SETUP
    if (!kv) return;
    for (int64_t r = rstart; r < n; r += rstride) a[r * K + k] = csub(a[r * K + k], b[r * K + k]);
}

// v_k /= sqrt(norm_k)
__global__ void k_bdiv_sqrt_norm(int64_t n, int K, double2 *v, const double *norm) {
    ELEM_SETUP
    if (!kv) return;
    const double d = sqrt(norm[k]);
    if (!(d > 0)) return;
    for (int64_t r = rstart; r < n; r += rstride) {
        const double2 t = v[r * K + k];
        v[r * K + k] = make_double2(t.x / d, t.y / d);
    }
}

// ------------------------------------------------------- batched Krylov

// The recurrence keeps its vectors unnormalized: the stored u_i is the
// residual w_i of the previous step, with v_i = u_i / s_i and s_i = beta_i
// (s_1 = 1 after the start vector is normalized).  The product of a step is
// y = (1/s_i) H u_i - (beta_i / s_{i-1}) u_{i-1} = H v_i - beta_i v_{i-1}, its
// dot <u_i|y> / s_i = alpha_i, and the residual y - (alpha_i / s_i) u_i is the
// next stored vector: no pass divides a vector by beta.
struct BKrylov {
    double *alpha;     // [maxn * K], alpha[(i-1)*K + k]
    double *beta;      // [(maxn+1) * K]
    double *norm;      // [K]
    int *neff;         // [K]
    int *run;          // [(maxn+1) * K]
    double2 *c1;       // [K]  -beta_i / s_{i-1}: coefficient of u_{i-1} in the product
    double2 *sums;     // [K]
    double2 *c0;       // [K]  1 / s_i: scale of H u_i in the product
    double *acoef;     // [K]  alpha_i / s_i: coefficient of u_i in the residual
    double *scur;      // [K]  s_i
};

// doubles of the per-RHS tail of the scalar block (norm .. scur) per RHS
constexpr int BK_TAIL = 1 + 2 + 2 + 2 + 1 + 1;

__global__ void k_bk_init(BKrylov s, int K, const double2 *sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    s.norm[k] = sums[k].x;
    s.c1[k] = make_double2(-0.0, -0.0);
    s.c0[k] = make_double2(1.0, 0.0);
    s.scur[k] = 1.0;
    s.neff[k] = 0;
    s.run[k] = sums[k].x > 0.0 ? 1 : 0;
}

__global__ void k_bk_alpha(BKrylov s, int K, int i, const double2 *sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    if (!s.run[(i - 1) * K + k]) return;
    const double inv = s.c0[k].x;
    const double a = sums[k].x * inv;
    s.alpha[(int64_t)(i - 1) * K + k] = a;
    s.acoef[k] = a * inv;
    s.neff[k] = i;
}

// y -= alpha_k x ; partial |y|^2
__global__ void k_baxpy_norm(int64_t n, int K, double2 *y, const double2 *x, const double *alpha_i,
                             double2 *partials) {
    ELEM_SETUP
    double2 s = make_double2(0, 0);
    if (kv) {
        const double al = alpha_i[k];
        for (int64_t r = rstart; r < n; r += rstride) {
            const int64_t o = r * K + k;
            const double2 v = make_double2(y[o].x - al * x[o].x, y[o].y - al * x[o].y);
            y[o] = v;
            s.x += v.x * v.x + v.y * v.y;
        }
    }
    elem_reduce(s, K, partials);
}

__global__ void k_bk_beta(BKrylov s, int K, int i, const double2 *sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    if (!s.run[(i - 1) * K + k]) {
        s.run[i * K + k] = 0;
        return;
    }
    const double b = sqrt(sums[k].x);
    s.beta[(int64_t)i * K + k] = b;
    s.c1[k] = make_double2(-b / s.scur[k], -0.0);
    s.scur[k] = b;
    s.c0[k] = make_double2(b > 0.0 ? 1.0 / b : 0.0, 0.0);
    s.run[i * K + k] = fabs(b) < 1e-10 ? 0 : 1;   // src/lanczos.f90:290-293
}

// ------------------------------------------------------- batched MINRES

// HIP-graph replay of fixed launch sequences (EDX_NO_GRAPH=1 turns it off)
bool graphs_enabled() {
    const char *off = getenv("EDX_NO_GRAPH");
    return !(off && off[0] == '1');
}

struct BMinres {
    double *beta0;   // [K]
    double2 *sigma0, *sigma1, *gamma0, *gamma1, *eta, *res;
    double2 *alpha, *rho2, *rho3, *inv_rho1, *xcoef;
    int *run;            // [(maxit+1) * K]
    int *iters;          // [K]
    double tol;
    int maxit;
    const int *istep;    // nullptr: the iteration is the launch argument i; else *istep
};

// the iteration a MINRES kernel works on
__device__ __forceinline__ int bm_step(const BMinres &s, int i) { return s.istep ? *s.istep : i; }

__global__ void k_bm_next(int *istep) { *istep += 1; }

__global__ void k_bm_init(BMinres s, int K, const double2 *sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const double b0 = sqrt(sums[k].x);
    s.beta0[k] = b0;
    s.sigma0[k] = s.sigma1[k] = make_double2(0, 0);
    s.gamma0[k] = s.gamma1[k] = make_double2(1, 0);
    s.eta[k] = make_double2(b0, 0);
    s.res[k] = make_double2(b0, 0);
    s.iters[k] = 0;
    s.run[k] = (b0 > s.tol && 0 < s.maxit) ? 1 : 0;
}

__global__ void k_bm_divbeta(int64_t n, int K, double2 *v, BMinres s, int i) {
    ELEM_SETUP
    if (!kv || !s.run[(int64_t)bm_step(s, i) * K + k]) return;
    const double d = s.beta0[k];
    for (int64_t r = rstart; r < n; r += rstride) {
        const double2 t = v[r * K + k];
        v[r * K + k] = make_double2(t.x / d, t.y / d);
    }
}

__global__ void k_bm_alpha(BMinres s, int K, int i, const double2 *sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    i = bm_step(s, i);
    if (k >= K || !s.run[i * K + k]) return;
    s.alpha[k] = sums[k];
}

// vp = tmp - alpha vc - beta0 vp ; partial |vp|^2   (src/linsys.f90:87)
__global__ void k_bm_vupdate(int64_t n, int K, const double2 *tmp, const double2 *vc, double2 *vp,
                             BMinres s, int i, double2 *partials) {
    ELEM_SETUP
    double2 acc = make_double2(0, 0);
    i = bm_step(s, i);
    if (kv && s.run[(int64_t)i * K + k]) {
        const double2 al = s.alpha[k];
        const double b0 = s.beta0[k];
        for (int64_t r = rstart; r < n; r += rstride) {
            const int64_t o = r * K + k;
            const double2 v = csub(csub(tmp[o], cmul(al, vc[o])), cscale(vp[o], b0));
            vp[o] = v;
            acc.x += v.x * v.x + v.y * v.y;
        }
    }
    elem_reduce(acc, K, partials);
}

__global__ void k_bm_scalars(BMinres s, int K, int i, const double2 *sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    i = bm_step(s, i);
    if (k >= K) return;
    if (!s.run[i * K + k]) {
        s.run[(i + 1) * K + k] = 0;
        return;
    }
    const double beta1 = sqrt(sums[k].x);
    const double beta0 = s.beta0[k];
    const double2 alpha = s.alpha[k];
    const double2 g0 = s.gamma0[k], g1 = s.gamma1[k], sg0 = s.sigma0[k], sg1 = s.sigma1[k];
    const double2 delta = csub(cmul(g0, alpha), cscale(cmul(g1, sg0), beta0));
    const double2 rho1 = csqrt_dev(cadd(cmul(delta, delta), make_double2(beta1 * beta1, 0.0)));
    const double2 rho2 = cadd(cmul(sg0, alpha), cscale(cmul(g1, g0), beta0));
    const double2 rho3 = cscale(sg1, beta0);
    s.gamma1[k] = g0;
    s.sigma1[k] = sg0;
    const double2 ng0 = cdiv(delta, rho1), ns0 = cdiv(make_double2(beta1, 0.0), rho1);
    s.gamma0[k] = ng0;
    s.sigma0[k] = ns0;
    s.rho2[k] = rho2;
    s.rho3[k] = rho3;
    s.inv_rho1[k] = cdiv(make_double2(1.0, 0.0), rho1);
    s.xcoef[k] = cmul(ng0, s.eta[k]);
    const double as0 = hypot(ns0.x, ns0.y);
    s.res[k] = cscale(s.res[k], as0);
    const double2 eta = s.eta[k];
    s.eta[k] = make_double2(-(ns0.x * eta.x - ns0.y * eta.y), -(ns0.x * eta.y + ns0.y * eta.x));
    s.beta0[k] = beta1;
    s.iters[k] = i + 1;
    s.run[(i + 1) * K + k] = (hypot(s.res[k].x, s.res[k].y) > s.tol && i + 1 < s.maxit) ? 1 : 0;
}

// w_prev = (v_curr - rho3 w_prev - rho2 w_curr) / rho1 ; x += gamma0 eta w_prev
__global__ void k_bm_wx(int64_t n, int K, const double2 *vc, double2 *wp, const double2 *wc,
                        double2 *x, BMinres s, int i) {
    ELEM_SETUP
    if (!kv || !s.run[(int64_t)bm_step(s, i) * K + k]) return;
    const double2 r3 = s.rho3[k], r2 = s.rho2[k], ir1 = s.inv_rho1[k], xc = s.xcoef[k];
    for (int64_t r = rstart; r < n; r += rstride) {
        const int64_t o = r * K + k;
        const double2 w = cmul(csub(csub(vc[o], cmul(r3, wp[o])), cmul(r2, wc[o])), ir1);
        wp[o] = w;
        x[o] = cadd(x[o], cmul(xc, w));
    }
}

int elem_blocks(int64_t n) {
    int64_t g = (n + 255) / 256;
    int64_t cap = (int64_t)edx::num_cus() * 16;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// the split (diagonal + complex64 off-diagonal) form of a square CSR
__global__ void k_split_count(int64_t m, const int64_t *rowptr, const int32_t *cols, int64_t *cnt) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = 0;
        for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) c += (cols[j] != r);
        cnt[r] = c;
    }
}

__global__ void k_split_fill(int64_t m, const int64_t *rowptr, const int32_t *cols, const double2 *vals,
                             const int64_t *orow, int32_t *ocols, float2 *ovals, double2 *diag) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = orow[r];
        double2 d = make_double2(0, 0);
        for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            if (cols[j] == r) {
                d = vals[j];
            } else {
                ocols[o] = cols[j];
                ovals[o] = herm_round(rowptr, cols, vals, r, j);
                ++o;
            }
        }
        diag[r] = d;
    }
}

// ---- indexed form: a hash table of the distinct off-diagonal values
constexpr int VTAB_SLOTS = 1 << 16;
constexpr int VTAB_PROBES = 1024;   // a longer probe means the table is too full

__device__ __forceinline__ uint64_t mix64(uint64_t z) {   // splitmix64 finalizer
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t value_fp(double2 v) {
    const uint64_t f = mix64((uint64_t)__double_as_longlong(v.x) ^
                             mix64((uint64_t)__double_as_longlong(v.y) + 0x9e3779b97f4a7c15ull));
    return f ? f : 1ull;
}

// claim (or find) the slot of every off-diagonal value (every value when
// skip_diag is 0) by its 64-bit fingerprint (compare-and-swap on an empty
// slot, linear probing) and store the value there; flag bit 1 = table full
__global__ void k_vt_insert(int64_t m, const int64_t *rowptr, const int32_t *cols,
                            const double2 *vals, unsigned long long *fp, double2 *tab, int *flag,
                            int skip_diag) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        if (*(volatile int *)flag) return;   // already failed: stop early
        unsigned long long last = 0;         // the value this thread placed last
        for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            if (skip_diag && cols[j] == r) continue;
            const double2 v = vals[j];
            const unsigned long long f = value_fp(v);
            if (f == last) continue;
            unsigned h = (unsigned)f & (VTAB_SLOTS - 1);
            int k = 0;
            for (; k < VTAB_PROBES; ++k) {
                // a plain read first: most values are already in the table, and a
                // compare-and-swap on the few hot slots would serialize
                const unsigned long long cur = ((volatile unsigned long long *)fp)[h];
                if (cur == f) break;
                if (cur == 0ull) {
                    const unsigned long long old = atomicCAS(&fp[h], 0ull, f);
                    if (old == 0ull) {
                        tab[h] = v;   // the claiming thread stores the value
                        break;
                    }
                    if (old == f) break;
                }
                h = (h + 1) & (VTAB_SLOTS - 1);
            }
            if (k == VTAB_PROBES) {
                atomicOr(flag, 1);
                return;
            }
            last = f;
        }
    }
}

// the slot number of every off-diagonal value (every value when skip_diag is
// 0), row r's from idx[orow[r]] on; flag bit 2 = two different values share a
// fingerprint
__global__ void k_vt_index(int64_t m, const int64_t *rowptr, const int32_t *cols,
                           const double2 *vals, const int64_t *orow,
                           const unsigned long long *fp, const double2 *tab, uint16_t *idx,
                           int *flag, int skip_diag) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = orow[r];
        for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            if (skip_diag && cols[j] == r) continue;
            const double2 v = vals[j];
            const unsigned long long f = value_fp(v);
            unsigned h = (unsigned)f & (VTAB_SLOTS - 1);
            int k = 0;
            while (fp[h] != f && k < VTAB_PROBES) {
                h = (h + 1) & (VTAB_SLOTS - 1);
                ++k;
            }
            const double2 t = tab[h];
            if (k == VTAB_PROBES || __double_as_longlong(t.x) != __double_as_longlong(v.x) ||
                __double_as_longlong(t.y) != __double_as_longlong(v.y))
                atomicOr(flag, 2);
            idx[o++] = (uint16_t)h;
        }
    }
}

// compact the table: slot s (fingerprint set) -> entry remap[s], so the
// nvtab distinct values sit in nvtab consecutive entries
__global__ void k_vt_compact(const unsigned long long *fp, const double2 *tab, const uint16_t *remap,
                             double2 *out) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < VTAB_SLOTS; s += gridDim.x * blockDim.x)
        if (fp[s]) out[remap[s]] = tab[s];
}

__global__ void k_vt_pack(int64_t nnz, const int32_t *cols, const uint16_t *idx, uint32_t *pk) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * blockDim.x)
        pk[j] = (uint32_t)cols[j] | ((uint32_t)idx[j] << PK_COL_BITS);
}

__global__ void k_vt_renumber(int64_t nnz, const uint16_t *remap, uint16_t *idx) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * blockDim.x)
        idx[j] = remap[idx[j]];
}

}  // namespace

namespace edx {

// schedule of a multi-RHS product: row-split (rl = 0, p lanes per row),
// RHS-parallel (rl = 1, tiles of 16 p right-hand sides), pipelined
// RHS-parallel (rl = 2, p workgroups per CU), for K <= 4 on a matrix with
// the slot-factored form in the indexed precision the factored product
// (rl = 3, p workgroups per CU), or the pipelined RHS-parallel product over
// the packed indexed words with the value table in LDS (rl = 4, p
// workgroups per CU)
struct BSched {
    int rl;
    int p;
};
int fac_spmm_launch(const edx_csr *a, int K, const double2 *X, const BEpi &e, int bpc);
double fac_bytes_k(const edx_csr *a, int K, bool z, bool dvec);
inline bool fac_spmm_ok(const edx_csr *a, int precision, int K) {
    return precision == EDX_C128I && a->d_fr_rowptr && a->nc_slots == 6 && K >= 2 && K <= 4;
}
inline bool packed_spmm_ok(const edx_csr *a, int precision) {
    return precision == EDX_C128I && a->d_off_pk && a->nvtab <= LDS_TAB_MAX;
}
int tune_spmv(edx_csr *a, int precision);
int tune_bspmm(const edx_csr *a, int precision, int K, BSched *sched);
int bsched_blocks(const edx_csr *a, BSched sched);

// One launch of the multi-RHS product with an explicit schedule: row-split
// (sched.rl == 0: sched.p lanes per row, tiles of BKT right-hand sides) or
// RHS-parallel (sched.rl == 1: tiles of 16 * sched.p right-hand sides).
// EDX_FORCE_BSCHED="rl:p" replaces the chosen multi-RHS schedule (rl 0
// row-split with p lanes, 1 RHS-parallel with tiles of 16 p for K >= 8, 2 and
// 4 pipelined RHS-parallel with p workgroups per CU, 3 slot-factored for
// K < 8): for tests that exercise every schedule on small matrices.
void forced_bsched(int K, BSched *sched) {
    const char *f = getenv("EDX_FORCE_BSCHED");
    int rl = 0, p = 0;
    if (!f || sscanf(f, "%d:%d", &rl, &p) != 2 || p <= 0) return;
    if (rl == 0 || rl == 2 || rl == 4 || (K >= 8 && rl == 1) || (K < 8 && rl == 3))
        *sched = BSched{rl, p};
}

// Workgroups (and dot partials per right-hand side) of a schedule's launch.
// The pipelined schedule needs a multiple of 8 and at most bspmm_blocks(m),
// the partials every caller provides.
int bsched_blocks(const edx_csr *a, BSched sched) {
    if (sched.rl >= 2) return (edx::num_cus() * sched.p + 7) / 8 * 8;
    return bspmm_blocks(a->m);
}

int bspmm_launch(const edx_csr *a, int precision, int K, const double2 *X64, const float2 *X32,
                 const BEpi &e, BSched sched) {
    const int nb = bspmm_blocks(a->m);
    if (sched.rl == 3) {   // slot-factored, sched.p workgroups per CU
        if (!fac_spmm_ok(a, precision, K))
            return set_error(EDX_ERR_ARG, "the factored multi-RHS product needs the slot-factored "
                             "form, the indexed precision and 2 <= K <= 4");
        if (bsched_blocks(a, sched) > nb)
            return set_error(EDX_ERR_ARG, "factored SpMM needs %d workgroups, %d available",
                             bsched_blocks(a, sched), nb);
        return fac_spmm_launch(a, K, X64, e, sched.p);
    }
    if (sched.rl == 4) {   // packed words, value table in LDS, sched.p workgroups per CU
        if (!packed_spmm_ok(a, precision))
            return set_error(EDX_ERR_ARG, "multi-RHS schedule 4 needs the packed indexed form");
        const int nbp = bsched_blocks(a, sched);
        if (nbp > nb)
            return set_error(EDX_ERR_ARG, "pipelined SpMM needs %d workgroups, %d available", nbp,
                             nb);
        dim3 grid((unsigned)nbp, (unsigned)((K + 15) / 16));
        hipLaunchKernelGGL((k_spmm_rlp<PkPtr, double2, double, true, true>), grid, dim3(256),
                           (size_t)a->nvtab * sizeof(double2), stream(), a->m, K, a->d_off_rowptr,
                           nullptr, PkPtr{a->d_off_pk, a->d_vtab}, a->d_diag, X64, X64, e,
                           (int)a->nvtab);
    } else if (sched.rl == 2) {   // pipelined RHS-parallel, sched.p workgroups per CU
        const int nbp = bsched_blocks(a, sched);
        if (nbp > nb)
            return set_error(EDX_ERR_ARG, "pipelined SpMM needs %d workgroups, %d available", nbp,
                             nb);
        dim3 grid((unsigned)nbp, (unsigned)((K + 15) / 16));
        if (precision == EDX_C128)
            hipLaunchKernelGGL((k_spmm_rlp<const double2 *, double2, double, false>), grid,
                               dim3(256), 0, stream(), a->m, K, a->d_rowptr, a->d_cols,
                               (const double2 *)a->d_vals, nullptr, X64, X64, e);
        else if (precision == EDX_C64)
            hipLaunchKernelGGL((k_spmm_rlp<const float2 *, float2, float, false>), grid, dim3(256),
                               0, stream(), a->m, K, a->d_rowptr, a->d_cols,
                               (const float2 *)a->d_vals32, nullptr, X32, X64, e);
        else if (precision == EDX_C128I)
            hipLaunchKernelGGL((k_spmm_rlp<IdxPtr, double2, double, true>), grid, dim3(256), 0,
                               stream(), a->m, K, a->d_off_rowptr, a->d_off_cols,
                               IdxPtr{a->d_off_idx, a->d_vtab}, a->d_diag, X64, X64, e);
        else
            hipLaunchKernelGGL((k_spmm_rlp<const float2 *, double2, double, true>), grid,
                               dim3(256), 0, stream(), a->m, K, a->d_off_rowptr, a->d_off_cols,
                               (const float2 *)a->d_off_vals32, a->d_diag, X64, X64, e);
    } else if (sched.rl) {
        const int npl = sched.p;
        dim3 grid((unsigned)nb, (unsigned)((K + 16 * npl - 1) / (16 * npl)));
        const int xmap = nb % 8 == 0 && xcd_map_enabled();
#define EDX_RL(NP)                                                                                 \
    case NP:                                                                                       \
        if (precision == EDX_C128)                                                                 \
            hipLaunchKernelGGL((k_spmm_rl<NP, const double2 *, double2, double, false>), grid,      \
                               dim3(256), 0, stream(), a->m, K, a->d_rowptr, a->d_cols,            \
                               (const double2 *)a->d_vals, nullptr, X64, X64, e, xmap);           \
        else if (precision == EDX_C64)                                                             \
            hipLaunchKernelGGL((k_spmm_rl<NP, const float2 *, float2, float, false>), grid,         \
                               dim3(256), 0, stream(), a->m, K, a->d_rowptr, a->d_cols,            \
                               (const float2 *)a->d_vals32, nullptr, X32, X64, e, xmap);          \
        else if (precision == EDX_C128I)                                                           \
            hipLaunchKernelGGL((k_spmm_rl<NP, IdxPtr, double2, double, true>), grid, dim3(256), 0,  \
                               stream(), a->m, K, a->d_off_rowptr, a->d_off_cols,                  \
                               IdxPtr{a->d_off_idx, a->d_vtab}, a->d_diag, X64, X64, e, xmap);     \
        else                                                                                       \
            hipLaunchKernelGGL((k_spmm_rl<NP, const float2 *, double2, double, true>), grid,        \
                               dim3(256), 0, stream(), a->m, K, a->d_off_rowptr, a->d_off_cols,    \
                               (const float2 *)a->d_off_vals32, a->d_diag, X64, X64, e, xmap);    \
        break;
        switch (npl) {
            EDX_RL(1)
            EDX_RL(2)
            EDX_RL(4)
            EDX_RL(8)
            default: return set_error(EDX_ERR_ARG, "bad RHS tile %d", 16 * npl);
        }
#undef EDX_RL
    } else {
        const int G = sched.p;
        dim3 grid((unsigned)nb, (unsigned)((K + BKT - 1) / BKT));
#define EDX_BSPMM(GG)                                                                              \
    case GG:                                                                                       \
        if (precision == EDX_C128)                                                                 \
            hipLaunchKernelGGL((k_spmm_il<GG, BKT, const double2 *, double2, double, false>), grid, \
                               dim3(256), 0, stream(), a->m, K, a->d_rowptr, a->d_cols,            \
                               (const double2 *)a->d_vals, nullptr, X64, X64, e);                 \
        else if (precision == EDX_C64)                                                             \
            hipLaunchKernelGGL((k_spmm_il<GG, BKT, const float2 *, float2, float, false>), grid,    \
                               dim3(256), 0, stream(), a->m, K, a->d_rowptr, a->d_cols,            \
                               (const float2 *)a->d_vals32, nullptr, X32, X64, e);                \
        else if (precision == EDX_C128I)                                                           \
            hipLaunchKernelGGL((k_spmm_il<GG, BKT, IdxPtr, double2, double, true>), grid,           \
                               dim3(256), 0, stream(), a->m, K, a->d_off_rowptr, a->d_off_cols,    \
                               IdxPtr{a->d_off_idx, a->d_vtab}, a->d_diag, X64, X64, e);           \
        else                                                                                       \
            hipLaunchKernelGGL((k_spmm_il<GG, BKT, const float2 *, double2, double, true>), grid,   \
                               dim3(256), 0, stream(), a->m, K, a->d_off_rowptr, a->d_off_cols,    \
                               (const float2 *)a->d_off_vals32, a->d_diag, X64, X64, e);          \
        break;
        switch (G) {
            EDX_BSPMM(4)
            EDX_BSPMM(8)
            EDX_BSPMM(16)
            EDX_BSPMM(32)
            default: return set_error(EDX_ERR_ARG, "unsupported lanes %d", G);
        }
#undef EDX_BSPMM
    }
    EDX_LAUNCH_CHECK();
    return EDX_OK;
}

// The schedule the mean row length and K suggest (used below TUNE_MIN_ROWS
// rows, where products are too short to time).
BSched bspmm_rule(const edx_csr *a, int K) {
    if (K >= 8) {
        const int kp = K < RL_MAXK ? K : RL_MAXK;
        return BSched{1, kp <= 16 ? 1 : (kp <= 32 ? 2 : (kp <= 64 ? 4 : 8))};
    }
    return BSched{0, pick_lanes(a)};
}

// Y = epilogue(H X) for K interleaved RHS.  For C64, X32 holds the complex64
// copy of X (interleaved); X64 is always the complex128 X (split diagonal and
// shift terms read it).
int bspmm(const edx_csr *a, int precision, int K, const double2 *X64, const float2 *X32,
          const BEpi &e, int *nblocks_out) {
    if (K == 1 && a->m >= TUNE_MIN_ROWS && !e.c2) {
        // one right-hand side of a large matrix: the tuned single-vector kernel
        // (its vector then fits the last-level cache, which an interleaved
        // block of several does not)
        SpmvEpi se{};
        se.y = e.y;
        se.z = e.z;
        se.c0 = e.c0;
        se.c1 = e.c1;
        se.dvec = e.dvec;
        se.partials = e.partials;
        return spmv_launch(a, precision, precision == EDX_C64 ? (const void *)X32 : (const void *)X64,
                           se, 0, nblocks_out);
    }
    if (nblocks_out) *nblocks_out = bspmm_blocks(a->m);
    if (a->m == 0 || K == 0) return EDX_OK;
    BSched sched = bspmm_rule(a, K);
    if (a->m >= TUNE_MIN_ROWS) EDX_TRY(tune_bspmm(a, precision, K, &sched));
    forced_bsched(K, &sched);
    if (sched.rl == 3 && !fac_spmm_ok(a, precision, K))
        return set_error(EDX_ERR_ARG, "the factored multi-RHS product needs the slot-factored "
                         "form, the indexed precision and 2 <= K <= 4");
    if (sched.rl == 4 && !packed_spmm_ok(a, precision))
        return set_error(EDX_ERR_ARG, "multi-RHS schedule 4 needs the packed indexed form");
    if (nblocks_out) *nblocks_out = bsched_blocks(a, sched);
    // the packed words read 4 bytes of index per nonzero instead of 6
    const double bytes = sched.rl == 3 ? fac_bytes_k(a, K, e.z != nullptr, e.dvec != nullptr)
                         : spmv_bytes_k(a, precision, K, e.z != nullptr, e.dvec != nullptr) -
                               (sched.rl == 4 ? 2.0 * (double)a->off_nnz : 0.0);
    hipEvent_t pe = prof_start(precision, sched.rl == 3 ? EDX_KERNEL_SPMM_FAC
                                          : sched.rl == 0 ? EDX_KERNEL_SPMM_IL
                                                          : EDX_KERNEL_SPMM_RL,
                               bytes, 0.0, K, csr_alg_bytes(a, K, e.z != nullptr));
    EDX_TRY(bspmm_launch(a, precision, K, X64, X32, e, sched));
    prof_stop(pe);
    return EDX_OK;
}

int csr_add_split(edx_csr *a) {
    if (a->d_diag) return EDX_OK;
    EDX_CHECK(a->m == a->n, EDX_ERR_ARG, "split form needs a square matrix");
    const int64_t m = a->m;
    int64_t *orow = (int64_t *)dmalloc((m + 1) * sizeof(int64_t), "split rowptr");
    double2 *diag = (double2 *)dmalloc((m ? m : 1) * sizeof(double2), "split diagonal");
    if (!orow || !diag) {
        dfree(orow);
        dfree(diag);
        return EDX_ERR_NOMEM;
    }
    const int g = elem_blocks(m);
    EDX_HIP(hipMemsetAsync(orow + m, 0, sizeof(int64_t), stream()));
    hipLaunchKernelGGL(k_split_count, dim3(g), dim3(256), 0, stream(), m, a->d_rowptr, a->d_cols, orow);
    EDX_LAUNCH_CHECK();
    EDX_TRY(exclusive_scan(orow, m + 1));
    int64_t onnz = 0;
    EDX_HIP(edx::copy_sync(&onnz, orow + m, sizeof(int64_t), hipMemcpyDeviceToHost));
    int32_t *oc = (int32_t *)dmalloc((onnz ? onnz : 1) * sizeof(int32_t), "split columns");
    float2 *ov = (float2 *)dmalloc((onnz ? onnz : 1) * sizeof(float2), "split values");
    if (!oc || !ov) {
        dfree(orow);
        dfree(diag);
        dfree(oc);
        dfree(ov);
        return EDX_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_split_fill, dim3(g), dim3(256), 0, stream(), m, a->d_rowptr, a->d_cols,
                       a->d_vals, orow, oc, ov, diag);
    EDX_LAUNCH_CHECK();
    EDX_HIP(hipStreamSynchronize(stream()));
    a->d_off_rowptr = orow;
    a->d_off_cols = oc;
    a->d_off_vals32 = ov;
    a->d_diag = diag;
    a->off_nnz = onnz;
    return EDX_OK;
}

// The indexed form: the split form's pattern and diagonal, and per
// off-diagonal nonzero the table slot of its exact value.
// Numbers the distinct values of a CSR (its off-diagonal ones with skip_diag)
// densely, 0 .. ntab-1 in hash-slot order: *tab_out gets the ntab values,
// *idx_out one number per numbered entry (row r's from orow[r] on, onnz in
// all).  EDX_ERR_CAPACITY when the values do not fit the VTAB_SLOTS-slot
// table or two of them share a fingerprint; nothing is kept then.
int value_numbering(int64_t m, const int64_t *rowptr, const int32_t *cols, const double2 *vals,
                    const int64_t *orow, int64_t onnz, int skip_diag, double2 **tab_out,
                    uint16_t **idx_out, int64_t *ntab_out) {
    auto *fp = (unsigned long long *)dmalloc(VTAB_SLOTS * sizeof(unsigned long long), "value fingerprints");
    auto *tab = (double2 *)dmalloc(VTAB_SLOTS * sizeof(double2), "value table");
    auto *idx = (uint16_t *)dmalloc((onnz ? onnz : 1) * sizeof(uint16_t), "value numbers");
    int *flag = (int *)dmalloc(sizeof(int), "index flag");
    int rc = EDX_OK;
    std::vector<unsigned long long> hfp;
    int hflag = 0;
    if (!fp || !tab || !idx || !flag) {
        rc = EDX_ERR_NOMEM;
        goto fail;
    }
    if (hipMemsetAsync(fp, 0, VTAB_SLOTS * sizeof(unsigned long long), stream()) != hipSuccess ||
        hipMemsetAsync(tab, 0, VTAB_SLOTS * sizeof(double2), stream()) != hipSuccess ||
        hipMemsetAsync(flag, 0, sizeof(int), stream()) != hipSuccess) {
        rc = set_error(EDX_ERR_HIP, "indexed form: memset failed");
        goto fail;
    }
    {
        const int g = elem_blocks(m);
        hipLaunchKernelGGL(k_vt_insert, dim3(g), dim3(256), 0, stream(), m, rowptr, cols, vals,
                           fp, tab, flag, skip_diag);
        if (hipGetLastError() != hipSuccess ||
            edx::copy_sync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = set_error(EDX_ERR_HIP, "indexed form: kernels failed");
            goto fail;
        }
        if (hflag & 1) {
            rc = set_error(EDX_ERR_CAPACITY,
                           "too many distinct %svalues for the %d-slot table: no indexed form",
                           skip_diag ? "off-diagonal " : "", VTAB_SLOTS);
            goto fail;
        }
        hipLaunchKernelGGL(k_vt_index, dim3(g), dim3(256), 0, stream(), m, rowptr, cols, vals,
                           orow, fp, tab, idx, flag, skip_diag);
        if (hipGetLastError() != hipSuccess ||
            edx::copy_sync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = set_error(EDX_ERR_HIP, "indexed form: kernels failed");
            goto fail;
        }
    }
    if (hflag & 2) {
        rc = set_error(EDX_ERR_CAPACITY, "two distinct values share a 64-bit fingerprint: no indexed form");
        goto fail;
    }
    hfp.resize(VTAB_SLOTS);
    if (edx::copy_sync(hfp.data(), fp, VTAB_SLOTS * sizeof(unsigned long long), hipMemcpyDeviceToHost) !=
        hipSuccess) {
        rc = set_error(EDX_ERR_HIP, "indexed form: download failed");
        goto fail;
    }
    {   // renumber the values 0 .. ntab-1 in slot order: a dense table of ntab entries
        std::vector<uint16_t> remap(VTAB_SLOTS, 0);
        int64_t k = 0;
        for (int sl = 0; sl < VTAB_SLOTS; ++sl)
            if (hfp[sl]) remap[sl] = (uint16_t)k++;
        auto *d_remap = (uint16_t *)dmalloc(VTAB_SLOTS * sizeof(uint16_t), "value renumbering");
        auto *tabc = (double2 *)dmalloc((size_t)std::max<int64_t>(k, 1) * sizeof(double2), "value table");
        if (!d_remap || !tabc) {
            dfree(d_remap);
            dfree(tabc);
            rc = EDX_ERR_NOMEM;
            goto fail;
        }
        bool ok = edx::copy_sync(d_remap, remap.data(), VTAB_SLOTS * sizeof(uint16_t), hipMemcpyHostToDevice) ==
                  hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(k_vt_compact, dim3(VTAB_SLOTS / 256), dim3(256), 0, stream(), fp, tab, d_remap,
                               tabc);
            if (onnz)
                hipLaunchKernelGGL(k_vt_renumber, dim3(elem_blocks(onnz)), dim3(256), 0, stream(),
                                   onnz, d_remap, idx);
            ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(stream()) == hipSuccess;
        }
        dfree(d_remap);
        dfree(tab);
        tab = tabc;
        if (!ok) {
            rc = set_error(EDX_ERR_HIP, "indexed form: renumbering failed");
            goto fail;
        }
        *ntab_out = k;
    }
    dfree(fp);
    dfree(flag);
    *tab_out = tab;
    *idx_out = idx;
    return EDX_OK;
fail:
    dfree(fp);
    dfree(tab);
    dfree(idx);
    dfree(flag);
    return rc;
}

int csr_add_indexed(edx_csr *a) {
    if (a->d_off_idx) return EDX_OK;
    EDX_TRY(csr_add_split(a));
    double2 *tab = nullptr;
    uint16_t *idx = nullptr;
    int64_t ntab = 0;
    EDX_TRY(value_numbering(a->m, a->d_rowptr, a->d_cols, a->d_vals, a->d_off_rowptr, a->off_nnz,
                            1, &tab, &idx, &ntab));
    a->nvtab = ntab;
    a->d_vtab = tab;
    a->d_off_idx = idx;
    // few values and columns: one 32-bit word per nonzero for the single-vector product
    if (a->n < (int64_t(1) << PK_COL_BITS) && a->nvtab <= (int64_t(1) << (32 - PK_COL_BITS)) &&
        a->off_nnz) {
        auto *pk = (uint32_t *)dmalloc((size_t)a->off_nnz * sizeof(uint32_t), "packed indices");
        if (pk) {   // optional: without the memory the unpacked form serves
            hipLaunchKernelGGL(k_vt_pack, dim3(elem_blocks(a->off_nnz)), dim3(256), 0, stream(),
                               a->off_nnz, a->d_off_cols, idx, pk);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream()) != hipSuccess) {
                dfree(pk);
                return set_error(EDX_ERR_HIP, "indexed form: packing failed");
            }
            a->d_off_pk = pk;
        }
    }
    return EDX_OK;
}

// Packed form of a whole (rectangular) CSR, for the projections of the
// multi-shift solve: one 32-bit word per nonzero, column | value number << 24,
// with its table of distinct values.  Optional: a matrix with 2^24 or more
// columns, more than 256 distinct values, or without the memory keeps
// none, and its users read the complex128 CSR.
int csr_add_packed_all(edx_csr *a) {
    if (a->d_apk || a->nnz == 0 || a->n >= (int64_t(1) << PK_COL_BITS)) return EDX_OK;
    const char *off = getenv("EDX_NO_PACKED_PROJ");
    if (off && off[0] == '1') return EDX_OK;
    double2 *tab = nullptr;
    uint16_t *idx = nullptr;
    int64_t ntab = 0;
    if (value_numbering(a->m, a->d_rowptr, a->d_cols, a->d_vals, a->d_rowptr, a->nnz, 0, &tab,
                        &idx, &ntab) != EDX_OK)
        return EDX_OK;
    auto *pk = ntab <= (int64_t(1) << (32 - PK_COL_BITS))
                   ? (uint32_t *)dmalloc((size_t)a->nnz * sizeof(uint32_t), "packed words")
                   : nullptr;
    if (pk) {
        hipLaunchKernelGGL(k_vt_pack, dim3(elem_blocks(a->nnz)), dim3(256), 0, stream(), a->nnz,
                           a->d_cols, idx, pk);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream()) != hipSuccess) {
            dfree(pk);
            dfree(tab);
            dfree(idx);
            return set_error(EDX_ERR_HIP, "packed form: packing failed");
        }
        a->d_apk = pk;
        a->d_atab = tab;
        a->natab = ntab;
    } else {
        dfree(tab);
    }
    dfree(idx);
    return EDX_OK;
}

int csr_add_factored(edx_csr *a);

int ensure_precision(edx_csr *a, int precision) {
    if (precision == EDX_C64) EDX_TRY(edx_csr_add_c64(a));
    else if (precision == EDX_C64D) EDX_TRY(csr_add_split(a));
    else if (precision == EDX_C128I) {
        EDX_TRY(csr_add_indexed(a));
        EDX_TRY(csr_add_factored(a));
    }
    else if (precision != EDX_C128) return set_error(EDX_ERR_ARG, "unknown precision %d", precision);
    if (a->m == a->n) EDX_TRY(tune_spmv(a, precision));
    return EDX_OK;
}

}  // namespace edx

namespace {

struct Dbuf {
    void *p = nullptr;
    ~Dbuf() { edx::dfree(p); }
    int alloc(size_t bytes, const char *what) {
        p = edx::dmalloc(bytes ? bytes : 16, what);
        return p ? EDX_OK : EDX_ERR_NOMEM;
    }
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

#define ELAUNCH(kernel, n, K, ...)                                                        \
    do {                                                                                  \
        const ELaunch el_ = elaunch((n), (K));                                            \
        hipLaunchKernelGGL(kernel, el_.grid, el_.block, 0, edx::stream(), __VA_ARGS__); \
    } while (0)

}  // namespace

// A batched Krylov recurrence that can be advanced in pieces: the vectors, the
// run flags and the (alpha, beta) history stay on the device between calls.
// The scalar arrays grow geometrically (device-to-device copy) when a call
// needs more steps than they hold.
struct edx_krylov_batch {
    const edx_csr *h = nullptr;
    int precision = 0;
    int K = 0;
    int64_t n = 0;
    int done = 0;        // recurrence steps taken
    int cap = 0;         // steps the scalar arrays hold
    bool live = true;    // some RHS has not broken down
    int curr = 0, prev = 1;
    Dbuf vb, x32, part, sc, flags;
    BKrylov st{};
};

namespace {

int kb_layout(edx_krylov_batch *q, int cap, const void *old_sc, const void *old_flags, int old_cap) {
    const int K = q->K;
    Dbuf sc, flags;
    // alpha, beta, then the per-RHS tail norm, c1, sums, c0, acoef, scur
    const size_t nsc = (size_t)cap * K + (size_t)(cap + 1) * K + (size_t)BK_TAIL * K;
    EDX_TRY(sc.alloc(nsc * sizeof(double), "batched scalars"));
    EDX_TRY(flags.alloc(((size_t)(cap + 1) * K + K) * sizeof(int), "batched flags"));
    hipStream_t s = edx::stream();
    EDX_HIP(hipMemsetAsync(sc.p, 0, nsc * sizeof(double), s));
    EDX_HIP(hipMemsetAsync(flags.p, 0, ((size_t)(cap + 1) * K + K) * sizeof(int), s));
    BKrylov st;
    st.alpha = sc.as<double>();
    st.beta = st.alpha + (size_t)cap * K;
    st.norm = st.beta + (size_t)(cap + 1) * K;
    st.c1 = (double2 *)(st.norm + K);    // 16-byte aligned: 2 (cap + 1) K doubles precede it
    st.sums = st.c1 + K;
    st.c0 = st.sums + K;
    st.acoef = (double *)(st.c0 + K);
    st.scur = st.acoef + K;
    st.run = flags.as<int>();
    st.neff = st.run + (size_t)(cap + 1) * K;
    if (old_sc) {
        const BKrylov &o = q->st;
        EDX_HIP(hipMemcpyAsync(st.alpha, o.alpha, (size_t)old_cap * K * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
        EDX_HIP(hipMemcpyAsync(st.beta, o.beta, (size_t)(old_cap + 1) * K * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
        EDX_HIP(hipMemcpyAsync(st.norm, o.norm, (size_t)BK_TAIL * K * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
        EDX_HIP(hipMemcpyAsync(st.run, o.run, (size_t)(old_cap + 1) * K * sizeof(int),
                               hipMemcpyDeviceToDevice, s));
        EDX_HIP(hipMemcpyAsync(st.neff, o.neff, K * sizeof(int), hipMemcpyDeviceToDevice, s));
        (void)old_flags;
    }
    std::swap(q->sc.p, sc.p);         // the old buffers are freed with the locals
    std::swap(q->flags.p, flags.p);
    q->st = st;
    q->cap = cap;
    EDX_HIP(hipStreamSynchronize(s));
    return EDX_OK;
}

int kb_start(edx_krylov_batch *q, const double2 *V0cm, int cap) {
    const int64_t n = q->n;
    const int K = q->K;
    hipStream_t s = edx::stream();
    EDX_TRY(q->vb.alloc((size_t)2 * n * K * sizeof(double2), "batched krylov vectors"));
    if (q->precision == EDX_C64)
        EDX_TRY(q->x32.alloc((size_t)n * K * sizeof(float2), "batched x32"));
    const int nbs = bspmm_blocks(n), nbe = elem_grid_x(n);
    EDX_TRY(q->part.alloc(std::max<size_t>((size_t)std::max(nbs, nbe) * K + 64,
                                           (size_t)edx::spmv_partials_len(q->h)) * sizeof(double2),
                          "partials"));
    EDX_TRY(kb_layout(q, cap, nullptr, nullptr, 0));
    const BKrylov &st = q->st;
    double2 *v0 = q->vb.as<double2>();
    EDX_HIP(hipMemsetAsync(q->vb.p, 0, (size_t)2 * n * K * sizeof(double2), s));
    const int kb = (K + 63) / 64;
    ELAUNCH(k_cm_to_il, n, K, n, K, V0cm, v0);
    ELAUNCH(k_bdot, n, K, n, K, v0, v0, q->part.as<double2>());
    hipLaunchKernelGGL(k_breduce, dim3(K), dim3(256), 0, s, q->part.as<double2>(), nbe, K, st.sums);
    hipLaunchKernelGGL(k_bk_init, dim3(kb), dim3(64), 0, s, st, K, st.sums);
    ELAUNCH(k_bdiv_sqrt_norm, n, K, n, K, v0, st.norm);
    EDX_LAUNCH_CHECK();
    q->done = 0;
    q->curr = 0;
    q->prev = 1;
    q->live = true;
    return EDX_OK;
}

// up to nsteps more steps of every live recurrence (build_krylov_mp,
// src/lanczos.f90:266-298, K at a time)
int kb_advance(edx_krylov_batch *q, int nsteps) {
    if (!q->live || nsteps <= 0) return EDX_OK;
    if (q->done + nsteps > q->cap) {
        int nc = std::max(q->cap * 2, q->done + nsteps);
        EDX_TRY(kb_layout(q, nc, q->sc.p, q->flags.p, q->cap));
    }
    const int64_t n = q->n;
    const int K = q->K;
    const int kb = (K + 63) / 64;
    hipStream_t s = edx::stream();
    const BKrylov &st = q->st;
    double2 *v[2] = {q->vb.as<double2>(), q->vb.as<double2>() + (size_t)n * K};
    const int nbe = elem_grid_x(n);
    const int nbv = vec_blocks(n);
    // the same condition routes the product to the single-vector SpMV (bspmm)
    const bool single = K == 1 && q->h->m >= TUNE_MIN_ROWS;
    const int end = q->done + nsteps;
    int checked = q->done;
    for (int i = q->done + 1; i <= end; ++i) {
        const int curr = q->curr, prev = q->prev;
        if (q->precision == EDX_C64) ELAUNCH(k_il_to_c64, n, K, n, K, v[curr], q->x32.as<float2>());
        BEpi e{};   // y = (1/s_i) H u_i - (beta_i/s_{i-1}) u_{i-1}; partial <u_i|y>
        e.y = v[prev];
        e.z = v[prev];
        e.c0 = st.c0;
        e.c1 = st.c1;
        e.dvec = v[curr];
        e.partials = q->part.as<double2>();
        int nb = 0;
        EDX_TRY(edx::bspmm(q->h, q->precision, K, v[curr], q->x32.as<float2>(), e, &nb));
        hipLaunchKernelGGL(k_breduce, dim3(K), dim3(256), 0, s, q->part.as<double2>(), nb, K, st.sums);
        hipLaunchKernelGGL(k_bk_alpha, dim3(kb), dim3(64), 0, s, st, K, i, st.sums);
        // u_{i+1} = y - (alpha_i/s_i) u_i, kept unnormalized (s_{i+1} = beta_{i+1})
        if (single) {   // one vector of a large matrix: the plain vector kernels
            hipLaunchKernelGGL(k_axpy_norm, dim3(nbv), dim3(VB), 0, s, n, v[prev], v[curr],
                               st.acoef, q->part.as<double2>(), nullptr);
            hipLaunchKernelGGL(k_breduce, dim3(1), dim3(256), 0, s, q->part.as<double2>(), nbv, 1,
                               st.sums);
        } else {
            ELAUNCH(k_baxpy_norm, n, K, n, K, v[prev], v[curr], st.acoef, q->part.as<double2>());
            hipLaunchKernelGGL(k_breduce, dim3(K), dim3(256), 0, s, q->part.as<double2>(), nbe, K,
                               st.sums);
        }
        hipLaunchKernelGGL(k_bk_beta, dim3(kb), dim3(64), 0, s, st, K, i, st.sums);
        EDX_LAUNCH_CHECK();
        std::swap(q->curr, q->prev);
        q->done = i;
        if (i - checked >= 32 || i == end) {
            std::vector<int> r(K);
            EDX_HIP(edx::copy_sync(r.data(), st.run + (size_t)i * K, K * sizeof(int), hipMemcpyDeviceToHost));
            checked = i;
            bool any = false;
            for (int k = 0; k < K; ++k) any |= r[k] != 0;
            if (!any) {
                q->live = false;
                break;
            }
        }
    }
    return EDX_OK;
}

// Keep the recurrences keep[0..nkeep) (strictly increasing) and drop the rest:
// both vector blocks are gathered on the device into blocks of nkeep, the
// scalar histories of the steps taken are rearranged on the host.  A kept
// recurrence continues with the same vectors and scalars, so with the same
// product schedule its later steps equal those of the uncompacted batch.
int kb_keep(edx_krylov_batch *q, int nkeep, const int *keep) {
    const int K = q->K;
    for (int j = 0; j < nkeep; ++j)
        EDX_CHECK(keep[j] >= 0 && keep[j] < K && (j == 0 || keep[j] > keep[j - 1]), EDX_ERR_ARG,
                  "keep[%d] = %d: indices must increase strictly within [0, %d)", j, keep[j], K);
    if (nkeep == K) return EDX_OK;
    const int64_t n = q->n;
    const int done = q->done;
    Dbuf vb, x32, kd;
    EDX_TRY(vb.alloc((size_t)2 * n * nkeep * sizeof(double2), "batched krylov vectors"));
    if (q->precision == EDX_C64)
        EDX_TRY(x32.alloc((size_t)n * nkeep * sizeof(float2), "batched x32"));
    EDX_TRY(kd.alloc(nkeep * sizeof(int), "kept columns"));
    EDX_HIP(edx::copy_sync(kd.p, keep, nkeep * sizeof(int), hipMemcpyHostToDevice));
    const double2 *src = q->vb.as<double2>();
    double2 *dst = vb.as<double2>();
    for (int h = 0; h < 2; ++h)
        ELAUNCH(k_il_keep, n, nkeep, n, K, nkeep, kd.as<int>(), src + (size_t)h * n * K,
                dst + (size_t)h * n * nkeep);
    EDX_LAUNCH_CHECK();
    // the history of the steps taken: alpha_1..done, beta_0..done, run_0..done
    const BKrylov &st = q->st;
    std::vector<double> ha((size_t)done * K), hb((size_t)(done + 1) * K), ht((size_t)BK_TAIL * K);
    std::vector<int> hr((size_t)(done + 1) * K), hn(K);
    if (done > 0)
        EDX_HIP(edx::copy_sync(ha.data(), st.alpha, ha.size() * sizeof(double), hipMemcpyDeviceToHost));
    EDX_HIP(edx::copy_sync(hb.data(), st.beta, hb.size() * sizeof(double), hipMemcpyDeviceToHost));
    EDX_HIP(edx::copy_sync(ht.data(), st.norm, ht.size() * sizeof(double), hipMemcpyDeviceToHost));
    EDX_HIP(edx::copy_sync(hr.data(), st.run, hr.size() * sizeof(int), hipMemcpyDeviceToHost));
    EDX_HIP(edx::copy_sync(hn.data(), st.neff, K * sizeof(int), hipMemcpyDeviceToHost));
    auto pick = [&](const auto &in, int rows, int width) {
        std::vector<typename std::decay_t<decltype(in)>::value_type> out((size_t)rows * nkeep * width);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < nkeep; ++j)
                for (int t = 0; t < width; ++t)
                    out[((size_t)i * nkeep + j) * width + t] = in[((size_t)i * K + keep[j]) * width + t];
        return out;
    };
    const std::vector<double> na = pick(ha, done, 1), nb = pick(hb, done + 1, 1);
    const std::vector<int> nr = pick(hr, done + 1, 1), nn = pick(hn, 1, 1);
    // the tail holds per-RHS fields of 1, 2, 2, 2, 1 and 1 doubles in turn
    std::vector<double> nt((size_t)BK_TAIL * nkeep);
    size_t off_old = 0, off_new = 0;
    for (int w : {1, 2, 2, 2, 1, 1}) {
        for (int j = 0; j < nkeep; ++j)
            for (int t = 0; t < w; ++t)
                nt[off_new + (size_t)j * w + t] = ht[off_old + (size_t)keep[j] * w + t];
        off_old += (size_t)w * K;
        off_new += (size_t)w * nkeep;
    }
    q->K = nkeep;
    const int rc = kb_layout(q, q->cap, nullptr, nullptr, 0);
    if (rc) {
        q->K = K;
        return rc;
    }
    // the vector blocks go in first: a failed upload below leaves a batch whose
    // vectors and scalars both have nkeep columns, marked as not advancing
    std::swap(q->vb.p, vb.p);        // the old blocks are freed with the locals
    std::swap(q->x32.p, x32.p);
    const BKrylov &ns = q->st;
    const bool ok =
        (done == 0 || edx::copy_sync(ns.alpha, na.data(), na.size() * sizeof(double),
                                     hipMemcpyHostToDevice) == hipSuccess) &&
        edx::copy_sync(ns.beta, nb.data(), nb.size() * sizeof(double), hipMemcpyHostToDevice) ==
            hipSuccess &&
        edx::copy_sync(ns.norm, nt.data(), nt.size() * sizeof(double), hipMemcpyHostToDevice) ==
            hipSuccess &&
        edx::copy_sync(ns.run, nr.data(), nr.size() * sizeof(int), hipMemcpyHostToDevice) ==
            hipSuccess &&
        edx::copy_sync(ns.neff, nn.data(), nkeep * sizeof(int), hipMemcpyHostToDevice) ==
            hipSuccess;
    if (!ok) {
        q->live = false;
        return edx::set_error(EDX_ERR_HIP, "batch compaction: upload of the kept histories failed");
    }
    return EDX_OK;
}

// alpha / beta of the first count steps, nrhs x count row-major (beta_1..beta_count)
int kb_read(const edx_krylov_batch *q, int count, int *neff, double *alpha, double *beta,
            double *norm) {
    const int K = q->K;
    const int c = std::min(count, q->done);
    std::vector<double> ha((size_t)std::max(c, 1) * K), hb((size_t)(c + 1) * K), hn(K);
    std::vector<int> hneff(K);
    const BKrylov &st = q->st;
    if (c > 0) {
        EDX_HIP(edx::copy_sync(ha.data(), st.alpha, (size_t)c * K * sizeof(double), hipMemcpyDeviceToHost));
        EDX_HIP(edx::copy_sync(hb.data(), st.beta, (size_t)(c + 1) * K * sizeof(double),
                          hipMemcpyDeviceToHost));
    }
    EDX_HIP(edx::copy_sync(hn.data(), st.norm, K * sizeof(double), hipMemcpyDeviceToHost));
    EDX_HIP(edx::copy_sync(hneff.data(), st.neff, K * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < K; ++k) {
        if (neff) neff[k] = std::min(hneff[k], count);
        if (norm) norm[k] = hn[k];
        for (int i = 0; i < count; ++i) {
            const bool have = i < c;
            if (alpha) alpha[(size_t)k * count + i] = have ? ha[(size_t)i * K + k] : 0.0;
            if (beta) beta[(size_t)k * count + i] = have ? hb[(size_t)(i + 1) * K + k] : 0.0;
        }
    }
    return EDX_OK;
}

int bminres_run(const edx_csr *h, int precision, int K, const double2 *sigma, const double2 *Bcm,
                double2 *Xcm, double tol, int maxit, int *iters, double *final_res) {
    const int64_t n = h->n;
    hipStream_t s = edx::stream();
    Dbuf vb, x32, part, sc, flags;
    const size_t nk = (size_t)n * K;
    EDX_TRY(vb.alloc(6 * nk * sizeof(double2), "batched minres vectors"));
    if (precision == EDX_C64) EDX_TRY(x32.alloc(nk * sizeof(float2), "batched x32"));
    const int nbs = bspmm_blocks(n), nbe = elem_grid_x(n);
    EDX_TRY(part.alloc((size_t)std::max(nbs, nbe) * K * sizeof(double2) + 64, "partials"));
    const int nvec_scalars = 14;
    EDX_TRY(sc.alloc((size_t)K * (nvec_scalars * sizeof(double2) + sizeof(double)), "minres scalars"));
    EDX_TRY(flags.alloc(((size_t)(maxit + 2) * K + K) * sizeof(int), "minres flags"));
    double2 *S2 = sc.as<double2>();
    BMinres st;
    st.sigma0 = S2;
    st.sigma1 = S2 + K;
    st.gamma0 = S2 + 2 * K;
    st.gamma1 = S2 + 3 * K;
    st.eta = S2 + 4 * K;
    st.res = S2 + 5 * K;
    st.alpha = S2 + 6 * K;
    st.rho2 = S2 + 7 * K;
    st.rho3 = S2 + 8 * K;
    st.inv_rho1 = S2 + 9 * K;
    st.xcoef = S2 + 10 * K;
    double2 *sums = S2 + 11 * K;
    double2 *neg1 = S2 + 12 * K;     // -1 for every RHS (c0 of sigma - H)
    double2 *sig = S2 + 13 * K;      // shifts
    st.beta0 = (double *)(S2 + 14 * K);
    st.run = flags.as<int>();
    st.iters = st.run + (size_t)(maxit + 2) * K;
    st.tol = tol;
    st.maxit = maxit;
    double2 *v[2] = {vb.as<double2>(), vb.as<double2>() + nk};
    double2 *w[2] = {vb.as<double2>() + 2 * nk, vb.as<double2>() + 3 * nk};
    double2 *tmp = vb.as<double2>() + 4 * nk, *x = vb.as<double2>() + 5 * nk, *b = tmp;
    EDX_HIP(hipMemsetAsync(vb.p, 0, 6 * nk * sizeof(double2), s));
    EDX_HIP(hipMemsetAsync(flags.p, 0, ((size_t)(maxit + 2) * K + K) * sizeof(int), s));
    {
        std::vector<double2> hneg(K, make_double2(-1.0, 0.0));
        EDX_HIP(edx::copy_sync(neg1, hneg.data(), K * sizeof(double2), hipMemcpyHostToDevice));
        EDX_HIP(edx::copy_sync(sig, sigma, K * sizeof(double2), hipMemcpyHostToDevice));
    }
    const int kb = (K + 63) / 64;
    ELAUNCH(k_cm_to_il, n, K, n, K, Bcm, v[0]);
    ELAUNCH(k_bdot, n, K, n, K, v[0], v[0], part.as<double2>());
    hipLaunchKernelGGL(k_breduce, dim3(K), dim3(256), 0, s, part.as<double2>(), nbe, K, sums);
    hipLaunchKernelGGL(k_bm_init, dim3(kb), dim3(64), 0, s, st, K, sums);
    EDX_LAUNCH_CHECK();
    st.istep = nullptr;
    // one iteration on edx::stream(); its step is i, or *sti.istep (then
    // advanced) when the launches are captured
    auto iterate = [&](int i, int cur, int prv, const BMinres &sti) -> int {
        hipStream_t ls = edx::stream();
        ELAUNCH(k_bm_divbeta, n, K, n, K, v[cur], sti, i);
        if (precision == EDX_C64) ELAUNCH(k_il_to_c64, n, K, n, K, v[cur], x32.as<float2>());
        BEpi e{};   // tmp = sigma_k v - H v ; partial <v|tmp>
        e.y = tmp;
        e.c0 = neg1;
        e.c2 = sig;
        e.dvec = v[cur];
        e.partials = part.as<double2>();
        int nb = 0;
        EDX_TRY(edx::bspmm(h, precision, K, v[cur], x32.as<float2>(), e, &nb));
        hipLaunchKernelGGL(k_breduce, dim3(K), dim3(256), 0, ls, part.as<double2>(), nb, K, sums);
        hipLaunchKernelGGL(k_bm_alpha, dim3(kb), dim3(64), 0, ls, sti, K, i, sums);
        ELAUNCH(k_bm_vupdate, n, K, n, K, tmp, v[cur], v[prv], sti, i, part.as<double2>());
        hipLaunchKernelGGL(k_breduce, dim3(K), dim3(256), 0, ls, part.as<double2>(), nbe, K, sums);
        hipLaunchKernelGGL(k_bm_scalars, dim3(kb), dim3(64), 0, ls, sti, K, i, sums);
        ELAUNCH(k_bm_wx, n, K, n, K, v[cur], w[prv], w[cur], x, sti, i);
        if (sti.istep) hipLaunchKernelGGL(k_bm_next, dim3(1), dim3(1), 0, ls, (int *)sti.istep);
        EDX_LAUNCH_CHECK();
        return EDX_OK;
    };
    // iteration 0 runs eagerly (it also times the SpMM schedules on first use);
    // then an odd/even pair of iterations is captured once into a HIP graph and
    // replayed, the kernels reading the step from device memory
    EDX_TRY(iterate(0, 0, 1, st));
    int i = 1, checked = 0;
    Dbuf stepb;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    if (graphs_enabled() && !edx::g_prof && maxit >= 3) {
        EDX_TRY(stepb.alloc(sizeof(int), "minres step"));
        EDX_HIP(edx::copy_sync(stepb.p, &i, sizeof(int), hipMemcpyHostToDevice));
        BMinres sg = st;
        sg.istep = stepb.as<int>();
        EDX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        const hipError_t eb = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int rc = eb == hipSuccess ? EDX_OK
                                  : edx::hip_fail(eb, "hipStreamBeginCapture", __FILE__, __LINE__);
        if (!rc) {
            {
                edx::StreamScope scope(cap);
                rc = iterate(0, 1, 0, sg);
                if (!rc) rc = iterate(0, 0, 1, sg);
            }
            hipGraph_t graph = nullptr;
            const hipError_t ee = hipStreamEndCapture(cap, &graph);
            if (!rc && ee != hipSuccess) rc = edx::hip_fail(ee, "hipStreamEndCapture", __FILE__, __LINE__);
            if (!rc) {
                const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (ei != hipSuccess) {
                    exec = nullptr;
                    rc = edx::hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
                }
            }
            if (graph) (void)hipGraphDestroy(graph);
        }
        if (rc) {
            (void)hipStreamDestroy(cap);
            return rc;
        }
    }
    int rc = EDX_OK;
    while (i < maxit && rc == EDX_OK) {
        if (exec && i + 2 <= maxit) {
            const hipError_t el = hipGraphLaunch(exec, s);
            if (el != hipSuccess) rc = edx::hip_fail(el, "hipGraphLaunch", __FILE__, __LINE__);
            i += 2;
        } else {
            rc = iterate(i, i & 1, (i & 1) ^ 1, st);
            i += 1;
        }
        if (!rc && i - checked >= 16) {
            std::vector<int> r(K);
            const hipError_t em = edx::copy_sync(r.data(), st.run + (size_t)i * K, K * sizeof(int),
                                            hipMemcpyDeviceToHost);
            if (em != hipSuccess) {
                rc = edx::hip_fail(em, "hipMemcpy", __FILE__, __LINE__);
                break;
            }
            checked = i;
            bool any = false;
            for (int k = 0; k < K; ++k) any |= r[k] != 0;
            if (!any) break;
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (cap) (void)hipStreamDestroy(cap);
    EDX_TRY(rc);
    // true residuals |(sigma_k - H) x_k - b_k|
    ELAUNCH(k_cm_to_il, n, K, n, K, Bcm, b);
    if (precision == EDX_C64) ELAUNCH(k_il_to_c64, n, K, n, K, x, x32.as<float2>());
    BEpi e{};
    e.y = v[0];
    e.c0 = neg1;
    e.c2 = sig;
    EDX_TRY(edx::bspmm(h, precision, K, x, x32.as<float2>(), e, nullptr));
    ELAUNCH(k_bsub, n, K, n, K, v[0], b);
    ELAUNCH(k_bdot, n, K, n, K, v[0], v[0], part.as<double2>());
    hipLaunchKernelGGL(k_breduce, dim3(K), dim3(256), 0, s, part.as<double2>(), nbe, K, sums);
    ELAUNCH(k_il_to_cm, n, K, n, K, x, Xcm);
    EDX_LAUNCH_CHECK();
    std::vector<double2> hs(K);
    EDX_HIP(edx::copy_sync(hs.data(), sums, K * sizeof(double2), hipMemcpyDeviceToHost));
    std::vector<int> hit(K);
    EDX_HIP(edx::copy_sync(hit.data(), st.iters, K * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < K; ++k) {
        if (iters) iters[k] = hit[k];
        if (final_res) final_res[k] = sqrt(fabs(hs[k].x));
    }
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_csr_prepare(edx_csr *a, int precision) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    return edx::ensure_precision(a, precision);
}

int edx_build_krylov_batch(const edx_csr *h, int precision, int nrhs, const void *d_V0, int maxn,
                           int *neff, double *alpha, double *beta, double *norm) {
    EDX_CHECK(h && d_V0 && neff && alpha && beta && norm && nrhs >= 1 && maxn >= 1, EDX_ERR_ARG,
              "bad argument");
    EDX_CHECK(nrhs <= 4096, EDX_ERR_ARG, "at most 4096 right-hand sides per call");
    EDX_CHECK(h->m == h->n, EDX_ERR_ARG, "matrix must be square");
    edx_krylov_batch *q = nullptr;
    EDX_TRY(edx_krylov_batch_start(h, precision, nrhs, d_V0, maxn, &q));
    int rc = kb_advance(q, maxn);
    if (!rc) rc = kb_read(q, maxn, neff, alpha, beta, norm);
    edx_krylov_batch_free(q);
    if (rc) return rc;
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    return EDX_OK;
}

int edx_krylov_batch_start(const edx_csr *h, int precision, int nrhs, const void *d_V0,
                           int capacity, edx_krylov_batch **out) {
    EDX_CHECK(out, EDX_ERR_ARG, "null output handle");
    *out = nullptr;
    EDX_CHECK(h && d_V0 && nrhs >= 1, EDX_ERR_ARG, "bad argument");
    EDX_CHECK(nrhs <= 4096, EDX_ERR_ARG, "at most 4096 right-hand sides per call");
    EDX_CHECK(h->m == h->n, EDX_ERR_ARG, "matrix must be square");
    EDX_TRY(edx::ensure_precision(const_cast<edx_csr *>(h), precision));
    edx_krylov_batch *q = new edx_krylov_batch;
    q->h = h;
    q->precision = precision;
    q->K = nrhs;
    q->n = h->n;
    int rc = kb_start(q, (const double2 *)d_V0, std::max(capacity, 16));
    if (rc) {
        delete q;
        return rc;
    }
    *out = q;
    return EDX_OK;
}

int edx_krylov_batch_advance(edx_krylov_batch *q, int nsteps, int *steps_done) {
    EDX_CHECK(q && nsteps >= 0, EDX_ERR_ARG, "bad argument");
    EDX_TRY(kb_advance(q, nsteps));
    if (steps_done) *steps_done = q->done;
    return EDX_OK;
}

int edx_krylov_batch_keep(edx_krylov_batch *q, int nkeep, const int *keep) {
    EDX_CHECK(q && keep && nkeep >= 1 && nkeep <= q->K, EDX_ERR_ARG, "bad argument");
    return kb_keep(q, nkeep, keep);
}

int edx_krylov_batch_read(const edx_krylov_batch *q, int count, int *neff, double *alpha,
                          double *beta, double *norm) {
    EDX_CHECK(q && count >= 0, EDX_ERR_ARG, "bad argument");
    return kb_read(q, count, neff, alpha, beta, norm);
}

void edx_krylov_batch_free(edx_krylov_batch *q) { delete q; }

int edx_minres_shifted_batch(const edx_csr *h, int precision, int nrhs, const double *sigma,
                             const void *d_B, void *d_X, double tol, int maxit, int *iters,
                             double *final_res) {
    EDX_CHECK(h && sigma && d_B && d_X && nrhs >= 1 && maxit >= 0, EDX_ERR_ARG, "bad argument");
    EDX_CHECK(nrhs <= 4096, EDX_ERR_ARG, "at most 4096 right-hand sides per call");
    EDX_CHECK(h->m == h->n, EDX_ERR_ARG, "matrix must be square");
    EDX_TRY(edx::ensure_precision(const_cast<edx_csr *>(h), precision));
    int rc = bminres_run(h, precision, nrhs, (const double2 *)sigma, (const double2 *)d_B,
                         (double2 *)d_X, tol, maxit, iters, final_res);
    if (rc) return rc;
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    return EDX_OK;
}

}  // extern "C"
