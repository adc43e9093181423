// CSR build of Hamiltonians and transition operators (part of edx_kernels.hip).
//
// Reference semantics (src/ham.f90:2-1115 with src/m_types.f90:59-121): for
// every row r, walk the terms in order (hoppings, then Coulombs, then omega on
// the diagonal when ham_sign < 0); for each term whose annihilated orbitals are
// occupied and created orbitals empty, the fermion sign is (-1)^(occupied
// orbitals below each operator, applied in sequence); the resulting word is
// looked up in the column basis; contributions are kept in ascending column
// order and equal columns are summed left to right in term order, exact zeros
// retained.  Values are val * sign * ham_sign.
//
// Device mapping: one 64-lane wavefront per row.  Lanes evaluate 64 terms at a
// time; survivors are compacted with a ballot into an LDS list that stays in
// term order; the list is bitonic-sorted on the key (column << 32 | term << 1 |
// sign), which orders by column and, within a column, by term; a head lane per
// column then sums its run sequentially, which is exactly the reference's
// accumulation order.  A count pass sizes the rows, an exclusive scan (int64)
// builds row pointers, a fill pass writes columns and values.

namespace {

enum { ROW_PLAIN = 0, ROW_COREHOLE = 1, ROW_FULLCORE = 2 };
enum { COL_PLAIN = 0, COL_COREHOLE = 1, COL_VALONLY = 2 };

struct BuildParams {
    int64_t M;             // rows
    const int64_t *left;   // row-space valence words
    int64_t nleft;
    BasisView right;       // column-space valence basis
    int nv, nc;
    const uint32_t *ops;   // packed 0-based operator positions, one byte each
    const double2 *tv;     // term values
    int nhop;              // terms [0, nhop) are hoppings, the rest Coulomb
    int T;
    int omega_on;
    double2 omega;
    int negate;            // ham_sign < 0
    int64_t *counts;       // count pass output (length M + 1)
    const int64_t *rowptr; // fill pass input
    int32_t *cols;
    double2 *vals;
    int *overflow;
    int inter;             // core-hole index valence * nc + slot (else slot * n + valence)
};

__device__ __forceinline__ bool op_ann(uint64_t &m, int p, int &par) {
    if (!((m >> p) & 1ull)) return false;
    par += __popcll(m & ((1ull << p) - 1ull));
    m ^= (1ull << p);
    return true;
}

__device__ __forceinline__ bool op_cre(uint64_t &m, int p, int &par) {
    if ((m >> p) & 1ull) return false;
    par += __popcll(m & ((1ull << p) - 1ull));
    m |= (1ull << p);
    return true;
}

template <int RK>
__device__ __forceinline__ uint64_t row_word(const BuildParams &p, int64_t r) {
    if (RK == ROW_PLAIN) return (uint64_t)p.left[r];
    const uint64_t full = (1ull << p.nc) - 1ull;
    if (RK == ROW_COREHOLE) {
        const int64_t k = p.inter ? r % p.nc : r / p.nleft;
        const int64_t v = p.inter ? r / p.nc : r - k * p.nleft;
        uint64_t core = full ^ (1ull << (p.nc - 1 - k));   // src/ham.f90:419-421
        return (core << p.nv) | (uint64_t)p.left[v];
    }
    return (full << p.nv) | (uint64_t)p.left[r];
}

template <int CK>
__device__ __forceinline__ int64_t col_of(const BuildParams &p, uint64_t w) {
    if (CK == COL_PLAIN) return basis_index(p.right, w);
    const uint64_t val = w & ((1ull << p.nv) - 1ull);
    if (CK == COL_VALONLY) return basis_index(p.right, val);   // src/ham.f90:806-808
    const uint64_t core = w >> p.nv;
    const uint64_t full = (1ull << p.nc) - 1ull;
    if (core > full) return -1;
    const uint64_t holes = full ^ core;
    if (__popcll(holes) != 1) return -1;
    int b = __ffsll((unsigned long long)holes) - 1;
    int64_t k = p.nc - 1 - b;
    int64_t l = basis_index(p.right, val);
    if (l < 0) return -1;
    return p.inter ? l * p.nc + k : k * p.right.n + l;
}

__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ double2 key_value(const BuildParams &p, uint64_t key) {
    unsigned t = (unsigned)((key >> 1) & 0x7fffffffu);
    if ((int)t == p.T) return p.omega;   // omega goes in unsigned
    double2 v = p.tv[t];
    if (key & 1ull) v = make_double2(-v.x, -v.y);
    return v;
}

// Bitonic sort of keys[0, n) (n a power of two) by one 64-lane workgroup.
__device__ __forceinline__ void wave_sort(uint64_t *keys, int n) {
    const int lane = threadIdx.x;
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = lane; q < (n >> 1); q += 64) {
                int i = (q / j) * 2 * j + (q % j);
                int ixj = i + j;
                uint64_t a = keys[i], b = keys[ixj];
                bool up = (i & k) == 0;
                if ((a > b) == up) {
                    keys[i] = b;
                    keys[ixj] = a;
                }
            }
            __syncthreads();
        }
    }
}

template <int RK, int CK, int CAP, bool FILL>
__global__ __launch_bounds__(64) void k_build(BuildParams p) {
    __shared__ uint64_t keys[CAP];
    const int lane = threadIdx.x;
    for (int64_t r = blockIdx.x; r < p.M; r += gridDim.x) {
        const uint64_t w = row_word<RK>(p, r);
        int ncand = 0;
        for (int t0 = 0; t0 < p.T; t0 += 64) {
            const int t = t0 + lane;
            bool ok = false;
            uint64_t key = 0;
            if (t < p.T) {
                const uint32_t op = p.ops[t];
                uint64_t m = w;
                int par = 0;
                bool s;
                if (t < p.nhop)
                    s = op_ann(m, op & 63, par) && op_cre(m, (op >> 8) & 63, par);
                else
                    s = op_ann(m, op & 63, par) && op_ann(m, (op >> 8) & 63, par) &&
                        op_cre(m, (op >> 16) & 63, par) && op_cre(m, (op >> 24) & 63, par);
                if (s) {
                    int64_t c = col_of<CK>(p, m);
                    if (c >= 0) {
                        ok = true;
                        key = ((uint64_t)c << 32) | ((uint64_t)t << 1) |
                              (uint64_t)((par ^ p.negate) & 1);
                    }
                }
            }
            const uint64_t mask = __ballot(ok);
            const int pos = ncand + lanes_below(mask);
            if (ok && pos < CAP) keys[pos] = key;
            ncand += __popcll(mask);
        }
        if (p.omega_on) {
            if (lane == 0 && ncand < CAP) keys[ncand] = ((uint64_t)r << 32) | ((uint64_t)p.T << 1);
            ++ncand;
        }
        if (ncand > CAP) {
            if (lane == 0) {
                atomicOr(p.overflow, 1);
                if (!FILL) p.counts[r] = 0;
            }
            __syncthreads();
            continue;
        }
        int n2 = 2;
        while (n2 < ncand) n2 <<= 1;
        for (int i = ncand + lane; i < n2; i += 64) keys[i] = ~0ull;
        __syncthreads();
        if (ncand > 1) wave_sort(keys, n2);
        int64_t uniq = 0;
        for (int base = 0; base < ncand; base += 64) {
            const int q = base + lane;
            bool head = false;
            uint32_t col = 0;
            if (q < ncand) {
                col = (uint32_t)(keys[q] >> 32);
                head = (q == 0) || ((uint32_t)(keys[q - 1] >> 32) != col);
            }
            const uint64_t hm = __ballot(head);
            if (FILL && head) {
                const int64_t out = p.rowptr[r] + uniq + lanes_below(hm);
                double2 acc = key_value(p, keys[q]);
                for (int s = q + 1; s < ncand && (uint32_t)(keys[s] >> 32) == col; ++s) {
                    double2 v = key_value(p, keys[s]);
                    acc.x = acc.x + v.x;
                    acc.y = acc.y + v.y;
                }
                p.cols[out] = (int32_t)col;
                p.vals[out] = acc;
            }
            uniq += __popcll(hm);
        }
        if (!FILL && lane == 0) p.counts[r] = uniq;
        __syncthreads();
    }
}

// ------------------------------------------------------ int64 exclusive scan

constexpr int SCAN_T = 256, SCAN_V = 8, SCAN_TILE = SCAN_T * SCAN_V;

__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(int64_t *a, int64_t n, int64_t *sums) {
    __shared__ int64_t part[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_V;
    int64_t v[SCAN_V];
    int64_t s = 0;
    for (int i = 0; i < SCAN_V; ++i) {
        v[i] = (base + i < n) ? a[base + i] : 0;
        s += v[i];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {   // inclusive Hillis-Steele
        int64_t t = (threadIdx.x >= off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t run = part[threadIdx.x] - s;
    for (int i = 0; i < SCAN_V; ++i) {
        if (base + i < n) a[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == SCAN_T - 1) sums[blockIdx.x] = part[SCAN_T - 1];
}

__global__ __launch_bounds__(SCAN_T) void k_scan_add(int64_t *a, int64_t n, const int64_t *offs) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    const int64_t add = offs[blockIdx.x];
    for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_T)
        if (base + i < n) a[base + i] += add;
}

int exclusive_scan(int64_t *d_a, int64_t n) {
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int64_t *sums = (int64_t *)edx::dmalloc(tiles * sizeof(int64_t), "scan sums");
    EDX_CHECK(sums, EDX_ERR_NOMEM, "%s", edx_last_error());
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_T), 0, edx::stream(), d_a,
                       n, sums);
    EDX_LAUNCH_CHECK();
    if (tiles > 1) {
        int rc = exclusive_scan(sums, tiles);
        if (rc) {
            edx::dfree(sums);
            return rc;
        }
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_T), 0, edx::stream(), d_a,
                           n, sums);
        EDX_LAUNCH_CHECK();
    }
    edx::dfree(sums);
    return EDX_OK;
}

// Hermitian-consistent complex64 rounding of a square CSR: an off-diagonal
// entry whose transposed partner equals its conjugate up to FP64 rounding is
// stored as round((a_rc + conj(a_cr)) / 2).  The partner rounds the conjugate of
// the same FP64 sum, so the stored matrix keeps a_rc == conj(a_cr) exactly (the
// FP64 sums of the two entries come from different term orders and can differ
// in the last bit, which independent rounding would turn into a 1-ulp float
// asymmetry and a non-Hermitian operator).  A diagonal entry whose imaginary
// part is rounding residue (<= 1e-13 of the real part) is stored real.  Other
// entries round independently.
__device__ __forceinline__ float2 herm_round(const int64_t *rowptr, const int32_t *cols,
                                             const double2 *vals, int64_t r, int64_t j) {
    const double2 a = vals[j];
    const int64_t c = cols[j];
    if (c == r) {
        const float re = __double2float_rn(a.x);
        return make_float2(re, fabs(a.y) <= 1e-13 * fabs(a.x) ? 0.0f : __double2float_rn(a.y));
    } else {
        int64_t lo = rowptr[c], hi = rowptr[c + 1] - 1;
        while (lo <= hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            const int64_t cm = cols[mid];
            if (cm < r) lo = mid + 1;
            else if (cm > r) hi = mid - 1;
            else {
                const double2 b = vals[mid];
                const double dx = a.x - b.x, dy = a.y + b.y;
                const double sc = fabs(a.x) + fabs(a.y) + fabs(b.x) + fabs(b.y);
                if (fabs(dx) + fabs(dy) <= 1e-13 * sc) {
                    const double2 s = make_double2((a.x + b.x) * 0.5, (a.y - b.y) * 0.5);
                    return make_float2(__double2float_rn(s.x), __double2float_rn(s.y));
                }
                break;
            }
        }
    }
    return make_float2(__double2float_rn(a.x), __double2float_rn(a.y));
}

__global__ void k_to_c64_herm(int64_t m, const int64_t *rowptr, const int32_t *cols,
                              const double2 *vals, float2 *out) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x)
        for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
            out[j] = herm_round(rowptr, cols, vals, r, j);
}

__global__ void k_to_c64(int64_t n, const double2 *in, float2 *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = make_float2(__double2float_rn(in[i].x), __double2float_rn(in[i].y));
}

// ------------------------------------------------------------ host driver

struct TermTable {
    uint32_t *d_ops = nullptr;
    double2 *d_vals = nullptr;
    int nhop = 0, T = 0;
    ~TermTable() {
        edx::dfree(d_ops);
        edx::dfree(d_vals);
    }
};

int upload_terms(const edx_tensor2 *hop, int64_t nhop, const edx_tensor4 *coul, int64_t ncoul,
                 TermTable &tt) {
    EDX_CHECK(nhop >= 0 && ncoul >= 0 && nhop + ncoul < (1 << 29), EDX_ERR_ARG,
              "term counts out of range");
    EDX_CHECK(nhop == 0 || hop, EDX_ERR_ARG, "null hopping terms");
    EDX_CHECK(ncoul == 0 || coul, EDX_ERR_ARG, "null Coulomb terms");
    const int64_t T = nhop + ncoul;
    std::vector<uint32_t> ops(T ? T : 1);
    std::vector<double2> vals(T ? T : 1);
    auto pos = [](int32_t i, int64_t t) -> int {
        return (i >= 1 && i <= 63) ? (i - 1) : -1;
    };
    for (int64_t t = 0; t < nhop; ++t) {
        int a = pos(hop[t].ind1, t), b = pos(hop[t].ind2, t);
        EDX_CHECK(a >= 0 && b >= 0, EDX_ERR_ARG,
                  "hopping term %lld has orbital index outside 1..63", (long long)t + 1);
        ops[t] = (uint32_t)a | ((uint32_t)b << 8);
        vals[t] = make_double2(hop[t].re, hop[t].im);
    }
    for (int64_t t = 0; t < ncoul; ++t) {
        const edx_tensor4 &c = coul[t];
        int a = pos(c.ind1, t), b = pos(c.ind2, t), d = pos(c.ind3, t), e = pos(c.ind4, t);
        EDX_CHECK(a >= 0 && b >= 0 && d >= 0 && e >= 0, EDX_ERR_ARG,
                  "Coulomb term %lld has orbital index outside 1..63", (long long)t + 1);
        ops[nhop + t] = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)d << 16) | ((uint32_t)e << 24);
        vals[nhop + t] = make_double2(c.re, c.im);
    }
    tt.nhop = (int)nhop;
    tt.T = (int)T;
    tt.d_ops = (uint32_t *)edx::dmalloc(ops.size() * sizeof(uint32_t), "term ops");
    tt.d_vals = (double2 *)edx::dmalloc(vals.size() * sizeof(double2), "term values");
    EDX_CHECK(tt.d_ops && tt.d_vals, EDX_ERR_NOMEM, "%s", edx_last_error());
    EDX_HIP(edx::copy_sync(tt.d_ops, ops.data(), ops.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    EDX_HIP(edx::copy_sync(tt.d_vals, vals.data(), vals.size() * sizeof(double2),
                      hipMemcpyHostToDevice));
    return EDX_OK;
}

template <int RK, int CK, int CAP, bool FILL>
int launch_build(const BuildParams &p) {
    int64_t g = (int64_t)edx::num_cus() * 16;
    if (g > p.M) g = p.M;
    if (g < 1) g = 1;
    hipLaunchKernelGGL((k_build<RK, CK, CAP, FILL>), dim3((unsigned)g), dim3(64), 0,
                       edx::stream(), p);
    EDX_LAUNCH_CHECK();
    return EDX_OK;
}

template <int RK, int CK, bool FILL>
int launch_cap(const BuildParams &p, int cap) {
    switch (cap) {
        case 256: return launch_build<RK, CK, 256, FILL>(p);
        case 1024: return launch_build<RK, CK, 1024, FILL>(p);
        case 4096: return launch_build<RK, CK, 4096, FILL>(p);
        default: return edx::set_error(EDX_ERR_ARG, "bad capacity %d", cap);
    }
}

template <int RK, int CK>
int build_csr(BuildParams p, int64_t ncols, int precision, edx_csr **out) {
    EDX_CHECK(out, EDX_ERR_ARG, "null output");
    EDX_CHECK(ncols < ((int64_t)1 << 31) && p.M < ((int64_t)1 << 31), EDX_ERR_ARG,
              "dimension %lld x %lld exceeds the int32 column index range", (long long)p.M,
              (long long)ncols);
    EDX_CHECK(precision == EDX_C128 || precision == EDX_C64, EDX_ERR_ARG, "bad precision %d",
              precision);
    EDX_TRY(edx::ensure_binomials());
    int64_t *rowptr = (int64_t *)edx::dmalloc((p.M + 1) * sizeof(int64_t), "row pointers");
    int *overflow = (int *)edx::dmalloc(sizeof(int), "overflow flag");
    if (!rowptr || !overflow) {
        edx::dfree(rowptr);
        edx::dfree(overflow);
        return EDX_ERR_NOMEM;
    }
    p.counts = rowptr;
    p.overflow = overflow;
    // the omega diagonal counts as one more term
    int maxc = p.T + (p.omega_on ? 1 : 0);
    int cap = maxc <= 256 ? 256 : (maxc <= 1024 ? 1024 : 4096);
    int rc = EDX_OK;
    for (;;) {
        EDX_HIP(hipMemsetAsync(rowptr + p.M, 0, sizeof(int64_t), edx::stream()));
        EDX_HIP(hipMemsetAsync(overflow, 0, sizeof(int), edx::stream()));
        if ((rc = launch_cap<RK, CK, false>(p, cap))) break;
        int h = 0;
        EDX_HIP(edx::copy_sync(&h, overflow, sizeof(int), hipMemcpyDeviceToHost));
        if (!h) break;
        if (cap == 4096) {
            rc = edx::set_error(EDX_ERR_CAPACITY,
                                "a row produced more than 4096 surviving terms; not supported");
            break;
        }
        cap = (cap == 256) ? 1024 : 4096;
    }
    if (!rc) rc = exclusive_scan(rowptr, p.M + 1);
    int64_t nnz = 0;
    if (!rc) {
        hipError_t e = edx::copy_sync(&nnz, rowptr + p.M, sizeof(int64_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = edx::hip_fail(e, "edx::copy_sync(nnz)", __FILE__, __LINE__);
    }
    int32_t *cols = nullptr;
    double2 *vals = nullptr;
    if (!rc) {
        cols = (int32_t *)edx::dmalloc((nnz ? nnz : 1) * sizeof(int32_t), "CSR columns");
        vals = (double2 *)edx::dmalloc((nnz ? nnz : 1) * sizeof(double2), "CSR values");
        if (!cols || !vals) rc = EDX_ERR_NOMEM;
    }
    if (!rc) {
        p.rowptr = rowptr;
        p.cols = cols;
        p.vals = vals;
        rc = launch_cap<RK, CK, true>(p, cap);
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(edx::stream());
        if (e != hipSuccess) rc = edx::hip_fail(e, "build", __FILE__, __LINE__);
    }
    edx::dfree(overflow);
    if (rc) {
        edx::dfree(rowptr);
        edx::dfree(cols);
        edx::dfree(vals);
        return rc;
    }
    edx_csr *a = new edx_csr{};
    a->corehole_layout = p.inter;
    if (RK == ROW_COREHOLE && CK == COL_COREHOLE && p.inter) a->nc_slots = p.nc;
    a->m = p.M;
    a->n = ncols;
    a->nnz = nnz;
    a->d_rowptr = rowptr;
    a->d_cols = cols;
    a->d_vals = vals;
    if (precision == EDX_C64) {
        rc = edx_csr_add_c64(a);
        if (rc) {
            edx_csr_free(a);
            return rc;
        }
    }
    *out = a;
    return EDX_OK;
}

BuildParams base_params(const TermTable &tt) {
    BuildParams p{};
    p.ops = tt.d_ops;
    p.tv = tt.d_vals;
    p.nhop = tt.nhop;
    p.T = tt.T;
    return p;
}

int check_ham_sign(double ham_sign) {
    EDX_CHECK(ham_sign == 1.0 || ham_sign == -1.0, EDX_ERR_ARG,
              "ham_sign must be +1 or -1 (got %g)", ham_sign);
    return EDX_OK;
}

int check_core(int nv, int nc) {
    EDX_CHECK(nv >= 1 && nc >= 1 && nv + nc <= 62, EDX_ERR_ARG,
              "need num_val_orbs >= 1, num_core_orbs >= 1, sum <= 62 (got %d, %d)", nv, nc);
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_build_ham_i(const edx_basis *fock, const edx_tensor2 *hop, int64_t nhopp,
                    const edx_tensor4 *coul, int64_t ncoul, double omega_re, double omega_im,
                    double ham_sign, int precision, edx_csr **out) {
    EDX_CHECK(fock, EDX_ERR_ARG, "null basis");
    EDX_TRY(check_ham_sign(ham_sign));
    TermTable tt;
    EDX_TRY(upload_terms(hop, nhopp, coul, ncoul, tt));
    BuildParams p = base_params(tt);
    p.M = fock->n;
    p.left = fock->d_words;
    p.nleft = fock->n;
    p.right = view_of(fock);
    p.omega_on = ham_sign < 0;
    p.omega = make_double2(omega_re, omega_im);
    p.negate = ham_sign < 0;
    return build_csr<ROW_PLAIN, COL_PLAIN>(p, fock->n, precision, out);
}

int edx_build_ham_n(const edx_basis *fock_n, int num_val_orbs, int num_core_orbs,
                    const edx_tensor2 *hop, int64_t nhopp, const edx_tensor4 *coul,
                    int64_t ncoul, double omega_re, double omega_im, double ham_sign,
                    int precision, edx_csr **out) {
    EDX_CHECK(fock_n, EDX_ERR_ARG, "null basis");
    EDX_TRY(check_ham_sign(ham_sign));
    EDX_TRY(check_core(num_val_orbs, num_core_orbs));
    TermTable tt;
    EDX_TRY(upload_terms(hop, nhopp, coul, ncoul, tt));
    BuildParams p = base_params(tt);
    p.M = fock_n->n * num_core_orbs;
    p.left = fock_n->d_words;
    p.nleft = fock_n->n;
    p.right = view_of(fock_n);
    p.nv = num_val_orbs;
    p.nc = num_core_orbs;
    p.inter = edx::corehole_layout() == EDX_COREHOLE_INTERLEAVED;
    p.omega_on = ham_sign < 0;
    p.omega = make_double2(omega_re, omega_im);
    p.negate = ham_sign < 0;
    return build_csr<ROW_COREHOLE, COL_COREHOLE>(p, p.M, precision, out);
}

int edx_build_transop_i(const edx_basis *fock_n, const edx_basis *fock_i, int num_val_orbs,
                        int num_core_orbs, const edx_tensor2 *trans, int64_t ntran,
                        int precision, edx_csr **out) {
    EDX_CHECK(fock_n && fock_i, EDX_ERR_ARG, "null basis");
    EDX_TRY(check_core(num_val_orbs, num_core_orbs));
    TermTable tt;
    EDX_TRY(upload_terms(trans, ntran, nullptr, 0, tt));
    BuildParams p = base_params(tt);
    p.M = fock_n->n * num_core_orbs;
    p.left = fock_n->d_words;
    p.nleft = fock_n->n;
    p.right = view_of(fock_i);
    p.nv = num_val_orbs;
    p.nc = num_core_orbs;
    p.inter = edx::corehole_layout() == EDX_COREHOLE_INTERLEAVED;
    return build_csr<ROW_COREHOLE, COL_VALONLY>(p, fock_i->n, precision, out);
}

int edx_build_transop_f(const edx_basis *fock_f, const edx_basis *fock_n, int num_val_orbs,
                        int num_core_orbs, const edx_tensor2 *trans, int64_t ntran,
                        int precision, edx_csr **out) {
    EDX_CHECK(fock_f && fock_n, EDX_ERR_ARG, "null basis");
    EDX_TRY(check_core(num_val_orbs, num_core_orbs));
    TermTable tt;
    EDX_TRY(upload_terms(trans, ntran, nullptr, 0, tt));
    BuildParams p = base_params(tt);
    p.M = fock_f->n;
    p.left = fock_f->d_words;
    p.nleft = fock_f->n;
    p.right = view_of(fock_n);
    p.nv = num_val_orbs;
    p.nc = num_core_orbs;
    p.inter = edx::corehole_layout() == EDX_COREHOLE_INTERLEAVED;
    return build_csr<ROW_FULLCORE, COL_COREHOLE>(p, fock_n->n * num_core_orbs, precision, out);
}

int edx_csr_info(const edx_csr *a, int64_t *m, int64_t *n, int64_t *nnz, int *precision) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    if (m) *m = a->m;
    if (n) *n = a->n;
    if (nnz) *nnz = a->nnz;
    if (precision) *precision = a->d_vals32 ? EDX_C64 : EDX_C128;
    return EDX_OK;
}

int edx_csr_download(const edx_csr *a, int64_t *rowptr, int32_t *cols, void *vals) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    if (rowptr)
        EDX_HIP(edx::copy_sync(rowptr, a->d_rowptr, (a->m + 1) * sizeof(int64_t),
                          hipMemcpyDeviceToHost));
    if (cols && a->nnz)
        EDX_HIP(edx::copy_sync(cols, a->d_cols, a->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (vals && a->nnz)
        EDX_HIP(edx::copy_sync(vals, a->d_vals, a->nnz * sizeof(double2), hipMemcpyDeviceToHost));
    return EDX_OK;
}

int edx_csr_download_values(const edx_csr *a, int precision, void *vals) {
    EDX_CHECK(a && vals, EDX_ERR_ARG, "null argument");
    if (!a->nnz) return EDX_OK;
    if (precision == EDX_C128) {
        EDX_HIP(edx::copy_sync(vals, a->d_vals, a->nnz * sizeof(double2), hipMemcpyDeviceToHost));
    } else {
        EDX_CHECK(precision == EDX_C64 && a->d_vals32, EDX_ERR_ARG,
                  "precision %d values not present", precision);
        EDX_HIP(edx::copy_sync(vals, a->d_vals32, a->nnz * sizeof(float2), hipMemcpyDeviceToHost));
    }
    return EDX_OK;
}

int edx_csr_add_c64(edx_csr *a) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    if (a->d_vals32) return EDX_OK;
    float2 *v = (float2 *)edx::dmalloc((a->nnz ? a->nnz : 1) * sizeof(float2), "CSR c64 values");
    EDX_CHECK(v, EDX_ERR_NOMEM, "%s", edx_last_error());
    if (a->nnz) {
        int64_t g = (a->nnz + 255) / 256;
        if (g > (int64_t)edx::num_cus() * 32) g = (int64_t)edx::num_cus() * 32;
        if (a->m == a->n) {
            int64_t gr = (a->m + 255) / 256;
            if (gr > (int64_t)edx::num_cus() * 32) gr = (int64_t)edx::num_cus() * 32;
            hipLaunchKernelGGL(k_to_c64_herm, dim3((unsigned)gr), dim3(256), 0, edx::stream(),
                               a->m, a->d_rowptr, a->d_cols, a->d_vals, v);
        } else {
            hipLaunchKernelGGL(k_to_c64, dim3((unsigned)g), dim3(256), 0, edx::stream(), a->nnz,
                               a->d_vals, v);
        }
        EDX_LAUNCH_CHECK();
        EDX_HIP(hipStreamSynchronize(edx::stream()));
    }
    a->d_vals32 = v;
    return EDX_OK;
}

void edx_csr_free(edx_csr *a) {
    if (!a) return;
    edx::dfree(a->d_rowptr);
    edx::dfree(a->d_cols);
    edx::dfree(a->d_vals);
    edx::dfree(a->d_vals32);
    edx::dfree(a->d_diag);
    edx::dfree(a->d_off_rowptr);
    edx::dfree(a->d_off_cols);
    edx::dfree(a->d_off_vals32);
    edx::dfree(a->d_off_idx);
    edx::dfree(a->d_vtab);
    edx::dfree(a->d_off_pk);
    edx::dfree(a->d_fv_rowptr);
    edx::dfree(a->d_fv_pk);
    edx::dfree(a->d_fr_rowptr);
    edx::dfree(a->d_fr_pk);
    edx::dfree(a->d_fb_rowptr);
    edx::dfree(a->d_fb_pk);
    edx::dfree(a->d_ftyp);
    edx::dfree(a->d_apk);
    edx::dfree(a->d_atab);
    delete a;
}

}  // extern "C"
