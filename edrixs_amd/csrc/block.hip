// Block Lanczos eigensolver for the lowest states (part of edx_kernels.hip).
//
//   edx_block_lanczos_eigs  <- diag_ham_lanczos, src/lanczos.f90:6-167 (the
//                              lowest nev eigenpairs of H_i for ed_fsolver,
//                              src/ed_driver.f90:216-239), for several states
//                              in one Krylov space
//
// A block of BW = 8 vectors advances per step: W = H Q_j is one 8-RHS SpMM
// over the CSR, W is reorthogonalized against the whole basis V = [Q_0 .. Q_j]
// by block classical Gram-Schmidt run twice, the block itself is
// orthonormalized by shifted Cholesky QR (Fukaya, Kannan, Nakatsukasa,
// Yamamoto and Yanagisawa, SIAM J. Sci. Comput. 42 (2020) A477, "shifted
// CholeskyQR for computing the QR factorization of ill-conditioned matrices"),
// and the Ritz pairs come from the Hermitian block-tridiagonal projection
//     T = V^H H V,   T_jj = Q_j^H H Q_j,   T_{j+1,j} = B_{j+1} = Q_{j+1}^H W.
// A Ritz pair (theta, V s) has residual |B_{j+1} s_last| (s_last: the last
// block of s), the block analogue of the reference's |beta_m z_m| test.
// Degenerate levels up to eight-fold are found without the restarts the
// single-vector solver needs.
//
// The reorthogonalization is the tall-skinny GEMM pair
//     C = V^H W   (k blocks of 8 x 8, a reduction over all n rows)
//     W = W - V C (n x 8)
// written for the FP64 matrix cores.  An 8-column complex block is the real
// n x 16 matrix [Re Q | Im Q]; with the interleaved layout (row r holds the 8
// complex values, 128 bytes) one v_mfma_f64_16x16x4f64 takes four rows of it:
//   * projection: lane l supplies real column l % 16 of row r + l / 16 for both
//     operands, so D = sum over rows of [Re Q|Im Q]^T [Re W|Im W], the 16 x 16
//     real Gram tile whose four 8 x 8 quarters give Re and Im of Q^H W;
//   * update: lane l supplies row r + l % 16 (two 16-byte loads cover its four
//     K-steps) against the real 16 x 16 form [[Re C, Im C], [-Im C, Re C]] of
//     the coefficients, so D is the 16-row tile of [Re VC | Im VC].
// Both read the basis once per pass in full 128-byte rows (HBM-bound: about
// 2 flops per byte); the projection keeps eight basis blocks' tiles in
// registers per wave and reduces over rows in a fixed order (waves, then
// workgroups), so results do not depend on timing.

namespace edx {
int herm_eig_lowest(int n, std::complex<double> *A, double *w, int nvec, std::complex<double> *Z);
}

namespace {

using cd = std::complex<double>;
typedef double bl_d4 __attribute__((ext_vector_type(4)));

constexpr int BW = 8;          // block width (complex columns)
constexpr int BL_GROUP = 8;    // basis blocks per projection workgroup
constexpr int BL_ROWS = 32;    // rows per wave pass of the projection (8 MFMA steps)
constexpr int BL_UTILES = 4;   // 16-row tiles per wave of the update

// Offset, in doubles within an interleaved row, of real column p of [Re | Im]:
// p < 8 is Re of complex column p, p >= 8 is Im of complex column p - 8.
__device__ __forceinline__ int bl_off(int p) { return p < 8 ? 2 * p : 2 * p - 15; }

// Partial Gram tiles of basis blocks blockIdx.y * BL_GROUP + b against X over
// the rows [blockIdx.x * rows_per_wg, +rows_per_wg): partials[chunk][block][256]
// holds D[i][j] = sum_r A'[r][i] X'[r][j] at i * 16 + j, A' = [Re Q | Im Q].
__global__ __launch_bounds__(256) void k_bproj(int64_t n, const double *V, int nblk,
                                               const double *X, int64_t rows_per_wg,
                                               double *__restrict__ partials) {
    __shared__ double red[BL_GROUP * 256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = lane & 15, q = lane >> 4;
    const int off = bl_off(p);
    const int b0 = blockIdx.y * BL_GROUP;
    const int nb = min(BL_GROUP, nblk - b0);
    const int64_t r_begin = (int64_t)blockIdx.x * rows_per_wg;
    const int64_t r_end = min(n, r_begin + rows_per_wg);
    bl_d4 acc[BL_GROUP];
#pragma unroll
    for (int b = 0; b < BL_GROUP; ++b) acc[b] = bl_d4{0.0, 0.0, 0.0, 0.0};
    for (int64_t t = r_begin + wave * BL_ROWS; t < r_end; t += 4 * BL_ROWS) {
        double xv[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int64_t r = t + 4 * s + q;
            xv[s] = r < r_end ? X[r * 16 + off] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < BL_GROUP; ++b) {
            if (b < nb) {
                const double *Vb = V + (int64_t)(b0 + b) * n * 16;
                double av[8];
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const int64_t r = t + 4 * s + q;
                    av[s] = r < r_end ? Vb[r * 16 + off] : 0.0;
                }
#pragma unroll
                for (int s = 0; s < 8; ++s)
                    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], xv[s], acc[b], 0, 0, 0);
            }
        }
    }
    // waves 0..3 in order; lane l holds D[4 * k + l / 16][l % 16] in acc[.][k]
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int b = 0; b < BL_GROUP; ++b) {
                if (b < nb) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int o = b * 256 + (4 * k + q) * 16 + p;
                        red[o] = w == 0 ? acc[b][k] : red[o] + acc[b][k];
                    }
                }
            }
        }
        __syncthreads();
    }
    double *out = partials + ((int64_t)blockIdx.x * nblk + b0) * 256;
    for (int i = threadIdx.x; i < nb * 256; i += 256) out[i] = red[i];
}

constexpr int BL_RED = 32;   // partial tiles summed per workgroup in the first reduction stage

// out[g][b] = sum of in[c][b] over c in [g * BL_RED, (g + 1) * BL_RED) (256
// doubles per tile), in chunk order: all loads of a thread are issued first.
__global__ __launch_bounds__(256) void k_bproj_stage(const double *__restrict__ in, int nchunk,
                                                     int nblk, double *__restrict__ out) {
    const int b = blockIdx.x, g = blockIdx.y, t = threadIdx.x;
    const int c0 = g * BL_RED;
    double v[BL_RED];
#pragma unroll
    for (int i = 0; i < BL_RED; ++i)
        v[i] = c0 + i < nchunk ? in[((int64_t)(c0 + i) * nblk + b) * 256 + t] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BL_RED; ++i) s += v[i];
    out[((int64_t)g * nblk + b) * 256 + t] = s;
}

// C[b] (8 x 8, row-major) = Q_b^H X from the summed Gram tile of block b.
__global__ __launch_bounds__(256) void k_bproj_reduce(const double *__restrict__ partials,
                                                      int nchunk, int nblk, double2 *C) {
    __shared__ double g[256];
    const int b = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partials[((int64_t)c * nblk + b) * 256 + t];
    g[t] = s;
    __syncthreads();
    if (t < 64) {
        const int c = t >> 3, d = t & 7;
        C[b * 64 + t] = make_double2(g[c * 16 + d] + g[(c + 8) * 16 + d + 8],
                                     g[c * 16 + d + 8] - g[(c + 8) * 16 + d]);
    }
}

// Wout = Win + sgn * sum_b Q_b C_b over nblk basis blocks (Win null: 0).  Each
// wave owns 16 * BL_UTILES rows and reads and writes only those, so Wout may
// be Win or one of the basis blocks.
__global__ __launch_bounds__(256) void k_bupd(int64_t n, const double2 *V, int nblk,
                                              const double2 *__restrict__ C, const double *Win,
                                              double *Wout, double sgn) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, q = lane >> 4;
    const int p = lane & 15, d = p & 7;
    const bool im_out = p >= 8;
    const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * (16 * BL_UTILES);
    if (base >= n) return;   // whole wave
    bl_d4 acc[BL_UTILES];
#pragma unroll
    for (int t = 0; t < BL_UTILES; ++t) acc[t] = bl_d4{0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < nblk; ++b) {
        // K-step s covers real basis columns: s=0 Re col q, s=1 Re col q+4,
        // s=2 Im col q, s=3 Im col q+4 (q = lane / 16)
        const double2 *Cb = C + b * 64;
        double bf[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const double2 cc = Cb[((s & 1) ? q + 4 : q) * 8 + d];
            bf[s] = s < 2 ? (im_out ? cc.y : cc.x) : (im_out ? cc.x : -cc.y);
        }
        const double2 *Vb = V + (int64_t)b * n * 8;
#pragma unroll
        for (int t = 0; t < BL_UTILES; ++t) {
            const int64_t r = base + 16 * t + i;
            double2 z0 = make_double2(0.0, 0.0), z1 = z0;
            if (r < n) {
                z0 = Vb[r * 8 + q];
                z1 = Vb[r * 8 + q + 4];
            }
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(z0.x, bf[0], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(z1.x, bf[1], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(z0.y, bf[2], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(z1.y, bf[3], acc[t], 0, 0, 0);
        }
    }
    const int off = bl_off(p);
#pragma unroll
    for (int t = 0; t < BL_UTILES; ++t) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t row = base + 16 * t + 4 * k + q;   // D[4k + l/16][l % 16]
            if (row < n) {
                const int64_t o = row * 16 + off;
                Wout[o] = (Win ? Win[o] : 0.0) + sgn * acc[t][k];
            }
        }
    }
}

// uniform real random values in [-0.5, 0.5) in column col of an interleaved block
__global__ void k_bl_random_col(int64_t n, int col, uint64_t seed, double2 *Q) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h = splitmix64(seed * 0x100000001b3ull + (uint64_t)r);
        Q[r * BW + col] = make_double2((double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5, 0.0);
    }
}

// out[c * n + r] = Q[r * BW + c] for c < ncols (interleaved -> column-major)
__global__ void k_bl_to_cols(int64_t n, const double2 *Q, int ncols, double2 *out) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < ncols; ++c) out[(int64_t)c * n + r] = Q[r * BW + c];
}

// ---------------------------------------------------------------- host side

struct BlockLz {
    const edx_csr *h;
    int precision;
    int64_t n;
    double2 *V;      // basis blocks, each n x BW interleaved, contiguous
    double2 *W;      // n x BW
    float2 *x32;     // complex64 copy of the block being multiplied (C64 only)
    double *part;    // projection partials: part_len doubles, then the stage output
    size_t part_len;
    double2 *dC;     // coefficients on the device, [blocks][64]
    uint64_t seed;
    int nrand = 0;   // random columns drawn so far
    std::vector<cd> tmpc;
};

// row chunks of the projection: at least 1024 rows per workgroup, about 1024
// workgroups in all
int bl_chunks(int64_t n, int nblk) {
    const int ngroups = (nblk + BL_GROUP - 1) / BL_GROUP;
    const int64_t by_rows = (n + 1023) / 1024;
    const int64_t target = std::max<int64_t>(64, 1024 / ngroups);
    return (int)std::max<int64_t>(1, std::min(by_rows, target));
}

// doubles of the partial tiles (first stage) for up to maxblk blocks
size_t bl_partials_len(int64_t n, int maxblk) {
    size_t m = 0;
    for (int k = 1; k <= maxblk; ++k) m = std::max(m, (size_t)bl_chunks(n, k) * k * 256);
    return m;
}

// dC[0..nblk) = Vb^H X; copied to hc (nblk * 64, row-major blocks) when given
int bl_proj(BlockLz &L, const double2 *Vb, int nblk, const double2 *X, std::vector<cd> *hc) {
    const int64_t n = L.n;
    int nchunk = bl_chunks(n, nblk);
    const int64_t unit = 4 * BL_ROWS;
    const int64_t rows = ((n + nchunk - 1) / nchunk + unit - 1) / unit * unit;
    nchunk = (int)((n + rows - 1) / rows);
    const dim3 grid((unsigned)nchunk, (unsigned)((nblk + BL_GROUP - 1) / BL_GROUP));
    hipStream_t s = edx::stream();
    // algorithmic: every basis block and X once; 512 flops per row and block
    hipEvent_t pe = edx::prof_start(EDX_C128, EDX_KERNEL_BPROJ, (nblk + 1.0) * n * 128.0,
                                    512.0 * n * nblk);
    hipLaunchKernelGGL(k_bproj, grid, dim3(256), 0, s, n, (const double *)Vb, nblk,
                       (const double *)X, rows, L.part);
    edx::prof_stop(pe);
    const double *tiles = L.part;
    int ntiles = nchunk;
    if (nchunk > BL_RED) {   // second level of partial tiles after the first stage
        double *out = L.part + L.part_len;
        ntiles = (nchunk + BL_RED - 1) / BL_RED;
        hipLaunchKernelGGL(k_bproj_stage, dim3((unsigned)nblk, (unsigned)ntiles), dim3(256), 0, s,
                           L.part, nchunk, nblk, out);
        tiles = out;
    }
    hipLaunchKernelGGL(k_bproj_reduce, dim3((unsigned)nblk), dim3(256), 0, s, tiles, ntiles, nblk,
                       L.dC);
    EDX_LAUNCH_CHECK();
    if (hc) {
        hc->resize((size_t)nblk * 64);
        EDX_HIP(edx::copy_sync(hc->data(), L.dC, (size_t)nblk * 64 * sizeof(double2),
                          hipMemcpyDeviceToHost));
    }
    return EDX_OK;
}

int bl_upd(BlockLz &L, const double2 *Vb, int nblk, const double2 *C, const double2 *Win,
           double2 *Wout, double sgn) {
    const int64_t rows_per_wg = 4 * 16 * BL_UTILES;
    const unsigned nwg = (unsigned)((L.n + rows_per_wg - 1) / rows_per_wg);
    hipEvent_t pe = edx::prof_start(EDX_C128, EDX_KERNEL_BUPD,
                                    (nblk + (Win ? 2.0 : 1.0)) * L.n * 128.0, 512.0 * L.n * nblk);
    hipLaunchKernelGGL(k_bupd, dim3(nwg), dim3(256), 0, edx::stream(), L.n, Vb, nblk, C,
                       (const double *)Win, (double *)Wout, sgn);
    edx::prof_stop(pe);
    EDX_LAUNCH_CHECK();
    return EDX_OK;
}

// X -= V V^H X twice over nblk basis blocks; the two coefficient sets are
// summed into coef (nblk * 64) when given.
int bl_cgs2(BlockLz &L, int nblk, double2 *X, std::vector<cd> *coef) {
    if (nblk <= 0) return EDX_OK;
    if (coef) coef->assign((size_t)nblk * 64, cd(0.0, 0.0));
    for (int pass = 0; pass < 2; ++pass) {
        EDX_TRY(bl_proj(L, L.V, nblk, X, coef ? &L.tmpc : nullptr));
        EDX_TRY(bl_upd(L, L.V, nblk, L.dC, X, X, -1.0));
        if (coef)
            for (size_t i = 0; i < coef->size(); ++i) (*coef)[i] += L.tmpc[i];
    }
    return EDX_OK;
}

// Cholesky of the 8 x 8 Hermitian G + shift I = R^H R (R upper, row-major).
// piv[k] receives the k-th pivot; false if one is not positive and finite.
bool bl_chol(const cd *G, double shift, cd *R, double *piv) {
    for (int i = 0; i < 64; ++i) R[i] = 0.0;
    for (int k = 0; k < BW; ++k) {
        double dk = G[k * BW + k].real() + shift;
        for (int l = 0; l < k; ++l) dk -= std::norm(R[l * BW + k]);
        piv[k] = dk;
        if (!(dk > 0.0) || !std::isfinite(dk)) return false;
        const double rk = std::sqrt(dk);
        R[k * BW + k] = rk;
        for (int j = k + 1; j < BW; ++j) {
            cd v = G[k * BW + j];
            for (int l = 0; l < k; ++l) v -= std::conj(R[l * BW + k]) * R[l * BW + j];
            R[k * BW + j] = v / rk;
        }
    }
    return true;
}

// inverse of an upper-triangular 8 x 8 matrix (row-major)
void bl_inv_upper(const cd *R, cd *Ri) {
    for (int i = 0; i < 64; ++i) Ri[i] = 0.0;
    for (int j = 0; j < BW; ++j) {
        Ri[j * BW + j] = 1.0 / R[j * BW + j];
        for (int i = j - 1; i >= 0; --i) {
            cd s = 0.0;
            for (int k = i + 1; k <= j; ++k) s += R[i * BW + k] * Ri[k * BW + j];
            Ri[i * BW + j] = -s / R[i * BW + i];
        }
    }
}

// Q = Q R^{-1} with R^H R = Q^H Q + shift I (G: Q^H Q when already known);
// *ok false if the Cholesky fails.
int bl_cholqr_pass(BlockLz &L, double2 *Q, double shift_rel, const std::vector<cd> *Gin, bool *ok,
                   double *piv_ratio) {
    std::vector<cd> G;
    if (Gin) G = *Gin;
    else EDX_TRY(bl_proj(L, Q, 1, Q, &G));
    double tr = 0.0;
    for (int k = 0; k < BW; ++k) tr += G[k * BW + k].real();
    cd R[64], Ri[64];
    double piv[BW];
    *ok = tr > 0.0 && std::isfinite(tr) && bl_chol(G.data(), shift_rel * tr, R, piv);
    if (!*ok) return EDX_OK;
    if (piv_ratio) {
        double lo = piv[0], hi = piv[0];
        for (int k = 1; k < BW; ++k) {
            lo = std::min(lo, piv[k]);
            hi = std::max(hi, piv[k]);
        }
        *piv_ratio = lo / hi;
    }
    bl_inv_upper(R, Ri);
    EDX_HIP(edx::copy_sync(L.dC, Ri, sizeof Ri, hipMemcpyHostToDevice));
    return bl_upd(L, Q, 1, L.dC, nullptr, Q, 1.0);
}

// Makes the block Q orthonormal and orthogonal to the first nprev basis
// blocks (Q is already orthogonal to them up to rounding).  proj2[k] is the
// squared norm the projections against the basis removed from column k, so
// proj2[k] + |q_k|^2 is its norm before them.  A column is deficient when its
// Cholesky pivot (the squared norm of its direction of its own) is below
// 1e-24 of that norm, or of the mean squared norm of the block: what is left
// is rounding noise of the projections or of the other columns, which holds
// no new direction (an invariant subspace, or a column dependent on the
// others).  Deficient columns are replaced by random vectors orthogonalized
// against the basis.  *exhausted: no such vector is found (the basis spans
// the space).
int bl_orth(BlockLz &L, int nprev, double2 *Q, const double *proj2_in, bool *exhausted) {
    *exhausted = false;
    const double u = 1.1102230246251565e-16;
    const double shift_rel = 11.0 * (8.0 * (double)L.n + 72.0) * u;
    double proj2[BW];
    for (int k = 0; k < BW; ++k) proj2[k] = proj2_in ? proj2_in[k] : 0.0;
    for (int attempt = 0; attempt < 4; ++attempt) {
        std::vector<cd> G;
        EDX_TRY(bl_proj(L, Q, 1, Q, &G));
        double tr = 0.0;
        for (int k = 0; k < BW; ++k) tr += G[k * BW + k].real();
        // pivots of the unshifted Cholesky, continued past deficient columns
        std::vector<int> defic;
        {
            cd R[64];
            for (int i = 0; i < 64; ++i) R[i] = 0.0;
            for (int k = 0; k < BW; ++k) {
                const double before = proj2[k] + G[k * BW + k].real();
                const double floor_piv = 1e-24 * std::max(std::max(tr / BW, before), 1e-300);
                double dk = G[k * BW + k].real();
                for (int l = 0; l < k; ++l) dk -= std::norm(R[l * BW + k]);
                if (!(dk > floor_piv) || !std::isfinite(dk) || !(tr > 0.0)) {
                    defic.push_back(k);
                    continue;   // row k of R stays zero
                }
                const double rk = std::sqrt(dk);
                R[k * BW + k] = rk;
                for (int j = k + 1; j < BW; ++j) {
                    cd v = G[k * BW + j];
                    for (int l = 0; l < k; ++l) v -= std::conj(R[l * BW + k]) * R[l * BW + j];
                    R[k * BW + j] = v / rk;
                }
            }
        }
        if (!defic.empty()) {
            for (int k : defic) {
                hipLaunchKernelGGL(k_bl_random_col, dim3((unsigned)vec_blocks(L.n)), dim3(VB), 0,
                                   edx::stream(), L.n, k, L.seed + 0x51ed27ull * (uint64_t)++L.nrand,
                                   Q);
                proj2[k] = (double)L.n / 12.0;   // expected |q|^2 of the uniform column
            }
            EDX_LAUNCH_CHECK();
            EDX_TRY(bl_cgs2(L, nprev, Q, nullptr));
            continue;
        }
        bool ok = false;
        double ratio = 1.0;
        EDX_TRY(bl_cholqr_pass(L, Q, shift_rel, &G, &ok, &ratio));
        if (!ok) continue;
        EDX_TRY(bl_cholqr_pass(L, Q, 0.0, nullptr, &ok, nullptr));
        if (!ok) continue;
        // an ill-conditioned block carries rounding noise from the basis into
        // its weak directions: orthogonalize once more before the last pass
        if (ratio < 1e-8) EDX_TRY(bl_cgs2(L, nprev, Q, nullptr));
        EDX_TRY(bl_cholqr_pass(L, Q, 0.0, nullptr, &ok, nullptr));
        if (!ok) continue;
        return EDX_OK;
    }
    *exhausted = true;
    return EDX_OK;
}

int block_lanczos_impl(const edx_csr *h, int precision, int nev, int maxblocks, double tol,
                       uint64_t seed, int nvec, double *evals, void *d_evecs, int *nblocks_out,
                       double *resid_out) {
    EDX_TRY(check_square(h, precision));
    const int64_t n = h->n;
    EDX_CHECK(n >= 2 * BW, EDX_ERR_ARG, "block Lanczos needs at least %d rows (matrix has %lld)",
              2 * BW, (long long)n);
    EDX_CHECK(maxblocks >= 1 && nev >= 1 && evals && nvec >= 0 && nvec <= nev &&
                  (nvec == 0 || d_evecs) && tol >= 0.0,
              EDX_ERR_ARG, "bad argument (nev=%d maxblocks=%d nvec=%d)", nev, maxblocks, nvec);
    const int cap = (int)std::min<int64_t>(maxblocks, n / BW);   // steps; T is at most 8 cap
    EDX_CHECK(nev <= BW * cap, EDX_ERR_ARG, "nev %d exceeds the %d Ritz values of %d block steps",
              nev, BW * cap, cap);
    const size_t blk = (size_t)n * BW;
    DevBuf vb, wb, x32b, partb, cb;
    EDX_TRY(vb.alloc((size_t)(cap + 1) * blk * sizeof(double2), "block lanczos basis"));
    EDX_TRY(wb.alloc(blk * sizeof(double2), "block lanczos W"));
    if (precision == EDX_C64) EDX_TRY(x32b.alloc(blk * sizeof(float2), "block lanczos x32"));
    const size_t plen = bl_partials_len(n, cap + 1);
    EDX_TRY(partb.alloc((plen + plen / BL_RED + (size_t)(cap + 1) * 256) * sizeof(double),
                        "projection partials"));
    EDX_TRY(cb.alloc((size_t)(cap + 1) * 64 * sizeof(double2), "block coefficients"));
    BlockLz L{h, precision, n, vb.as<double2>(), wb.as<double2>(), x32b.as<float2>(),
              partb.as<double>(), plen, cb.as<double2>(), seed};
    hipStream_t s = edx::stream();

    // Q_0: real uniform random block, orthonormalized
    hipLaunchKernelGGL(k_random_real, dim3((unsigned)vec_blocks((int64_t)blk)), dim3(VB), 0, s,
                       (int64_t)blk, seed, L.V);
    EDX_LAUNCH_CHECK();
    bool exhausted = false;
    EDX_TRY(bl_orth(L, 0, L.V, nullptr, &exhausted));
    EDX_CHECK(!exhausted, EDX_ERR_NUMERIC, "could not orthonormalize the start block");

    std::vector<cd> Ablk, Bblk;   // A_j (hermitized) and B_{j+1}, 64 each per step
    std::vector<double> ritz;
    std::vector<cd> Z;
    std::vector<double> res(nev, 0.0);
    int steps = 0, m = 0, last_check = 0;
    for (int j = 0; j < cap; ++j) {
        double2 *Qj = L.V + (size_t)j * blk;
        BEpi e{};
        e.y = L.W;
        if (precision == EDX_C64) ELAUNCH(k_il_to_c64, n, BW, n, BW, Qj, L.x32);
        EDX_TRY(edx::bspmm(h, precision, BW, Qj, L.x32, e, nullptr));
        std::vector<cd> coef;
        EDX_TRY(bl_cgs2(L, j + 1, L.W, &coef));
        for (int a = 0; a < BW; ++a)
            for (int b = 0; b < BW; ++b) {
                const cd x = coef[(size_t)j * 64 + a * BW + b];
                const cd y = std::conj(coef[(size_t)j * 64 + b * BW + a]);
                Ablk.push_back(0.5 * (x + y));
            }
        steps = j + 1;
        m = BW * steps;
        double2 *Qn = L.V + (size_t)(j + 1) * blk;
        EDX_HIP(hipMemcpyAsync(Qn, L.W, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
        double proj2[BW] = {0.0};   // |V^H w_k|^2, removed by the projections
        for (int b = 0; b <= j; ++b)
            for (int r = 0; r < BW; ++r)
                for (int k = 0; k < BW; ++k) proj2[k] += std::norm(coef[(size_t)b * 64 + r * BW + k]);
        EDX_TRY(bl_orth(L, j + 1, Qn, proj2, &exhausted));
        std::vector<cd> B;
        if (exhausted) B.assign(64, cd(0.0, 0.0));
        else EDX_TRY(bl_proj(L, Qn, 1, L.W, &B));
        Bblk.insert(Bblk.end(), B.begin(), B.end());
        const bool last = exhausted || steps == cap || m >= n;
        // the projected eigensolve costs O(m^3) on the host: every step up to
        // 16 steps, then at 1/8 growth of the basis
        const int every = steps <= 16 ? 1 : std::max(1, steps / 8);
        if (m < nev || (!last && steps - last_check < every)) continue;
        last_check = steps;
        // T = V^H H V, block tridiagonal
        std::vector<cd> T((size_t)m * m, cd(0.0, 0.0));
        for (int b = 0; b < steps; ++b)
            for (int r = 0; r < BW; ++r)
                for (int c = 0; c < BW; ++c) {
                    T[(size_t)(b * BW + r) * m + b * BW + c] = Ablk[(size_t)b * 64 + r * BW + c];
                    if (b + 1 < steps) {
                        const cd x = Bblk[(size_t)b * 64 + r * BW + c];   // T_{b+1,b}[r][c]
                        T[(size_t)((b + 1) * BW + r) * m + b * BW + c] = x;
                        T[(size_t)(b * BW + c) * m + (b + 1) * BW + r] = std::conj(x);
                    }
                }
        ritz.assign(m, 0.0);
        Z.assign((size_t)m * nev, cd(0.0, 0.0));   // m x nev, column k: Ritz vector k in V
        EDX_CHECK(edx::herm_eig_lowest(m, T.data(), ritz.data(), nev, Z.data()) == 0,
                  EDX_ERR_NUMERIC, "projected eigensolver did not converge");
        const cd *Bl = Bblk.data() + (size_t)(steps - 1) * 64;
        bool conv = true;
        for (int k = 0; k < nev; ++k) {
            double r2 = 0.0;
            for (int r = 0; r < BW; ++r) {
                cd acc = 0.0;
                for (int c = 0; c < BW; ++c) acc += Bl[r * BW + c] * Z[(size_t)(m - BW + c) * nev + k];
                r2 += std::norm(acc);
            }
            res[k] = std::sqrt(r2);
            if (!(res[k] < tol)) conv = false;
        }
        if (conv || last) break;
    }
    EDX_CHECK(!ritz.empty() && (int)ritz.size() == m, EDX_ERR_NUMERIC,
              "no Ritz values (%d block steps)", steps);
    for (int k = 0; k < nev; ++k) evals[k] = ritz[k];
    if (resid_out)
        for (int k = 0; k < nev; ++k) resid_out[k] = res[k];
    // Ritz vectors X = V S, eight at a time
    for (int g = 0; g * BW < nvec; ++g) {
        std::vector<cd> S((size_t)steps * 64, cd(0.0, 0.0));
        for (int b = 0; b < steps; ++b)
            for (int r = 0; r < BW; ++r)
                for (int c = 0; c < BW && g * BW + c < nev; ++c)
                    S[(size_t)b * 64 + r * BW + c] = Z[(size_t)(b * BW + r) * nev + g * BW + c];
        EDX_HIP(edx::copy_sync(L.dC, S.data(), S.size() * sizeof(cd), hipMemcpyHostToDevice));
        EDX_TRY(bl_upd(L, L.V, steps, L.dC, nullptr, L.W, 1.0));
        const int nc = std::min(BW, nvec - g * BW);
        hipLaunchKernelGGL(k_bl_to_cols, dim3((unsigned)vec_blocks(n)), dim3(VB), 0, s, n, L.W, nc,
                           (double2 *)d_evecs + (size_t)g * BW * n);
        EDX_LAUNCH_CHECK();
    }
    EDX_HIP(hipStreamSynchronize(s));
    if (nblocks_out) *nblocks_out = steps;
    return EDX_OK;
}

}  // namespace

extern "C" {

int edx_block_lanczos_eigs(const edx_csr *h, int precision, int nev, int maxblocks, double tol,
                           uint64_t seed, int nvec, double *evals, void *d_evecs,
                           int *nblocks_out, double *resid_out) {
    return block_lanczos_impl(h, precision, nev, maxblocks, tol, seed, nvec, evals, d_evecs,
                              nblocks_out, resid_out);
}

}  // extern "C"
