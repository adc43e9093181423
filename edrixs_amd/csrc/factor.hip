// Slot-factored form of a core-hole Hamiltonian (part of edx_kernels.hip).
//
// In the interleaved core-hole layout row r = v * nc + s holds valence
// configuration v with the core hole in slot s.  A term on valence orbitals
// only leaves the core word alone, and its fermion sign counts occupied
// orbitals below its own position, which are all valence orbitals (the core
// bits sit above them, src/ham.f90:419-446); so its contribution is the same
// for every slot: that part of H is exactly H_v (x) I_nc.  The build
// (src/ham.f90:352-711) sums every term of an entry, so the form is found
// from the finished CSR, not from the terms: for each valence row v, the
// off-diagonal entries that keep the slot (column c with c % nc == s) are
// compared across the nc rows v * nc + s, value bit for value bit.  Where
// they agree they are stored once, as valence row v of H_v (valence column,
// value number); every other off-diagonal entry stays in a per-row remainder
// R.  A valence row whose slot rows disagree keeps all its entries in R.
//
//     (H x)[v nc + s] = diag[r] x[r] + sum_k Hv[v, k] x[k nc + s] + sum_j R[r, j] x[j]
//
// The values are the entries of the exact complex128 CSR, referenced by their
// numbers in the indexed form's table, so products are those of the FP64
// operator (the summation order differs from the row-wise product, as it
// does between any two schedules).  The payoff is in the gathers: the nc
// lanes of a valence row read x[k nc + 0 .. nc - 1], nc * 16 contiguous
// bytes, for each H_v entry, where the row-wise product makes nc separate
// 16-byte gathers from nc different nonzeros.

namespace {

constexpr uint32_t FAC_COL_MASK = (1u << PK_COL_BITS) - 1;

// Typed blocks of the slot-changing entries.  For a valence row v whose slot
// rows share their slot-keeping entries, every other off-diagonal entry moves
// the core hole: row v NC + s, column v' NC + s + d with d != 0.  Grouped by
// (v', d), the entries of one group form a block of at most NC entries, one
// per row s; its "type" is d and the value number of each row's entry
// (0xff where row s has none).  The distinct types are few (core-valence
// exchange couples each valence move to one core shift, with coefficients
// from a handful of Coulomb integrals), so a block is one word, valence
// column | type << FAC_TYPE_SHIFT, and the types sit in a small table in LDS.
// The NC lanes of valence row v then gather one x segment per block, where
// the row-wise remainder made one gather per entry and lane.
constexpr int FAC_TYPE_SHIFT = 22;
constexpr uint32_t FAC_VCOL_MASK = (1u << FAC_TYPE_SHIFT) - 1;
constexpr int FAC_MAX_TYPES = 1 << (32 - FAC_TYPE_SHIFT);
constexpr int FAC_TYPE_SET = 4096;          // hash set slots for the types (power of 2)
constexpr uint64_t FAC_EMPTY = ~0ull;
constexpr uint64_t FAC_NO_VALUES = 0xffffffffffffull;   // six 0xff value numbers

__device__ __forceinline__ uint32_t fac_type_hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return (uint32_t)k & (FAC_TYPE_SET - 1);
}

// Inserts a type into the set; false when the set is full.  A plain read
// comes first: nearly every block's type is already there, and a
// compare-and-swap on the few slots all threads hit would serialize them.
__device__ bool fac_type_insert(unsigned long long *set, uint64_t key) {
    uint32_t h = fac_type_hash(key);
    for (int p = 0; p < FAC_TYPE_SET; ++p) {
        const unsigned long long cur = ((volatile unsigned long long *)set)[h];
        if (cur == key) return true;
        if (cur == FAC_EMPTY) {
            const unsigned long long old = atomicCAS(set + h, (unsigned long long)FAC_EMPTY,
                                                     (unsigned long long)key);
            if (old == FAC_EMPTY || old == key) return true;
        }
        h = (h + 1) & (FAC_TYPE_SET - 1);
    }
    return false;
}

// Slot of a type present in the set.
__device__ int fac_type_find(const unsigned long long *set, uint64_t key) {
    uint32_t h = fac_type_hash(key);
    for (int p = 0; p < FAC_TYPE_SET; ++p) {
        const unsigned long long k = set[h];
        if (k == key) return (int)h;
        if (k == FAC_EMPTY) return -1;
        h = (h + 1) & (FAC_TYPE_SET - 1);
    }
    return -1;
}

// Calls emit(valence column, type) for each block of valence row v, in
// increasing valence column and then shift: a merge over the NC rows' column
// lists (sorted), visiting the entries that change the slot.
template <int NC, class Emit>
__device__ void fac_visit_blocks(int64_t v, const int64_t *orow, const int32_t *ocols,
                                 const uint16_t *oidx, Emit emit) {
    int64_t j[NC], e[NC];
#pragma unroll
    for (int s = 0; s < NC; ++s) {
        j[s] = orow[v * NC + s];
        e[s] = orow[v * NC + s + 1];
    }
    while (true) {
        int64_t vmin = INT64_MAX;
#pragma unroll
        for (int s = 0; s < NC; ++s) {
            while (j[s] < e[s] && ocols[j[s]] % NC == s) ++j[s];   // slot-keeping: H_v
            if (j[s] < e[s]) vmin = min(vmin, (int64_t)(ocols[j[s]] / NC));
        }
        if (vmin == INT64_MAX) break;
        uint64_t desc[2 * NC - 1];
#pragma unroll
        for (int d = 0; d < 2 * NC - 1; ++d) desc[d] = FAC_NO_VALUES;
        uint32_t used = 0;
#pragma unroll
        for (int s = 0; s < NC; ++s) {
            for (; j[s] < e[s] && ocols[j[s]] / NC == vmin; ++j[s]) {
                const int sc = ocols[j[s]] % NC;
                if (sc == s) continue;
                const int d = sc - s + NC - 1;
                desc[d] = (desc[d] & ~(0xffull << (8 * s))) | ((uint64_t)oidx[j[s]] << (8 * s));
                used |= 1u << d;
            }
        }
#pragma unroll
        for (int d = 0; d < 2 * NC - 1; ++d)
            if (used >> d & 1) emit(vmin, desc[d] | ((uint64_t)d << 48));
    }
}

// per valence row: whether the slot-keeping entries agree across the nc slot
// rows, and the entry counts of H_v and of the typed blocks (per valence row)
// and of R (per row).  With use_b (nc == 6), a row that agrees keeps its
// slot-changing entries as typed blocks, whose types go into the hash set;
// otherwise they stay in R.  A failed insertion (set full) raises *overflow.
__global__ void k_fac_count(int64_t mv, int nc, const int64_t *orow, const int32_t *ocols,
                            const uint16_t *oidx, uint8_t *kron, int64_t *vcnt, int64_t *bcnt,
                            int64_t *rcnt, int use_b, unsigned long long *tset, int *overflow) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < mv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r0 = v * nc;
        // slot-0 entries of row r0 that keep the slot: the reference list
        int64_t n0 = 0;
        for (int64_t j = orow[r0]; j < orow[r0 + 1]; ++j) n0 += (ocols[j] % nc == 0);
        bool same = true;
        for (int s = 1; s < nc && same; ++s) {
            const int64_t r = r0 + s;
            int64_t j0 = orow[r0];
            const int64_t e0 = orow[r0 + 1];
            int64_t ns = 0;
            for (int64_t j = orow[r]; j < orow[r + 1] && same; ++j) {
                const int32_t c = ocols[j];
                if (c % nc != s) continue;
                ++ns;
                while (j0 < e0 && ocols[j0] % nc != 0) ++j0;   // next slot-0 reference entry
                if (j0 == e0 || ocols[j0] / nc != c / nc || oidx[j0] != oidx[j]) same = false;
                ++j0;
            }
            if (ns != n0) same = false;
        }
        kron[v] = same ? 1 : 0;
        vcnt[v] = same ? n0 : 0;
        const bool blocks = same && use_b;
        int64_t nb = 0;
        if (blocks) {
            uint64_t last[2] = {FAC_EMPTY, FAC_EMPTY};   // the types this thread placed last
            fac_visit_blocks<6>(v, orow, ocols, oidx, [&](int64_t, uint64_t type) {
                ++nb;
                if (type == last[0] || type == last[1]) return;
                if (!fac_type_insert(tset, type)) atomicOr(overflow, 1);
                last[1] = last[0];
                last[0] = type;
            });
        }
        bcnt[v] = nb;
        for (int s = 0; s < nc; ++s) {
            const int64_t r = r0 + s;
            int64_t k = 0;
            if (!blocks)
                for (int64_t j = orow[r]; j < orow[r + 1]; ++j)
                    k += !(same && ocols[j] % nc == s);
            rcnt[r] = k;
        }
    }
}

__global__ void k_fac_fill(int64_t mv, int nc, const int64_t *orow, const int32_t *ocols,
                           const uint16_t *oidx, const uint8_t *kron, const int64_t *vrow,
                           uint32_t *vpk, const int64_t *brow, uint32_t *bpk, const int64_t *rrow,
                           uint32_t *rpk, int use_b, const unsigned long long *tset,
                           const int16_t *tid) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < mv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r0 = v * nc;
        const bool same = kron[v] != 0;
        if (same) {
            int64_t o = vrow[v];
            for (int64_t j = orow[r0]; j < orow[r0 + 1]; ++j)
                if (ocols[j] % nc == 0)
                    vpk[o++] = (uint32_t)(ocols[j] / nc) | ((uint32_t)oidx[j] << PK_COL_BITS);
        }
        const bool blocks = same && use_b;
        if (blocks) {
            int64_t o = brow[v];
            fac_visit_blocks<6>(v, orow, ocols, oidx, [&](int64_t vc, uint64_t type) {
                // every type was inserted by k_fac_count, so the lookup finds it
                const int slot = fac_type_find(tset, type);
                bpk[o++] = (uint32_t)vc | ((uint32_t)tid[slot < 0 ? 0 : slot] << FAC_TYPE_SHIFT);
            });
        } else {
            for (int s = 0; s < nc; ++s) {
                const int64_t r = r0 + s;
                int64_t o = rrow[r];
                for (int64_t j = orow[r]; j < orow[r + 1]; ++j)
                    if (!(same && ocols[j] % nc == s))
                        rpk[o++] = (uint32_t)ocols[j] | ((uint32_t)oidx[j] << PK_COL_BITS);
            }
        }
    }
}

struct FacArgs {
    int64_t mv;              // valence rows (m / NC)
    const int64_t *vrow;     // mv + 1
    const uint32_t *vpk;     // valence column | value number << 24
    const int64_t *brow;     // mv + 1, or null: no typed blocks
    const uint32_t *bpk;     // valence column | type << FAC_TYPE_SHIFT
    const int64_t *rrow;     // m + 1, or null: no remainder
    const uint32_t *rpk;     // column | value number << 24
    const double2 *tab;      // value table (ntab entries, staged in LDS)
    int ntab;
    const uint64_t *typ;     // block types (ntyp entries, staged in LDS after the table)
    int ntyp;
    const double2 *diag;     // exact diagonal, m entries
};

// Stages the value table and the block types of a factored matrix in LDS
// (s_tab: ntab double2, then ntyp uint64); returns the types' LDS address.
__device__ __forceinline__ const uint64_t *fac_stage_lds(const FacArgs &f, double2 *s_tab) {
    uint64_t *s_typ = reinterpret_cast<uint64_t *>(s_tab + f.ntab);
    for (int i = threadIdx.x; i < f.ntab; i += blockDim.x) s_tab[i] = f.tab[i];
    for (int i = threadIdx.x; i < f.ntyp; i += blockDim.x) s_typ[i] = f.typ[i];
    __syncthreads();
    return s_typ;
}

// Off-diagonal sum of row r = v NC + s in the slot-factored form: the H_v
// entries of valence row v gathered at slot s, the valence row's typed blocks,
// then the row's remainder R.  The first UR words of the remainder row are
// loaded into registers before H_v is walked, so those loads overlap the
// valence gathers; H_v and the blocks are walked four entries at a time (four
// independent gathers in flight).  s_tab / s_typ: the value table and the
// block types staged in LDS.
template <int NC, int UR>
__device__ __forceinline__ double2 fac_offdiag_sum(const FacArgs &f, const double2 *__restrict__ x,
                                                   const double2 *s_tab, const uint64_t *s_typ,
                                                   int64_t v, int s, int64_t r) {
    int64_t jb = 0, je = 0;
    if (f.rrow) {
        jb = f.rrow[r];
        je = f.rrow[r + 1];
    }
    uint32_t rw[UR];
#pragma unroll
    for (int u = 0; u < UR; ++u) rw[u] = jb + u < je ? f.rpk[jb + u] : 0u;
    const int64_t kb = f.vrow[v], ke = f.vrow[v + 1];
    double sr = 0.0, si = 0.0;
    auto fma_c = [&](double2 a, double2 xv) {
        sr += a.x * xv.x - a.y * xv.y;
        si += a.x * xv.y + a.y * xv.x;
    };
    int64_t k = kb;
    for (; k + 4 <= ke; k += 4) {
        const uint32_t w0 = f.vpk[k], w1 = f.vpk[k + 1], w2 = f.vpk[k + 2], w3 = f.vpk[k + 3];
        const double2 x0 = x[(int64_t)(w0 & FAC_COL_MASK) * NC + s];
        const double2 x1 = x[(int64_t)(w1 & FAC_COL_MASK) * NC + s];
        const double2 x2 = x[(int64_t)(w2 & FAC_COL_MASK) * NC + s];
        const double2 x3 = x[(int64_t)(w3 & FAC_COL_MASK) * NC + s];
        fma_c(s_tab[w0 >> PK_COL_BITS], x0);
        fma_c(s_tab[w1 >> PK_COL_BITS], x1);
        fma_c(s_tab[w2 >> PK_COL_BITS], x2);
        fma_c(s_tab[w3 >> PK_COL_BITS], x3);
    }
    for (; k < ke; ++k) {
        const uint32_t w = f.vpk[k];
        fma_c(s_tab[w >> PK_COL_BITS], x[(int64_t)(w & FAC_COL_MASK) * NC + s]);
    }
    if (f.brow) {
        const int64_t bb = f.brow[v], be = f.brow[v + 1];
        int64_t q = bb;
        for (; q + 4 <= be; q += 4) {
            uint32_t n[4];
            int64_t c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t w = f.bpk[q + u];
                const uint64_t d = s_typ[w >> FAC_TYPE_SHIFT];
                n[u] = (uint32_t)(d >> (8 * s)) & 0xffu;
                c[u] = (int64_t)(w & FAC_VCOL_MASK) * NC + s + (int)(d >> 48) - (NC - 1);
            }
            double2 xv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) xv[u] = n[u] != 0xffu ? x[c[u]] : make_double2(0.0, 0.0);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (n[u] != 0xffu) fma_c(s_tab[n[u]], xv[u]);
        }
        for (; q < be; ++q) {
            const uint32_t w = f.bpk[q];
            const uint64_t d = s_typ[w >> FAC_TYPE_SHIFT];
            const uint32_t nn = (uint32_t)(d >> (8 * s)) & 0xffu;
            if (nn != 0xffu)
                fma_c(s_tab[nn], x[(int64_t)(w & FAC_VCOL_MASK) * NC + s + (int)(d >> 48) - (NC - 1)]);
        }
    }
#pragma unroll
    for (int u = 0; u < UR; ++u)
        if (jb + u < je) fma_c(s_tab[rw[u] >> PK_COL_BITS], x[rw[u] & FAC_COL_MASK]);
    for (int64_t j = jb + UR; j < je; ++j) {
        const uint32_t w = f.rpk[j];
        fma_c(s_tab[w >> PK_COL_BITS], x[w & FAC_COL_MASK]);
    }
    return make_double2(sr, si);
}

// y = c0 * (H x) + c1 * z (+ <d|y> block partials) in the slot-factored form.
// Lane l of a wave owns valence row v = wave_base + l / NC and slot s = l % NC
// (64 / NC valence rows per wave, 64 % NC lanes idle), so the NC lanes of a
// valence row read each H_v entry's x[k NC + s] as one contiguous NC * 16-byte
// segment, and the rows a wave writes are consecutive.  The epilogue's
// operands (diagonal, x[r], addend, dot vector) are loaded before the sums so
// they arrive with the first gathers.  One grid pass, no stride: workgroups
// are dispatched in row order, so the rows in flight stay in one window whose
// gathers share cache lines.
template <int NC, int UR, int WPE>
__global__ __launch_bounds__(256, WPE) void k_spmv_fac(FacArgs f, const double2 *__restrict__ x,
                                                       SpmvEpi e) {
    if (e.run && *e.run == 0) return;
    extern __shared__ double2 s_tab[];
    const uint64_t *s_typ = fac_stage_lds(f, s_tab);
    constexpr int VPW = 64 / NC;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int vl = lane / NC, s = lane - vl * NC;
    const int64_t bid = e.xmap ? (int64_t)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8
                               : (int64_t)blockIdx.x;
    const int64_t v = (bid * 4 + wave) * VPW + vl;
    double2 dot = make_double2(0.0, 0.0);
    if (vl < VPW && v < f.mv) {
        const int64_t r = v * NC + s;
        // the row's single-use operands and its output go non-temporal, so they
        // do not push the reused x lines out of L2 (2 % faster on H_n, norb 24)
        const double2 dg = nt_load2(f.diag + r), xr = x[r];
        double2 pz = make_double2(0, 0), pd = pz;
        if (e.z) pz = nt_load2(e.z + r);
        if (e.dvec) pd = nt_load2(e.dvec + r);
        const double2 ax = cadd(fac_offdiag_sum<NC, UR>(f, x, s_tab, s_typ, v, s, r), cmul(dg, xr));
        const double2 c0 = e.c0 ? *e.c0 : make_double2(1.0, 0.0);
        const double2 c1 = e.c1 ? *e.c1 : make_double2(1.0, 0.0);
        double2 out = cmul(c0, ax);
        if (e.z) out = cadd(out, cmul(c1, pz));
        nt_store2(e.y + r, out);
        if (e.dvec) dot = cconjmul(pd, out);
    }
    if (e.dvec) {
        const double2 bs = block_sum_256(dot);
        if (threadIdx.x == 0) e.partials[blockIdx.x] = bs;
    }
}

// Multi-RHS factored product for small K (interleaved right-hand sides,
// X[r K + k], and the batched epilogue of batch.hip): lane (v, s) as in
// k_spmv_fac computes its row for all K right-hand sides, so an H_v entry
// gathers NC K 16-byte values contiguous over the valence row's lanes and an
// R entry K contiguous values.  A capped grid (a multiple of 8 workgroups)
// walks XCD-contiguous eighths of the valence rows, 4 (64 / NC) valence rows
// per workgroup per pass; dot partials per workgroup and right-hand side.
template <int NC, int K>
__global__ __launch_bounds__(256) void k_spmm_fac(FacArgs f, const double2 *__restrict__ X, BEpi e) {
    extern __shared__ double2 s_tab[];
    __shared__ double2 red[4][K];
    const uint64_t *s_typ = fac_stage_lds(f, s_tab);
    constexpr int VPW = 64 / NC;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int vl = lane / NC, s = lane - vl * NC;
    const int64_t per_x = gridDim.x / 8, xcd = blockIdx.x % 8, wg = blockIdx.x / 8;
    const int64_t vrows_x = (f.mv + 7) / 8;
    const int64_t vbeg = xcd * vrows_x, vend = min(f.mv, vbeg + vrows_x);
    double2 dot[K];
#pragma unroll
    for (int k = 0; k < K; ++k) dot[k] = make_double2(0, 0);
    if (vl < VPW) {
        for (int64_t v = vbeg + (wg * 4 + wave) * VPW + vl; v < vend; v += per_x * 4 * VPW) {
            const int64_t r = v * NC + s;
            double sr[K], si[K];
#pragma unroll
            for (int k = 0; k < K; ++k) sr[k] = si[k] = 0.0;
            auto add = [&](double2 a, int64_t base) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const double2 xv = X[base + k];
                    sr[k] += a.x * xv.x - a.y * xv.y;
                    si[k] += a.x * xv.y + a.y * xv.x;
                }
            };
            // the values of K right-hand sides at one column, then their sums:
            // several entries' gathers are issued before any is summed (the
            // sums still run entry by entry)
            auto acc = [&](double2 a, const double2 *xv) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    sr[k] += a.x * xv[k].x - a.y * xv[k].y;
                    si[k] += a.x * xv[k].y + a.y * xv[k].x;
                }
            };
            const int64_t kb = f.vrow[v], ke = f.vrow[v + 1];
            int64_t q = kb;
            for (; q + 4 <= ke; q += 4) {
                uint32_t w[4];
                double2 xv[4][K];
#pragma unroll
                for (int u = 0; u < 4; ++u) w[u] = f.vpk[q + u];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        xv[u][k] = X[((int64_t)(w[u] & FAC_COL_MASK) * NC + s) * K + k];
#pragma unroll
                for (int u = 0; u < 4; ++u) acc(s_tab[w[u] >> PK_COL_BITS], xv[u]);
            }
            for (; q < ke; ++q) {
                const uint32_t w = f.vpk[q];
                add(s_tab[w >> PK_COL_BITS], ((int64_t)(w & FAC_COL_MASK) * NC + s) * K);
            }
            if (f.brow) {
                const int64_t bb = f.brow[v], be = f.brow[v + 1];
                int64_t p = bb;
                for (; p + 2 <= be; p += 2) {
                    uint32_t nn[2];
                    int64_t base[2];
                    double2 xv[2][K];
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const uint32_t w = f.bpk[p + u];
                        const uint64_t d = s_typ[w >> FAC_TYPE_SHIFT];
                        nn[u] = (uint32_t)(d >> (8 * s)) & 0xffu;
                        base[u] = ((int64_t)(w & FAC_VCOL_MASK) * NC + s + (int)(d >> 48) - (NC - 1)) * K;
                    }
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int k = 0; k < K; ++k)
                            xv[u][k] = nn[u] != 0xffu ? X[base[u] + k] : make_double2(0.0, 0.0);
#pragma unroll
                    for (int u = 0; u < 2; ++u)
                        if (nn[u] != 0xffu) acc(s_tab[nn[u]], xv[u]);
                }
                for (; p < be; ++p) {
                    const uint32_t w = f.bpk[p];
                    const uint64_t d = s_typ[w >> FAC_TYPE_SHIFT];
                    const uint32_t nn = (uint32_t)(d >> (8 * s)) & 0xffu;
                    if (nn != 0xffu)
                        add(s_tab[nn], ((int64_t)(w & FAC_VCOL_MASK) * NC + s + (int)(d >> 48) - (NC - 1)) * K);
                }
            }
            if (f.rrow) {
                const int64_t jb = f.rrow[r], je = f.rrow[r + 1];
                for (int64_t j = jb; j < je; ++j) {
                    const uint32_t w = f.rpk[j];
                    add(s_tab[w >> PK_COL_BITS], (int64_t)(w & FAC_COL_MASK) * K);
                }
            }
            add(f.diag[r], r * K);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int64_t o = r * K + k;
                const double2 out = epilogue(e, k, o, make_double2(sr[k], si[k]), X);
                e.y[o] = out;
                if (e.dvec) dot[k] = cadd(dot[k], cconjmul(e.dvec[o], out));
            }
        }
    }
    if (e.dvec) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double2 t = dot[k];
            for (int off = 32; off > 0; off >>= 1) {
                t.x += __shfl_xor(t.x, off, 64);
                t.y += __shfl_xor(t.y, off, 64);
            }
            if (lane == 0) red[wave][k] = t;
        }
        __syncthreads();
        if (threadIdx.x < K) {
            const int k = threadIdx.x;
            e.partials[(int64_t)blockIdx.x * K + k] =
                cadd(cadd(red[0][k], red[1][k]), cadd(red[2][k], red[3][k]));
        }
    }
}

}  // namespace

namespace edx {

// Valence rows per workgroup of the factored product (4 waves).
inline int64_t fac_rows_per_block(int nc) { return 4 * (64 / nc); }

int fac_blocks(const edx_csr *a) {
    const int64_t mv = a->m / a->nc_slots;
    int64_t blocks = (mv + fac_rows_per_block(a->nc_slots) - 1) / fac_rows_per_block(a->nc_slots);
    if (blocks < 1) blocks = 1;
    if (blocks >= 64) blocks = (blocks + 7) / 8 * 8;
    return (int)blocks;
}

// Kernel arguments of a matrix's factored form: the typed blocks and the
// remainder only where they hold entries.
FacArgs fac_args(const edx_csr *a) {
    return FacArgs{a->m / a->nc_slots,
                   a->d_fv_rowptr,
                   a->d_fv_pk,
                   a->fb_nnz ? a->d_fb_rowptr : nullptr,
                   a->d_fb_pk,
                   a->fr_nnz ? a->d_fr_rowptr : nullptr,
                   a->d_fr_pk,
                   a->d_vtab,
                   (int)a->nvtab,
                   a->d_ftyp,
                   a->nftyp,
                   a->d_diag};
}

// LDS bytes of the factored kernels: the value table, then the block types.
size_t fac_lds_bytes(const edx_csr *a) {
    return (size_t)a->nvtab * sizeof(double2) + (size_t)a->nftyp * sizeof(uint64_t);
}

// Launch of the factored product (schedule FAC_CODE); returns the block count.
int fac_launch(const edx_csr *a, const double2 *x, const SpmvEpi &e_in, int *nblocks_out) {
    if (!a->d_fr_rowptr)
        return set_error(EDX_ERR_ARG, "SpMV schedule %d needs the slot-factored form", FAC_CODE);
    const int nb = fac_blocks(a);
    if (a->m == 0) return EDX_OK;
    if (nblocks_out) *nblocks_out = nb;
    SpmvEpi e = e_in;
    e.xmap = nb % 8 == 0 && xcd_map_enabled();
    if (a->nc_slots != 6)
        return set_error(EDX_ERR_ARG, "slot-factored product for %d core slots not built",
                         a->nc_slots);
    const FacArgs f = fac_args(a);
    const size_t lds = fac_lds_bytes(a);
    hipLaunchKernelGGL((k_spmv_fac<6, 8, 4>), dim3(nb), dim3(256), lds, stream(), f, x, e);
    EDX_LAUNCH_CHECK();
    return EDX_OK;
}

// Launch of the multi-RHS factored product with bpc workgroups per CU
// (K = 2, 3 or 4; the caller has checked the form exists).
int fac_spmm_launch(const edx_csr *a, int K, const double2 *X, const BEpi &e, int bpc) {
    const int nb = (num_cus() * bpc + 7) / 8 * 8;
    const FacArgs f = fac_args(a);
    const size_t lds = fac_lds_bytes(a);
    if (a->m == 0) return EDX_OK;
    switch (K) {
        case 2: hipLaunchKernelGGL((k_spmm_fac<6, 2>), dim3(nb), dim3(256), lds, stream(), f, X, e); break;
        case 3: hipLaunchKernelGGL((k_spmm_fac<6, 3>), dim3(nb), dim3(256), lds, stream(), f, X, e); break;
        case 4: hipLaunchKernelGGL((k_spmm_fac<6, 4>), dim3(nb), dim3(256), lds, stream(), f, X, e); break;
        default: return set_error(EDX_ERR_ARG, "factored multi-RHS product for K = %d not built", K);
    }
    EDX_LAUNCH_CHECK();
    return EDX_OK;
}

// Stored bytes of the factored form that one product reads: the words and
// row pointers of H_v, of the typed blocks and of R (where they hold
// entries), the value table, the block types and the diagonal.
static double fac_stored_bytes(const edx_csr *a) {
    const double mv = a->nc_slots > 0 ? (double)(a->m / a->nc_slots) : 0.0;
    return (double)(a->fv_nnz + a->fb_nnz + a->fr_nnz) * 4.0 + (mv + 1.0) * 8.0 +
           (a->fb_nnz ? (mv + 1.0) * 8.0 : 0.0) + (a->fr_nnz ? (double)(a->m + 1) * 8.0 : 0.0) +
           (double)a->nvtab * 16.0 + (double)a->nftyp * 8.0 + (double)a->m * 16.0;
}

// Algorithmic bytes of the factored product: the stored form, x once, y
// written and the epilogue's addend and dot vector.
double fac_bytes(const edx_csr *a, bool z, bool dvec) {
    return fac_stored_bytes(a) + (double)a->n * 16.0 +
           (double)a->m * 16.0 * (1.0 + (z ? 1.0 : 0.0) + (dvec ? 1.0 : 0.0));
}

// The same for K interleaved right-hand sides: the stored form once, the
// vector terms K times.
double fac_bytes_k(const edx_csr *a, int K, bool z, bool dvec) {
    return fac_stored_bytes(a) +
           K * ((double)a->n * 16.0 + (double)a->m * 16.0 * (1.0 + (z ? 1.0 : 0.0) + (dvec ? 1.0 : 0.0)));
}

// Builds the slot-factored form from the indexed form.  Only square matrices
// in the interleaved core-hole layout with a packable indexed form (fewer
// than 2^24 columns, at most 256 values) and a built core-slot count qualify;
// for others nothing is made and the row-wise forms serve.  The typed blocks
// need 6 core slots, fewer than 2^22 valence rows, at most 255 values (0xff
// marks an absent entry) and at most 1024 distinct types; without them (or
// with EDX_NO_FAC_BLOCKS=1) the slot-changing entries stay in R.
int csr_add_factored(edx_csr *a) {
    if (a->d_fr_rowptr) return EDX_OK;
    const int nc = a->nc_slots;
    if (nc != 6 || a->m != a->n || a->m % nc != 0 || !a->d_off_idx) return EDX_OK;
    if (a->n >= (int64_t(1) << PK_COL_BITS) || a->nvtab > (int64_t(1) << (32 - PK_COL_BITS)))
        return EDX_OK;
    const char *off = getenv("EDX_NO_FACTOR");
    if (off && off[0] == '1') return EDX_OK;
    const int64_t m = a->m, mv = m / nc;
    const char *nob = getenv("EDX_NO_FAC_BLOCKS");
    int use_b = !(nob && nob[0] == '1') && mv <= (int64_t(1) << FAC_TYPE_SHIFT) && a->nvtab <= 255;
    auto *kron = (uint8_t *)dmalloc((size_t)mv, "factored flags");
    auto *vrow = (int64_t *)dmalloc((size_t)(mv + 1) * sizeof(int64_t), "factored H_v rowptr");
    auto *brow = (int64_t *)dmalloc((size_t)(mv + 1) * sizeof(int64_t), "factored block rowptr");
    auto *rrow = (int64_t *)dmalloc((size_t)(m + 1) * sizeof(int64_t), "factored R rowptr");
    auto *tset = (unsigned long long *)dmalloc(FAC_TYPE_SET * sizeof(uint64_t) + sizeof(int),
                                               "factored type set");
    int16_t *tid = nullptr;
    uint64_t *typ = nullptr;
    uint32_t *vpk = nullptr, *bpk = nullptr, *rpk = nullptr;
    int64_t vn = 0, bn = 0, rn = 0;
    int ntyp = 0;
    int rc = EDX_OK;
    int *overflow = reinterpret_cast<int *>(tset + FAC_TYPE_SET);
    std::vector<uint64_t> hset(FAC_TYPE_SET);
    int hover = 0;
    if (!kron || !vrow || !brow || !rrow || !tset) {
        rc = EDX_ERR_NOMEM;
        goto fail;
    }
    for (int pass = 0; pass < 2; ++pass) {
        if (hipMemsetAsync(vrow + mv, 0, sizeof(int64_t), stream()) != hipSuccess ||
            hipMemsetAsync(brow + mv, 0, sizeof(int64_t), stream()) != hipSuccess ||
            hipMemsetAsync(rrow + m, 0, sizeof(int64_t), stream()) != hipSuccess ||
            hipMemsetAsync(tset, 0xff, FAC_TYPE_SET * sizeof(uint64_t), stream()) != hipSuccess ||
            hipMemsetAsync(overflow, 0, sizeof(int), stream()) != hipSuccess) {
            rc = set_error(EDX_ERR_HIP, "factored form: memset failed");
            goto fail;
        }
        hipLaunchKernelGGL(k_fac_count, dim3(elem_blocks(mv)), dim3(256), 0, stream(), mv, nc,
                           a->d_off_rowptr, a->d_off_cols, a->d_off_idx, kron, vrow, brow, rrow,
                           use_b, tset, overflow);
        if (hipGetLastError() != hipSuccess) {
            rc = set_error(EDX_ERR_HIP, "factored form: count kernel failed");
            goto fail;
        }
        if (!use_b) break;
        if (edx::copy_sync(hset.data(), tset, FAC_TYPE_SET * sizeof(uint64_t), hipMemcpyDeviceToHost) !=
                hipSuccess ||
            edx::copy_sync(&hover, overflow, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = set_error(EDX_ERR_HIP, "factored form: type download failed");
            goto fail;
        }
        ntyp = 0;
        for (uint64_t k : hset) ntyp += k != FAC_EMPTY;
        if (!hover && ntyp <= FAC_MAX_TYPES) break;
        use_b = 0;   // too many types: count again with the slot-changing entries in R
        ntyp = 0;
    }
    if (use_b) {
        // type ids in increasing order of the type, whatever slots the inserts took
        std::vector<uint64_t> sorted;
        for (uint64_t k : hset)
            if (k != FAC_EMPTY) sorted.push_back(k);
        std::sort(sorted.begin(), sorted.end());
        std::vector<int16_t> hid(FAC_TYPE_SET, 0);
        for (int i = 0; i < FAC_TYPE_SET; ++i)
            if (hset[i] != FAC_EMPTY)
                hid[i] = (int16_t)(std::lower_bound(sorted.begin(), sorted.end(), hset[i]) -
                                   sorted.begin());
        tid = (int16_t *)dmalloc(FAC_TYPE_SET * sizeof(int16_t), "factored type ids");
        typ = (uint64_t *)dmalloc(std::max<size_t>(sorted.size(), 1) * sizeof(uint64_t),
                                  "factored types");
        if (!tid || !typ) {
            rc = EDX_ERR_NOMEM;
            goto fail;
        }
        if (edx::copy_sync(tid, hid.data(), FAC_TYPE_SET * sizeof(int16_t), hipMemcpyHostToDevice) !=
                hipSuccess ||
            (!sorted.empty() && edx::copy_sync(typ, sorted.data(), sorted.size() * sizeof(uint64_t),
                                          hipMemcpyHostToDevice) != hipSuccess)) {
            rc = set_error(EDX_ERR_HIP, "factored form: type upload failed");
            goto fail;
        }
    }
    if ((rc = exclusive_scan(vrow, mv + 1)) || (rc = exclusive_scan(brow, mv + 1)) ||
        (rc = exclusive_scan(rrow, m + 1)))
        goto fail;
    if (edx::copy_sync(&vn, vrow + mv, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
        edx::copy_sync(&bn, brow + mv, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
        edx::copy_sync(&rn, rrow + m, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess) {
        rc = set_error(EDX_ERR_HIP, "factored form: download failed");
        goto fail;
    }
    vpk = (uint32_t *)dmalloc((size_t)std::max<int64_t>(vn, 1) * sizeof(uint32_t), "factored H_v words");
    bpk = (uint32_t *)dmalloc((size_t)std::max<int64_t>(bn, 1) * sizeof(uint32_t), "factored block words");
    rpk = (uint32_t *)dmalloc((size_t)std::max<int64_t>(rn, 1) * sizeof(uint32_t), "factored R words");
    if (!vpk || !bpk || !rpk) {
        rc = EDX_ERR_NOMEM;
        goto fail;
    }
    hipLaunchKernelGGL(k_fac_fill, dim3(elem_blocks(mv)), dim3(256), 0, stream(), mv, nc,
                       a->d_off_rowptr, a->d_off_cols, a->d_off_idx, kron, vrow, vpk, brow, bpk,
                       rrow, rpk, use_b, tset, tid);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream()) != hipSuccess) {
        rc = set_error(EDX_ERR_HIP, "factored form: fill kernel failed");
        goto fail;
    }
    dfree(kron);
    dfree(tset);
    dfree(tid);
    a->d_fv_rowptr = vrow;
    a->d_fv_pk = vpk;
    a->fv_nnz = vn;
    a->d_fb_rowptr = brow;
    a->d_fb_pk = bpk;
    a->fb_nnz = bn;
    a->d_ftyp = typ;
    a->nftyp = ntyp;
    a->d_fr_rowptr = rrow;
    a->d_fr_pk = rpk;
    a->fr_nnz = rn;
    return EDX_OK;
fail:
    dfree(kron);
    dfree(tset);
    dfree(tid);
    dfree(typ);
    dfree(vrow);
    dfree(brow);
    dfree(rrow);
    dfree(vpk);
    dfree(bpk);
    dfree(rpk);
    return rc;
}

}  // namespace edx

extern "C" {

int edx_csr_factored_info(const edx_csr *a, int *nc_slots, int64_t *hv_nnz, int64_t *r_nnz) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    const bool on = a->d_fr_rowptr != nullptr;
    if (nc_slots) *nc_slots = on ? a->nc_slots : 0;
    if (hv_nnz) *hv_nnz = on ? a->fv_nnz : 0;
    if (r_nnz) *r_nnz = on ? a->fr_nnz : 0;
    return EDX_OK;
}

int edx_csr_factored_blocks(const edx_csr *a, int64_t *blocks, int *types) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    const bool on = a->d_fr_rowptr != nullptr;
    if (blocks) *blocks = on ? a->fb_nnz : 0;
    if (types) *types = on ? a->nftyp : 0;
    return EDX_OK;
}

}  // extern "C"
