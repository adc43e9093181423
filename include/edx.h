/*
 * edx.h -- C ABI of libedx, the MI355X (gfx950) exact-diagonalization core.
 *
 * Every entry point returns 0 on success and a nonzero edx_status code on
 * failure; edx_last_error() then holds a message for the calling thread.
 * Pointers named d_* are device pointers (HIP global memory, e.g. the
 * data_ptr() of a torch tensor on the current device); everything else is
 * host memory.  Complex numbers are (re, im) pairs of double (complex128) or
 * float (complex64), interleaved.  All indices crossing the ABI are 0-based.
 *
 * The drop-in layer (edx_*_fsolver) reproduces the f2py module the reference
 * builds from src/pyapi.f90: same three integer arguments, same input files in
 * the current directory, same output files.  The lower-level entry points
 * replace the Fortran routines the drivers call, and are what the parity
 * tests drive.
 */
#ifndef EDX_H
#define EDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum edx_status {
    EDX_OK = 0,
    EDX_ERR_ARG = 1,        /* invalid argument / shape */
    EDX_ERR_HIP = 2,        /* HIP runtime error (message has the HIP string) */
    EDX_ERR_IO = 3,         /* missing or malformed input file */
    EDX_ERR_NOMEM = 4,      /* device allocation failed / over budget */
    EDX_ERR_NUMERIC = 5,    /* eigensolver / breakdown failure */
    EDX_ERR_CAPACITY = 6    /* a row produced more candidates than supported */
};

/* value storage / arithmetic of products with a CSR matrix:
 *   EDX_C128  exact complex128 values, complex128 vectors, FP64 accumulation;
 *   EDX_C64   complex64 values and vectors, FP32 accumulation;
 *   EDX_C64D  complex64 off-diagonal values, exact complex128 diagonal kept
 *             apart, complex128 vectors, FP64 accumulation;
 *   EDX_C128I exact: the diagonal kept apart in complex128, each off-diagonal
 *             nonzero stored as the 16-bit number of its value in a table of
 *             the matrix's distinct off-diagonal values (a Hamiltonian built
 *             from a few hundred terms has a few dozen), complex128 vectors,
 *             FP64 accumulation.  Preparing it fails (EDX_ERR_CAPACITY) when
 *             the distinct off-diagonal values do not fit the 65536-slot
 *             table (in practice: more than a few tens of thousands).  With
 *             fewer than 2^24 columns and at most 256 values a packed copy
 *             holds column and value number in one 32-bit word per nonzero. */
enum edx_precision { EDX_C128 = 0, EDX_C64 = 1, EDX_C64D = 2, EDX_C128I = 3 };

/* Term records, layout-compatible with T_tensor2 / T_tensor4
 * (src/m_types.f90:7-20): 1-based spin-orbital indices, complex value. */
typedef struct { int32_t ind1, ind2; double re, im; } edx_tensor2;
typedef struct { int32_t ind1, ind2, ind3, ind4; double re, im; } edx_tensor4;

typedef struct edx_basis edx_basis;   /* sorted Fock words resident in HBM */
typedef struct edx_csr edx_csr;       /* CSR matrix resident in HBM */

/* ------------------------------------------------------------- runtime */

const char *edx_last_error(void);
/* Select the HIP device used by this thread's calls (default 0). */
int edx_init(int device);
/* Device-memory bytes currently held by library objects. */
int edx_mem_in_use(int64_t *bytes);
/* The stream this thread's library work is queued on: a hipStream_t, or NULL
 * for the legacy default stream (the initial choice of every thread).  All
 * launches and host-device copies of the thread's later calls go to it, and a
 * call that returns results to the host waits for that stream only, so two
 * threads on two streams overlap.  Use streams of the default priority.  No
 * reference counterpart (the reference runs one MPI rank per process). */
int edx_set_stream(void *stream);

/* --------------------------------------------------------- Fock basis */

/* All words of nbits bits with popcount nocc in ascending order, enumerated on
 * the device.  Replaces edrixs/fock_basis.py:398-451 (write_fock_dec_by_N) as the
 * source of fock_*.in contents. */
int edx_basis_combinations(int nbits, int nocc, edx_basis **out);
/* Upload a sorted basis as read from fock_*.in (src/io.f90:174-199, 308-361).
 * When it is the complete popcount set it is looked up by combinatorial rank,
 * otherwise by binary search (src/utils.f90:2-40). */
int edx_basis_from_words(const int64_t *words, int64_t n, int nbits, edx_basis **out);
int edx_basis_size(const edx_basis *b, int64_t *n, int *is_complete);
int edx_basis_download(const edx_basis *b, int64_t *words);
/* Index of each word (or -1): replaces binary_search (src/utils.f90:2-40). */
int edx_basis_lookup(const edx_basis *b, const int64_t *words, int64_t n, int64_t *index);
void edx_basis_free(edx_basis *b);

/* ------------------------------------------------------------ CSR build */

/* Index layout of the core-hole space (core_slot, valence) of H_n and of the
 * core-hole side of the transition operators:
 *   EDX_COREHOLE_REFERENCE   core_slot * dim(fock_n) + valence, the reference's
 *                            (src/ham.f90:419-446);
 *   EDX_COREHOLE_INTERLEAVED valence * num_core_orbs + core_slot: the matrix is
 *                            P H P^T with the same value at every permuted
 *                            entry, and the terms that move the core hole (core
 *                            spin-orbit, core-valence exchange) land next to the
 *                            diagonal instead of dim(fock_n) rows away.
 * Only the internal order of core-hole-space vectors changes: spectra and pole
 * files from a consistent set of matrices differ by rounding alone.
 * edx_set_corehole_layout sets the layout of the following edx_build_ham_n,
 * edx_build_transop_i and edx_build_transop_f calls on the calling thread
 * (default EDX_COREHOLE_REFERENCE) and returns the previous one.  No reference
 * counterpart. */
enum edx_corehole_layout { EDX_COREHOLE_REFERENCE = 0, EDX_COREHOLE_INTERLEAVED = 1 };
int edx_set_corehole_layout(int layout, int *previous);

/* H = sum hop + sum coul on one space; rows/columns = basis
 * (src/ham.f90:2-349 build_ham_i).  ham_sign must be +1 or -1; with -1 the
 * result is omega*I - H, omega added last on the diagonal (src/ham.f90:156-176). */
int edx_build_ham_i(const edx_basis *fock, const edx_tensor2 *hop, int64_t nhopp,
                    const edx_tensor4 *coul, int64_t ncoul, double omega_re,
                    double omega_im, double ham_sign, int precision, edx_csr **out);
/* Core-hole space H_n, rows (k, v) -> k*dim(fock_n)+v (src/ham.f90:352-711). */
int edx_build_ham_n(const edx_basis *fock_n, int num_val_orbs, int num_core_orbs,
                    const edx_tensor2 *hop, int64_t nhopp, const edx_tensor4 *coul,
                    int64_t ncoul, double omega_re, double omega_im, double ham_sign,
                    int precision, edx_csr **out);
/* Absorption operator: rows = core-hole space over fock_n, columns = fock_i
 * (src/ham.f90:714-912 build_transop_i). */
int edx_build_transop_i(const edx_basis *fock_n, const edx_basis *fock_i, int num_val_orbs,
                        int num_core_orbs, const edx_tensor2 *trans, int64_t ntran,
                        int precision, edx_csr **out);
/* Emission operator: rows = fock_f with full core, columns = core-hole space
 * over fock_n (src/ham.f90:915-1115 build_transop_f). */
int edx_build_transop_f(const edx_basis *fock_f, const edx_basis *fock_n, int num_val_orbs,
                        int num_core_orbs, const edx_tensor2 *trans, int64_t ntran,
                        int precision, edx_csr **out);
int edx_csr_info(const edx_csr *a, int64_t *m, int64_t *n, int64_t *nnz, int *precision);
/* Copy to host: rowptr[m+1], cols[nnz], vals[2*nnz] complex128 (any may be NULL). */
int edx_csr_download(const edx_csr *a, int64_t *rowptr, int32_t *cols, void *vals);
/* Copy the stored values of one precision: 2*nnz doubles (EDX_C128) or floats (EDX_C64). */
int edx_csr_download_values(const edx_csr *a, int precision, void *vals);
/* Add a complex64 value copy (rounded from complex128) for the mixed path. */
int edx_csr_add_c64(edx_csr *a);
/* Make the storage a precision needs resident (EDX_C64 copy, EDX_C64D split). */
int edx_csr_prepare(edx_csr *a, int precision);
void edx_csr_free(edx_csr *a);

/* ------------------------------------------------------------------ SpMV */

/* d_y += A d_x, complex128 vectors (src/spmv.f90:73-103 matvec_csr).
 * precision selects the stored values used: EDX_C128 (FP64 accumulate) or
 * EDX_C64 (complex64 values and x, FP32 accumulation, result added in FP64). */
int edx_spmv(const edx_csr *a, int precision, const void *d_x, void *d_y);
/* Multi-RHS: d_Y[:, k] += A d_X[:, k] for k < nrhs, column-major, leading
 * dimensions m and n.  Same precision rules as edx_spmv. */
int edx_spmm(const edx_csr *a, int precision, int nrhs, const void *d_X, void *d_Y);

/* Lanes per row of the single-vector SpMV of one precision: set = 4..64
 * fixes it, set = 0 only reads it; *lanes = 0 means the mean-row-length rule
 * (matrices under 131072 rows, or before the first edx_csr_prepare, which
 * times the choices on larger matrices).  Codes above 64 name other
 * schedules: 143, 144, 184 (pipelined), 145 (pipelined, packed indexed form
 * only), 206 (slot-factored form, EDX_C128I only).  No reference counterpart. */
int edx_csr_spmv_lanes(edx_csr *a, int precision, int set, int *lanes);

/* The slot-factored form of a square core-hole matrix in the interleaved
 * layout, made by edx_csr_prepare(EDX_C128I) where the matrix qualifies (6
 * core slots, fewer than 2^24 rows, at most 256 distinct off-diagonal
 * values): the off-diagonal entries that are the same for every core slot of
 * a valence row (terms on valence orbitals only, src/ham.f90:419-446) stored
 * once per valence row, the entries that move the core hole as typed blocks
 * (edx_csr_factored_blocks), and the valence rows whose slot rows differ
 * whole per row (R).  *nc_slots = 0 when the form is absent; *hv_nnz counts
 * the entries stored once per valence row, *r_nnz those stored per row.  No reference
 * counterpart (a storage form of the matrix of src/ham.f90:352-711). */
int edx_csr_factored_info(const edx_csr *a, int *nc_slots, int64_t *hv_nnz, int64_t *r_nnz);

/* The typed blocks of the slot-factored form: for each valence row whose
 * slot rows share their slot-keeping entries, its entries that move the core
 * hole grouped by (valence column, slot shift), one 32-bit word per block
 * (valence column | type << 22); a type is the shift and the value number of
 * each slot's entry.  Made with the form where there are fewer than 2^22
 * valence rows, at most 255 distinct values and at most 1024 types
 * (EDX_NO_FAC_BLOCKS=1 keeps those entries per row instead).  *blocks counts
 * the block words, *types the distinct types; both 0 without blocks.  No
 * reference counterpart (a storage form of the matrix of src/ham.f90:352-711). */
int edx_csr_factored_blocks(const edx_csr *a, int64_t *blocks, int *types);

/* Live profiling of SpMV/SpMM launches (no reference counterpart): while
 * enabled, every launch is bracketed by HIP events on its stream and its
 * algorithmic HBM bytes
 *   nnz*(s_v + s_i) + (m+1)*s_p + K*(n*s_x + m*s_y*(reads + writes of y-like vectors))
 * are recorded; a ring of event pairs is folded into running totals as it
 * wraps, so every launch of a profiled region counts, however many there are.
 * edx_profile(1) clears and starts, edx_profile(0) stops;
 * edx_profile_read synchronises and returns the launch count, summed kernel
 * milliseconds and summed bytes of the launches in one precision
 * (edx_precision) and one kernel (edx_kernel, or -1 for all);
 * edx_profile_read_rhs the same for the launches with nrhs right-hand sides
 * (-1: all; any count above 16 reads every launch of more than 16);
 * edx_profile_read_alg the summed algorithmic bytes of those launches counted
 * as the complex128 CSR of the same operator, whatever form they read:
 *   nnz*(16 + 4) + (m+1)*8 + K*(n*16 + m*16*(1 + acc)),
 * acc = 1 where the epilogue accumulates into y. */
enum edx_kernel {
    EDX_KERNEL_SPMV = 0,      /* k_spmv: one vector */
    EDX_KERNEL_SPMM_IL = 1,   /* k_spmm_il: K < 8 interleaved right-hand sides */
    EDX_KERNEL_SPMM_RL = 2,   /* k_spmm_rl: K >= 8, RHS-parallel lanes */
    EDX_KERNEL_SPMM_CM = 3,   /* k_spmm: column-major blocks (edx_spmm) */
    EDX_KERNEL_MF = 4,        /* k_mf: matrix-free apply (recorded as precision EDX_C128) */
    EDX_KERNEL_BPROJ = 5,     /* k_bproj: block projection V^H W on the matrix cores (EDX_C128) */
    EDX_KERNEL_BUPD = 6,      /* k_bupd: block update W - V C on the matrix cores (EDX_C128) */
    EDX_KERNEL_SPMV_FAC = 7,  /* k_spmv_fac: one vector, slot-factored form (EDX_C128I) */
    EDX_KERNEL_SPMM_FAC = 8   /* k_spmm_fac: 2-4 interleaved vectors, slot-factored form */
};
int edx_profile(int enable);
int edx_profile_read(int precision, int kernel, int64_t *launches, double *total_ms,
                     double *total_bytes);
int edx_profile_read_rhs(int precision, int kernel, int nrhs, int64_t *launches,
                         double *total_ms, double *total_bytes);
int edx_profile_read_alg(int precision, int kernel, int nrhs, double *alg_bytes);
/* Summed algorithmic floating-point operations of the same launches (nonzero
 * for the matrix-core kernels only: 512 real flops per row and basis block). */
int edx_profile_read_flops(int precision, int kernel, double *total_flops);

/* ------------------------------------------------- matrix-free operators */

/* H applied on the fly, without storing it (no reference counterpart; the
 * semantics are those of build_ham_i / build_ham_n, src/ham.f90:2-711, followed
 * by matvec_csr, src/spmv.f90:73-103): every term is tested, signed and ranked
 * exactly as the CSR build does, so columns and signs are the build's; only
 * the order in which duplicate terms are summed differs.  The basis must
 * outlive the operator.  Complex128 vectors, FP64 accumulation. */
typedef struct edx_mfop edx_mfop;
int edx_mf_ham_i(const edx_basis *fock, const edx_tensor2 *hop, int64_t nhopp,
                 const edx_tensor4 *coul, int64_t ncoul, double omega_re, double omega_im,
                 double ham_sign, edx_mfop **out);
int edx_mf_ham_n(const edx_basis *fock_n, int num_val_orbs, int num_core_orbs,
                 const edx_tensor2 *hop, int64_t nhopp, const edx_tensor4 *coul, int64_t ncoul,
                 double omega_re, double omega_im, double ham_sign, edx_mfop **out);
int edx_mf_size(const edx_mfop *op, int64_t *n);
/* d_y += H d_x */
int edx_mf_apply(const edx_mfop *op, const void *d_x, void *d_y);
/* build_krylov_mp (src/lanczos.f90:210-303) over the matrix-free operator; same
 * outputs as edx_build_krylov. */
int edx_build_krylov_mf(const edx_mfop *op, const void *d_v0, int maxn, int *neff, double *alpha,
                        double *beta, double *norm);
void edx_mf_free(edx_mfop *op);

/* --------------------------------------------------------------- solvers */

/* Lowest nev eigenpairs of a Hermitian CSR (replaces diag_ham_lanczos,
 * src/lanczos.f90:6-167).  FP64 Lanczos with full reorthogonalization and a
 * seeded start vector; degenerate copies are recovered by deflated restarts
 * when deflate != 0.  evals[nev]; d_evecs is n x nvec complex128 column-major
 * (may be NULL when nvec == 0).  niter_out: Krylov steps used (all restarts). */
int edx_lanczos_eigs(const edx_csr *h, int precision, int nev, int maxiter, double tol,
                     uint64_t seed, int deflate, int nvec, double *evals, void *d_evecs,
                     int *niter_out);
/* Lowest nev eigenpairs by the reference's two-pass procedure
 * (diag_ham_lanczos, src/lanczos.f90:6-167): the three-term recurrence with
 * no reorthogonalization from the same seeded start vector as
 * edx_lanczos_eigs, tested every 10 steps (converged when the nev lowest
 * Ritz residual estimates |beta_m y_m| are below tol), then run again from
 * the same start to accumulate the Ritz vectors.  Keeps three vectors of n
 * instead of the basis; finds one copy of a degenerate level (use
 * edx_lanczos_eigs with deflate for multiplets).  evals[nev]; d_evecs n x nvec
 * complex128 column-major (NULL when nvec == 0); niter_out: steps of pass 1. */
int edx_lanczos_ground(const edx_csr *h, int precision, int nev, int maxiter, double tol,
                       uint64_t seed, int nvec, double *evals, void *d_evecs, int *niter_out);
/* Same solver started from the normalised sum of nstart given vectors
 * (d_start, n x nstart column-major) instead of a random vector, without
 * deflated restarts: the FP64 refinement that finishes a ground state found
 * with complex64 products (no counterpart in the reference). */
int edx_lanczos_refine(const edx_csr *h, int precision, int nev, int maxiter, double tol,
                       int nstart, const void *d_start, int nvec, double *evals, void *d_evecs,
                       int *niter_out);
/* Lowest nev eigenpairs by thick-restart Lanczos (replaces diag_ham_arpack,
 * src/arpack.f90:2-112, which='SR', the ed_solver=2 path): a basis of
 * m = max(ncv, 2 nev + 16) fully reorthogonalized vectors, restarted
 * at most maxiter times, each restart keeping the wanted Ritz vectors and half
 * of the others (Wu & Simon 2000).  Converged when every wanted Ritz residual
 * |beta_m y_m| <= tol * max(eps^(2/3), |theta|), ARPACK's test.  deflate != 0
 * adds restarts from vectors orthogonal to the pairs found, for degenerate
 * levels.  evals[nev]; d_evecs n x nvec complex128 column-major (NULL when
 * nvec == 0); nmatvec_out: products with H (ARPACK's iparam(3) counts
 * restarts instead); nconv_out: converged values of the first run (iparam(5)).
 * Keeps about nev + 2m vectors of 16 n bytes. */
int edx_trlan_eigs(const edx_csr *h, int precision, int nev, int ncv, int maxiter, double tol,
                   uint64_t seed, int deflate, int nvec, double *evals, void *d_evecs,
                   int *nmatvec_out, int *nconv_out);
/* Lowest nev eigenpairs by block Lanczos (replaces diag_ham_lanczos,
 * src/lanczos.f90:6-167, for several states in one Krylov space): blocks of
 * 8 vectors, one 8-RHS SpMM per step, full reorthogonalization of each block
 * against the whole basis on the FP64 matrix cores (block classical
 * Gram-Schmidt twice, shifted Cholesky QR), Ritz pairs of the block-tridiagonal
 * projection.  Stops when the nev lowest Ritz residuals |B s_last| are below
 * tol, after maxblocks steps, or when the basis spans the space.  Needs
 * n >= 16.  evals[nev]; d_evecs n x nvec complex128 column-major (NULL when
 * nvec == 0); nblocks_out: block steps taken; resid_out[nev] (may be NULL):
 * final residual estimates.  Keeps maxblocks + 1 blocks of 128 n bytes. */
int edx_block_lanczos_eigs(const edx_csr *h, int precision, int nev, int maxblocks, double tol,
                           uint64_t seed, int nvec, double *evals, void *d_evecs,
                           int *nblocks_out, double *resid_out);
/* Krylov tridiagonalization from an unnormalised start vector (replaces
 * build_krylov_mp, src/lanczos.f90:210-303): alpha[maxn], beta[maxn]
 * (beta_1..beta_neff as written to the pole files), norm = <v|v>. */
int edx_build_krylov(const edx_csr *h, int precision, const void *d_v0, int maxn, int *neff,
                     double *alpha, double *beta, double *norm);
/* edx_build_krylov without device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * the same results bit for bit; its vectors, partials and history live in a
 * static 64 MiB device array per GPU (held until the process ends), which
 * bounds a call to about 2 n + 2 maxn <= 4e6 complex128 values (requests
 * beyond it fail with EDX_ERR_CAPACITY).  The precision's form must already be
 * prepared (edx_csr_prepare); EDX_C64 is not offered.  Calls that overlap on
 * one GPU share the array and race. */
int edx_build_krylov_noalloc(const edx_csr *h, int precision, const void *d_v0, int maxn,
                             int *neff, double *alpha, double *beta, double *norm);
/* The same tridiagonalization with every new vector reorthogonalized against
 * all previous ones (two classical Gram-Schmidt passes): the exact Krylov
 * projection, against which the rounding of the plain recurrence is measured
 * (no counterpart in the reference; keeps maxn+1 vectors of length n). */
int edx_build_krylov_reorth(const edx_csr *h, int precision, const void *d_v0, int maxn,
                            int *neff, double *alpha, double *beta, double *norm);
/* MINRES for A x = b (replaces pminres_csr, src/linsys.f90:3-128); stops when
 * the recurrence residual estimate <= tol or after maxit steps.  d_x receives
 * x; info[0] = iterations, final_res[0] = ||A x - b||. */
int edx_minres(const edx_csr *a, int precision, const void *d_b, void *d_x, double tol,
               int maxit, int *iters, double *final_res);

/* Batched Krylov: nrhs independent build_krylov_mp recurrences in lockstep
 * over one matrix (interleaved multi-RHS products).  d_V0 is n x nrhs
 * column-major; alpha and beta are nrhs x maxn row-major (beta_1..beta_neff per
 * RHS as in the pole files), neff[nrhs], norm[nrhs].  Each RHS gets the same
 * numbers it would get alone. */
int edx_build_krylov_batch(const edx_csr *h, int precision, int nrhs, const void *d_V0, int maxn,
                           int *neff, double *alpha, double *beta, double *norm);
/* Batched MINRES for (sigma_k I - H) x_k = b_k, k < nrhs, with the recurrence
 * of pminres_csr (src/linsys.f90:3-128) per RHS; sigma is nrhs complex
 * (re, im) pairs, d_B and d_X are n x nrhs column-major.  Per RHS: iters[k],
 * final_res[k] = |(sigma_k - H) x_k - b_k|.  This replaces the reference's
 * one-fsolver-call-per-(omega_in, polarization) loop with one shifted solve of
 * the unshifted H_n (edrixs/solvers.py:2651-2693, src/rixs_driver.f90:130-150). */
/* The same batched recurrence, resumable: start normalizes the nrhs columns
 * of d_V0 (n x nrhs column-major; copied, so the caller's buffer may go) and
 * keeps vectors, flags and the (alpha, beta) history on the device; advance
 * takes up to nsteps more steps of every recurrence that has not broken down
 * (steps_done = total steps so far); read returns the first count steps in
 * the layout of edx_build_krylov_batch (zeros past a recurrence's end).  A
 * caller stops once the continued fractions it evaluates have converged,
 * instead of fixing nkryl up front as build_krylov_mp does
 * (src/lanczos.f90:266-298).  h must outlive the state. */
typedef struct edx_krylov_batch edx_krylov_batch;
int edx_krylov_batch_start(const edx_csr *h, int precision, int nrhs, const void *d_V0,
                           int capacity, edx_krylov_batch **out);
int edx_krylov_batch_advance(edx_krylov_batch *state, int nsteps, int *steps_done);
int edx_krylov_batch_read(const edx_krylov_batch *state, int count, int *neff, double *alpha,
                          double *beta, double *norm);
/* Keep the recurrences keep[0..nkeep) (indices into the current columns,
 * strictly increasing) and drop the others, so that a caller whose spectra
 * of some columns have converged stops paying for their products; the kept
 * recurrences continue unchanged and become columns 0..nkeep-1 (no reference
 * counterpart: build_krylov_mp, src/lanczos.f90:266-298, runs one vector). */
int edx_krylov_batch_keep(edx_krylov_batch *state, int nkeep, const int *keep);
void edx_krylov_batch_free(edx_krylov_batch *state);

/* Grouped recurrences for a serving loop (no reference counterpart): many
 * independent single-vector build_krylov_mp recurrences (src/lanczos.f90:
 * 266-298), each over its own square matrix, advanced together so that one
 * step of every live member is one SpMV launch plus four small launches.
 * Precision EDX_C128, EDX_C64D or EDX_C128I for every member; each member keeps up to
 * max_steps (alpha, beta).  add copies and normalizes d_v0 (length rows(h))
 * and returns a member id; h must stay alive while it is a member.  advance
 * takes nsteps steps of every member still running (live = members still
 * running).  read returns a member's steps so far, alpha_1..steps,
 * beta_1..steps (pole-file convention), <v0|v0> and whether it is still
 * running.  A member's numbers do not depend on the other members. */
typedef struct edx_group edx_group;
int edx_group_create(int precision, int max_steps, edx_group **out);
int edx_group_add(edx_group *g, const edx_csr *h, const void *d_v0, int *member);
int edx_group_remove(edx_group *g, int member);
int edx_group_advance(edx_group *g, int nsteps, int *live);
int edx_group_read(const edx_group *g, int member, int *steps, double *alpha, double *beta,
                   double *norm, int *running);
void edx_group_free(edx_group *g);

/* Multi-shift Galerkin solve of (sigma_k I - H) x_k = b for nshift shifts from
 * ONE Lanczos recurrence on H, keeping only projections: for each of nproj
 * operators P_p (a CSR with H's row count as its column count, or NULL for the
 * identity) d_out[p] receives P_p x_k for every k, rows(P_p) x nshift
 * column-major complex128.  Shift k stops at the first iterate whose residual
 * |b - (sigma_k - H) x_k| (the recurrence's own value beta_m |e_m^T y_m|) is
 * <= tol |b|; iters[k] and final_res[k] (relative to |b|) describe that iterate.
 * Every sigma_k needs a nonzero imaginary part (the core-hole width); nshift <=
 * 256.  This replaces the reference's per-incident-energy rebuild of
 * omega - H_n and pminres_csr solve (src/rixs_driver.f90:127-150,
 * src/linsys.f90:3-128) followed by the emission operator applied to each
 * solution (src/rixs_driver.f90:163-185): one recurrence serves all incident
 * energies of a polarization (edrixs/solvers.py:2651-2693 loops over them). */
int edx_shifted_fom_project(const edx_csr *h, int precision, const void *d_b, int nshift,
                            const double *sigma, int nproj, const edx_csr *const *proj,
                            void *const *d_out, double tol, int maxit, int *iters,
                            double *final_res);
/* The same for nrhs right-hand sides at once (column r of d_B, n x nrhs
 * column-major; nrhs <= 64), all with the same nshift shifts, in one
 * interleaved recurrence whose product is the multi-RHS SpMM: proj[r] (NULL:
 * the identity) projects right-hand side r into d_out[r] (rows(proj[r]) x
 * nshift column-major); iters and final_res are nrhs x nshift, row r for
 * right-hand side r.  Each right-hand side gets the solution of its own
 * edx_shifted_fom_project call up to rounding (the reference's RIXS driver
 * solves the incident polarizations one after another, src/rixs_driver.f90:
 * 127-150). */
int edx_shifted_fom_project_batch(const edx_csr *h, int precision, int nrhs, const void *d_B,
                                  int nshift, const double *sigma, const edx_csr *const *proj,
                                  void *const *d_out, double tol, int maxit, int *iters,
                                  double *final_res);
int edx_minres_shifted_batch(const edx_csr *h, int precision, int nrhs, const double *sigma,
                             const void *d_B, void *d_X, double tol, int maxit, int *iters,
                             double *final_res);

/* Continued-fraction spectra on the device (edrixs/plot_spectrum.py:9-61):
 * spec[w] += weight * (-1/pi) Im[norm / (z - a_1 - b_1^2/(z - a_2 - ...))],
 * z = omega[w] + i*gamma[w] + e_gs; host arrays in and out. */
int edx_cf_spectrum(int neff, const double *alpha, const double *beta, double norm,
                    double e_gs, double weight, int nw, const double *omega,
                    const double *gamma, double *spec);

/* Density matrix <v|c+_i c_j|v> of a device vector over fock
 * (src/ed_driver.f90:219-248); denmat is norb*norb complex128 row-major. */
int edx_denmat(const edx_basis *fock, int num_val_orbs, const void *d_vec, double *denmat);

/* ----------------------------------------------------- drop-in drivers */

/* The four f2py entry points of edrixs.fedrixs (src/pyapi.f90:1-56, 58-118,
 * 120-182, 184-240): read config.in and the term / fock files in the current
 * directory, run on the device, write eigvals.dat, denmat.dat, eigvec.N,
 * xas_poles.N, rixs_poles.N, opavg.dat with the reference formats.  comm,
 * my_id and num_procs have the reference meaning; ranks other than 0 return
 * at once (one GPU computes for the communicator). */
int edx_ed_fsolver(int comm, int my_id, int num_procs);
int edx_xas_fsolver(int comm, int my_id, int num_procs);
int edx_rixs_fsolver(int comm, int my_id, int num_procs);
int edx_opavg_fsolver(int comm, int my_id, int num_procs);

#ifdef __cplusplus
}
#endif
#endif /* EDX_H */
