// Internal runtime pieces shared by the libedx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "../../include/edx.h"

namespace edx {

int set_error(int code, const char *fmt, ...);
// layout of core-hole-space indices for the next builds on this thread
// (edx_set_corehole_layout)
int corehole_layout();
int hip_fail(hipError_t e, const char *what, const char *file, int line);

// Device allocation with accounting; returns nullptr (error set) on failure.
void *dmalloc(size_t bytes, const char *what);
void dfree(void *p);
int64_t mem_in_use();

// The stream all library work of the calling thread is queued on: by default
// the legacy default stream, so it is ordered with torch's default stream; a
// thread may choose another (edx_set_stream), and while a solver captures its
// iterations into a HIP graph the launches of this thread go to the capture
// stream (StreamScope).  Host-device copies and fills go through copy_sync /
// fill_sync, ordered on the same stream, so the library's results are complete
// when they return whichever stream the thread uses.
extern thread_local hipStream_t g_launch_stream;
inline hipStream_t stream() { return g_launch_stream; }
inline hipError_t copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, g_launch_stream);
    return e == hipSuccess ? hipStreamSynchronize(g_launch_stream) : e;
}
inline hipError_t fill_sync(void *dst, int value, size_t bytes) {
    const hipError_t e = hipMemsetAsync(dst, value, bytes, g_launch_stream);
    return e == hipSuccess ? hipStreamSynchronize(g_launch_stream) : e;
}
struct StreamScope {
    hipStream_t saved;
    explicit StreamScope(hipStream_t s) : saved(g_launch_stream) { g_launch_stream = s; }
    ~StreamScope() { g_launch_stream = saved; }
};
int current_device();
int num_cus();

}  // namespace edx

#define EDX_HIP(call)                                                            \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) return ::edx::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define EDX_CHECK(cond, code, ...)                                               \
    do {                                                                         \
        if (!(cond)) return ::edx::set_error((code), __VA_ARGS__);               \
    } while (0)

#define EDX_TRY(call)                                                            \
    do {                                                                         \
        int r_ = (call);                                                         \
        if (r_) return r_;                                                       \
    } while (0)

#define EDX_LAUNCH_CHECK() EDX_HIP(hipGetLastError())

// Sorted Fock words resident on the device.
struct edx_basis {
    int64_t n;          // number of words
    int nbits;          // valence orbital count the words live in
    int nocc;           // popcount shared by every word (when complete)
    int complete;       // 1: words == all nbits-bit words with popcount nocc, ascending
    int64_t *d_words;   // device, ascending
};

// CSR matrix resident on the device.  Rows are 0-based, columns are 0-based
// global indices.  d_vals is always present (complex128, the exact sums of the
// reference builder); d_vals32 is an optional complex64 rounding of it.
struct edx_csr {
    int64_t m, n, nnz;
    int64_t *d_rowptr;
    int32_t *d_cols;
    double2 *d_vals;
    float2 *d_vals32;
    int spmv_lanes;      // lanes per row chosen for SpMV (0 = not tuned)
    int spmv_lanes32;
    int spmv_lanes64d;
    int spmv_lanes128i;
    int spmm_lanes;
    // split form (EDX_C64D): exact diagonal + complex64 off-diagonal CSR
    double2 *d_diag;
    int64_t *d_off_rowptr;
    int32_t *d_off_cols;
    float2 *d_off_vals32;
    int64_t off_nnz;
    // indexed form (EDX_C128I): the split form's off-diagonal pattern with, per
    // nonzero, the number of its exact value in a table of distinct values
    uint16_t *d_off_idx;
    double2 *d_vtab;     // nvtab entries: value number k -> its exact value
    int64_t nvtab;       // distinct off-diagonal values
    // packed indexed form (n < 2^24, nvtab <= 256): column | value number << 24
    uint32_t *d_off_pk;
    // index layout of the core-hole space on the rows and/or columns that are
    // in it (edx_corehole_layout); 0 for matrices without one
    int corehole_layout;
    // core slots of a square core-hole matrix in the interleaved layout (row
    // v * nc_slots + s), 0 otherwise
    int nc_slots;
    // slot-factored form (factor.hip): the off-diagonal entries shared by all
    // slots of a valence row once per valence row (H_v: valence column | value
    // number << 24), the rest per row (R: column | value number << 24)
    int64_t *d_fv_rowptr;   // m / nc_slots + 1
    uint32_t *d_fv_pk;
    int64_t fv_nnz;
    int64_t *d_fr_rowptr;   // m + 1
    uint32_t *d_fr_pk;
    int64_t fr_nnz;
    // typed blocks of the factored form: per valence row whose slot rows share
    // their slot-keeping entries, its slot-changing entries grouped by
    // (valence column, slot shift), one word each (valence column | type << 22);
    // a type is the shift and each slot's value number (factor.hip)
    int64_t *d_fb_rowptr;   // m / nc_slots + 1
    uint32_t *d_fb_pk;
    int64_t fb_nnz;
    uint64_t *d_ftyp;       // nftyp types
    int nftyp;
    // packed form of every entry (projections of the multi-shift solve):
    // column | value number << 24, with the table of distinct values
    uint32_t *d_apk;
    double2 *d_atab;
    int64_t natab;
};

// Lookup descriptor for a column space basis.
struct BasisView {
    const int64_t *words;
    int64_t n;
    int nbits;
    int nocc;
    int complete;
};

inline BasisView view_of(const edx_basis *b) {
    return BasisView{b->d_words, b->n, b->nbits, b->nocc, b->complete};
}

namespace edx {
// basis.hip
int ensure_binomials();
}  // namespace edx
