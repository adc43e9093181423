// Launch-shape autotuner for the single-vector SpMV (part of edx_kernels.hip).
//
// The best number of lanes per row depends on the row-length distribution and
// on how well the gathers of x hit the caches, which the mean row length does
// not predict (H_n at norb=24: 24 nonzeros per row, yet 32 lanes per row beat
// 8 because a row's loads are then all in flight at once).  So each matrix is
// timed once per precision on itself: every supported G runs a warm-up and
// three timed launches on scratch vectors, and the fastest is kept in the
// matrix (spmv_lanes / spmv_lanes32 / spmv_lanes64d), which every later
// single-vector product of that matrix uses.  The timed launches carry the
// epilogue of a Krylov step (addend and dot partials): without it the ranking
// of the choices differs from the one inside the solvers.  Matrices below TUNE_MIN_ROWS
// keep the mean-row-length rule: their products are too short to time and too
// cheap to matter.  EDX_NO_TUNE=1 in the environment turns tuning off.

namespace {

constexpr int64_t TUNE_MIN_ROWS = 1 << 17;
constexpr int TUNE_REPS = 5;

int *lanes_slot(edx_csr *a, int precision) {
    if (precision == EDX_C64) return &a->spmv_lanes32;
    if (precision == EDX_C64D) return &a->spmv_lanes64d;
    if (precision == EDX_C128I) return &a->spmv_lanes128i;
    return &a->spmv_lanes;
}

}  // namespace

namespace edx {

int tune_spmv(edx_csr *a, int precision) {
    int *slot = lanes_slot(a, precision);
    if (*slot) return EDX_OK;
    const char *off = getenv("EDX_NO_TUNE");
    if (a->m < TUNE_MIN_ROWS || (off && off[0] == '1')) return EDX_OK;
    const size_t sx = precision == EDX_C64 ? sizeof(float2) : sizeof(double2);
    void *x = dmalloc((size_t)a->n * sx, "tune x");
    void *y = dmalloc((size_t)a->m * sizeof(double2), "tune y");
    void *d = dmalloc((size_t)a->m * sizeof(double2), "tune d");
    void *part = dmalloc((size_t)spmv_partials_len(a) * sizeof(double2), "tune partials");
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = EDX_OK;
    int best = 0;
    float best_ms = 0.f;
    bool prof_saved = g_prof;
    if (!x || !y || !d || !part) {
        rc = EDX_ERR_NOMEM;
        goto done;
    }
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
        rc = set_error(EDX_ERR_HIP, "tune: event create failed");
        goto done;
    }
    if (hipMemsetAsync(x, 0, (size_t)a->n * sx, stream()) != hipSuccess ||
        hipMemsetAsync(y, 0, (size_t)a->m * sizeof(double2), stream()) != hipSuccess ||
        hipMemsetAsync(d, 0, (size_t)a->m * sizeof(double2), stream()) != hipSuccess) {
        rc = set_error(EDX_ERR_HIP, "tune: memset failed");
        goto done;
    }
    prof_saved = g_prof;   // tuning launches stay out of the live profile
    g_prof = false;
    for (int G : {4, 8, 16, 32, 64}) {
        // the epilogue of a Krylov step: y = A x + z, with the dot <d|y>
        SpmvEpi e{};
        e.y = (double2 *)y;
        e.z = (const double2 *)y;
        e.dvec = (const double2 *)d;
        e.partials = (double2 *)part;
        int nb = 0;
        if ((rc = spmv_launch(a, precision, x, e, G, &nb)) != EDX_OK) goto done;
        (void)hipEventRecord(ev[0], stream());
        for (int r = 0; r < TUNE_REPS; ++r)
            if ((rc = spmv_launch(a, precision, x, e, G, &nb)) != EDX_OK) goto done;
        (void)hipEventRecord(ev[1], stream());
        if (hipEventSynchronize(ev[1]) != hipSuccess) {
            rc = set_error(EDX_ERR_HIP, "tune: synchronize failed");
            goto done;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (!best || ms < best_ms) {
            best = G;
            best_ms = ms;
        }
    }
    *slot = best;
done:
    g_prof = prof_saved;
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    dfree(x);
    dfree(y);
    dfree(d);
    dfree(part);
    return rc;
}

}  // namespace edx

extern "C" {

int edx_csr_spmv_lanes(edx_csr *a, int precision, int set, int *lanes) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    EDX_CHECK(precision >= EDX_C128 && precision <= EDX_C128I, EDX_ERR_ARG, "bad precision %d",
              precision);
    EDX_CHECK(set == 0 || set == 4 || set == 8 || set == 16 || set == 32 || set == 64, EDX_ERR_ARG,
              "lanes per row must be 4, 8, 16, 32 or 64 (got %d)", set);
    int *slot = lanes_slot(a, precision);
    if (set) *slot = set;
    if (lanes) *lanes = *slot;
    return EDX_OK;
}

}  // extern "C"
