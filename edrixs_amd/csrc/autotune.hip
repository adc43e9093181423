// Launch-shape autotuner (part of edx_kernels.hip).
//
// Single-vector SpMV: the best number of lanes per row depends on the
// row-length distribution and on how well the gathers of x hit the caches,
// which the mean row length does not predict.  Multi-RHS products: the
// schedule (row-split with 4-RHS tiles and 8 or 16 lanes per row, or
// RHS-parallel with tiles of 16 to 128 right-hand sides) trades reading the
// matrix once per tile against keeping each tile's slice of the vector block
// in the last-level cache, which depends on the vector length and K.  So each
// choice is timed on the matrix itself -- a warm-up and TUNE_REPS timed
// launches on scratch vectors, with the epilogue of a Krylov step (addend and
// dot partials), since without it the ranking differs from the one inside the
// solvers -- and the fastest is kept.  Choices are cached by matrix shape
// (rows, columns, nonzeros), precision, K and the storage forms present (the
// packed indexed and slot-factored forms add candidates), so a matrix rebuilt
// with the same structure reuses them.  Matrices below TUNE_MIN_ROWS keep the
// rules; EDX_NO_TUNE=1 in the environment turns tuning off.
//
// The schedule fixes the order of every sum (lanes per row, dot partials per
// workgroup), so two processes that picked different schedules for the same
// matrix would round differently.  Choices therefore persist in a text file,
// one line per choice keyed also by the device (architecture and CU count):
// $EDX_TUNE_CACHE, else $XDG_CACHE_HOME/edrixs_amd/tune-v2.txt, else
// $HOME/.cache/edrixs_amd/tune-v2.txt.  A process reads it on its first
// lookup and appends what it tunes; the first line of a key wins, so every
// later process runs the schedule the first one chose.  EDX_TUNE_CACHE=0 (or
// empty) keeps choices in memory only.

namespace {

constexpr int TUNE_REPS = 5;

// m, n, nnz, precision, K, forms present
using TuneKey = std::tuple<int64_t, int64_t, int64_t, int, int, int>;
std::map<TuneKey, int> g_tuned;   // single vector: lanes; multi-RHS: rl * 256 + p
std::mutex g_tune_mu;
bool g_tune_loaded = false;
std::string g_tune_path;          // empty: no file
std::string g_tune_device;        // "<arch>:<CUs>"

int forms_of(const edx_csr *a) {
    return (a->d_off_pk ? 1 : 0) | (a->d_fr_rowptr ? 2 : 0);
}

TuneKey tune_key(const edx_csr *a, int precision, int K) {
    return TuneKey{a->m, a->n, a->nnz, precision, K, forms_of(a)};
}

std::string tune_cache_path() {
    const char *p = getenv("EDX_TUNE_CACHE");
    if (p) return (p[0] == 0 || (p[0] == '0' && p[1] == 0)) ? std::string() : std::string(p);
    const char *x = getenv("XDG_CACHE_HOME");
    if (x && x[0]) return std::string(x) + "/edrixs_amd/tune-v2.txt";
    const char *h = getenv("HOME");
    if (h && h[0]) return std::string(h) + "/.cache/edrixs_amd/tune-v2.txt";
    return std::string();
}

// Reads the cache file once (caller holds g_tune_mu).
void tune_load_locked() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, edx::current_device()) == hipSuccess)
        g_tune_device = std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount);
    else
        g_tune_device = "unknown:0";
    for (char &c : g_tune_device)
        if (c == ' ') c = '_';
    g_tune_path = tune_cache_path();
    if (g_tune_path.empty()) return;
    FILE *f = fopen(g_tune_path.c_str(), "r");
    if (!f) return;
    char dev[256];
    long long m, n, nnz;
    int prec, K, forms, choice;
    char line[512];
    while (fgets(line, sizeof(line), f)) {
        if (sscanf(line, "%255s %lld %lld %lld %d %d %d %d", dev, &m, &n, &nnz, &prec, &K, &forms,
                   &choice) != 8)
            continue;
        if (g_tune_device != dev) continue;
        g_tuned.emplace(TuneKey{m, n, nnz, prec, K, forms}, choice);   // first line wins
    }
    fclose(f);
}

// Records a new choice in memory and appends it to the cache file (caller
// holds g_tune_mu).  A file that cannot be written leaves the choice in memory.
void tune_store_locked(const TuneKey &k, int choice) {
    if (!g_tuned.emplace(k, choice).second) return;
    if (g_tune_path.empty()) return;
    for (size_t i = 1; i < g_tune_path.size(); ++i)   // create the directories
        if (g_tune_path[i] == '/') mkdir(g_tune_path.substr(0, i).c_str(), 0755);
    FILE *f = fopen(g_tune_path.c_str(), "a");
    if (!f) return;
    fprintf(f, "%s %lld %lld %lld %d %d %d %d\n", g_tune_device.c_str(),
            (long long)std::get<0>(k), (long long)std::get<1>(k), (long long)std::get<2>(k),
            std::get<3>(k), std::get<4>(k), std::get<5>(k), choice);
    fclose(f);
}

bool tuning_off() {
    const char *off = getenv("EDX_NO_TUNE");
    return off && off[0] == '1';
}

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
    if (a->m < TUNE_MIN_ROWS || tuning_off()) return EDX_OK;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_locked();
        auto it = g_tuned.find(tune_key(a, precision, 1));
        if (it != g_tuned.end() && schedule_available(a, precision, it->second)) {
            *slot = it->second;
            return EDX_OK;
        }
    }
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
    for (int G : {4, 8, 16, 32, 64, 143, 144, 145, 184, FAC_CODE}) {
        if (!schedule_available(a, precision, G)) continue;
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
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_store_locked(tune_key(a, precision, 1), best);
        // another thread may have stored the key first: take the stored choice
        best = g_tuned[tune_key(a, precision, 1)];
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

int tune_bspmm(const edx_csr *a, int precision, int K, BSched *sched) {
    if (a->m < TUNE_MIN_ROWS || tuning_off()) return EDX_OK;
    const TuneKey key = tune_key(a, precision, K);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_locked();
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            *sched = BSched{it->second / 256, it->second % 256};
            return EDX_OK;
        }
    }
    std::vector<BSched> cand = {BSched{0, 8}, BSched{0, 16}};
    if (K >= 8)
        for (int npl : {1, 2, 4, 8})
            if (16 * npl <= ((K + 15) / 16) * 16) cand.push_back(BSched{1, npl});
    // the pipelined RHS-parallel schedules for every K: a block compacted below
    // 8 vectors keeps them (lanes past K of each 16-lane group idle)
    for (int bpc : {4, 5, 8})   // workgroups per CU
        if (bsched_blocks(a, BSched{2, bpc}) <= bspmm_blocks(a->m))
            cand.push_back(BSched{2, bpc});
    if (packed_spmm_ok(a, precision))   // the same over the packed words, table in LDS
        for (int bpc : {4, 5, 8})
            if (bsched_blocks(a, BSched{4, bpc}) <= bspmm_blocks(a->m))
                cand.push_back(BSched{4, bpc});
    if (fac_spmm_ok(a, precision, K))   // slot-factored, workgroups per CU
        for (int bpc : {4, 5, 6, 8})
            if (bsched_blocks(a, BSched{3, bpc}) <= bspmm_blocks(a->m))
                cand.push_back(BSched{3, bpc});
    const size_t nk = (size_t)a->n * K;
    double2 *X = (double2 *)dmalloc(nk * sizeof(double2), "tune X");
    float2 *X32 = precision == EDX_C64 ? (float2 *)dmalloc(nk * sizeof(float2), "tune X32") : nullptr;
    double2 *Y = (double2 *)dmalloc((size_t)a->m * K * sizeof(double2), "tune Y");
    double2 *part = (double2 *)dmalloc((size_t)bspmm_blocks(a->m) * K * sizeof(double2), "tune partials");
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = EDX_OK;
    BSched best = *sched;
    float best_ms = 0.f;
    bool have = false;
    const bool prof_saved = g_prof;
    g_prof = false;
    if (!X || !Y || !part || (precision == EDX_C64 && !X32)) {
        rc = EDX_ERR_NOMEM;
        goto done;
    }
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess ||
        hipMemsetAsync(X, 0, nk * sizeof(double2), stream()) != hipSuccess ||
        hipMemsetAsync(Y, 0, (size_t)a->m * K * sizeof(double2), stream()) != hipSuccess ||
        (X32 && hipMemsetAsync(X32, 0, nk * sizeof(float2), stream()) != hipSuccess)) {
        rc = set_error(EDX_ERR_HIP, "tune: setup failed");
        goto done;
    }
    for (const BSched &c : cand) {
        BEpi e{};
        e.y = Y;
        e.z = Y;
        e.dvec = X;
        e.partials = part;
        if ((rc = bspmm_launch(a, precision, K, X, X32, e, c)) != EDX_OK) goto done;
        (void)hipEventRecord(ev[0], stream());
        for (int r = 0; r < TUNE_REPS; ++r)
            if ((rc = bspmm_launch(a, precision, K, X, X32, e, c)) != EDX_OK) goto done;
        (void)hipEventRecord(ev[1], stream());
        if (hipEventSynchronize(ev[1]) != hipSuccess) {
            rc = set_error(EDX_ERR_HIP, "tune: synchronize failed");
            goto done;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (!have || ms < best_ms) {
            best = c;
            best_ms = ms;
            have = true;
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_store_locked(key, best.rl * 256 + best.p);
        const int v = g_tuned[key];
        best = BSched{v / 256, v % 256};
    }
    *sched = best;
done:
    g_prof = prof_saved;
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    dfree(X);
    dfree(X32);
    dfree(Y);
    dfree(part);
    return rc;
}

}  // namespace edx

extern "C" {

int edx_csr_spmv_lanes(edx_csr *a, int precision, int set, int *lanes) {
    EDX_CHECK(a, EDX_ERR_ARG, "null matrix");
    EDX_CHECK(precision >= EDX_C128 && precision <= EDX_C128I, EDX_ERR_ARG, "bad precision %d",
              precision);
    EDX_CHECK(set == 0 || set == 4 || set == 8 || set == 16 || set == 32 || set == 64 ||
                  edx::is_pipe_code(set) || set == edx::FAC_CODE,
              EDX_ERR_ARG,
              "SpMV schedule must be 4, 8, 16, 32 or 64 lanes per row, the pipelined 143, 144, "
              "145 (packed indexed form) or 184, or 206 (slot-factored form) (got %d)", set);
    int *slot = lanes_slot(a, precision);
    if (set) *slot = set;
    if (lanes) *lanes = *slot;
    return EDX_OK;
}

}  // extern "C"
