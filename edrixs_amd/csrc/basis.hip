// Fock-basis enumeration and ranking (part of the edx_kernels.hip unit).
//
// The reference stores each basis as the ascending list of words with a fixed
// popcount (edrixs/fock_basis.py:443-444) and finds a word's index by binary
// search (src/utils.f90:2-40).  Ascending order of equal-popcount words is
// colex order, so index(w) = sum_k C(p_k, k) over the set-bit positions
// p_1 < p_2 < ... (0-based); the device computes it from a binomial table in
// constant memory instead of searching.  Bases read from files that are not the
// complete popcount set keep the binary search.

namespace edx {

__constant__ int64_t c_binom[64 * 64];

static bool g_binom_ready[64];

int ensure_binomials() {
    int dev = 0;
    EDX_HIP(hipGetDevice(&dev));
    if (dev < 64 && g_binom_ready[dev]) return EDX_OK;
    static int64_t table[64 * 64];
    for (int n = 0; n < 64; ++n) {
        for (int k = 0; k < 64; ++k) {
            if (k > n) table[n * 64 + k] = 0;
            else if (k == 0 || k == n) table[n * 64 + k] = 1;
            else table[n * 64 + k] = table[(n - 1) * 64 + k - 1] + table[(n - 1) * 64 + k];
        }
    }
    EDX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_binom), table, sizeof(table), 0,
                                   hipMemcpyHostToDevice, stream()));
    EDX_HIP(hipStreamSynchronize(stream()));
    if (dev < 64) g_binom_ready[dev] = true;
    return EDX_OK;
}

int64_t binom_host(int n, int k) {
    if (k < 0 || k > n) return 0;
    int64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

}  // namespace edx

// word of colex rank r among nbits-bit words of popcount nocc
__device__ __forceinline__ uint64_t colex_unrank(int64_t r, int nbits, int nocc) {
    uint64_t w = 0;
    int p = nbits - 1;
    for (int k = nocc; k >= 1; --k) {
        while (binom_dev(p, k) > r) --p;
        w |= 1ull << p;
        r -= binom_dev(p, k);
        --p;
    }
    return w;
}

__global__ void k_unrank(int64_t n, int nbits, int nocc, int64_t *words) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        words[i] = (int64_t)colex_unrank(i, nbits, nocc);
}

__global__ void k_check_complete(int64_t n, int nbits, int nocc, const int64_t *words,
                                 int *mismatch) {
    int bad = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        bad |= (words[i] != (int64_t)colex_unrank(i, nbits, nocc));
    if (__any(bad) && (threadIdx.x % 64) == 0) atomicOr(mismatch, 1);
}

__global__ void k_check_sorted(int64_t n, const int64_t *words, int *bad) {
    int b = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i + 1 < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b |= !(words[i] < words[i + 1]);
    if (__any(b) && (threadIdx.x % 64) == 0) atomicOr(bad, 1);
}

__global__ void k_lookup(BasisView b, const int64_t *in, int64_t n, int64_t *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = basis_index(b, (uint64_t)in[i]);
}

static int grid_for(int64_t n, int block = 256) {
    int64_t g = (n + block - 1) / block;
    int64_t cap = (int64_t)edx::num_cus() * 32;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

extern "C" {

int edx_basis_combinations(int nbits, int nocc, edx_basis **out) {
    EDX_CHECK(out, EDX_ERR_ARG, "null output");
    EDX_CHECK(nbits >= 1 && nbits <= 62 && nocc >= 0 && nocc <= nbits, EDX_ERR_ARG,
              "basis needs 1 <= nbits <= 62 and 0 <= nocc <= nbits (got %d, %d)", nbits, nocc);
    EDX_TRY(edx::ensure_binomials());
    int64_t n = edx::binom_host(nbits, nocc);
    int64_t *d = (int64_t *)edx::dmalloc(n * sizeof(int64_t), "basis words");
    EDX_CHECK(d, EDX_ERR_NOMEM, "%s", edx_last_error());
    hipLaunchKernelGGL(k_unrank, dim3(grid_for(n)), dim3(256), 0, edx::stream(), n, nbits, nocc,
                       d);
    EDX_LAUNCH_CHECK();
    EDX_HIP(hipStreamSynchronize(edx::stream()));
    edx_basis *b = new edx_basis{n, nbits, nocc, 1, d};
    *out = b;
    return EDX_OK;
}

int edx_basis_from_words(const int64_t *words, int64_t n, int nbits, edx_basis **out) {
    EDX_CHECK(out && (words || n == 0) && n >= 0, EDX_ERR_ARG, "bad basis arguments");
    EDX_CHECK(nbits >= 1 && nbits <= 62, EDX_ERR_ARG, "nbits %d out of range", nbits);
    EDX_TRY(edx::ensure_binomials());
    int64_t *d = (int64_t *)edx::dmalloc((n ? n : 1) * sizeof(int64_t), "basis words");
    EDX_CHECK(d, EDX_ERR_NOMEM, "%s", edx_last_error());
    if (n) EDX_HIP(edx::copy_sync(d, words, n * sizeof(int64_t), hipMemcpyHostToDevice));
    int nocc = n ? __builtin_popcountll((unsigned long long)words[0]) : 0;
    int complete = 0;
    if (n > 0 && n == edx::binom_host(nbits, nocc)) {
        int *flag = (int *)edx::dmalloc(sizeof(int), "flag");
        EDX_CHECK(flag, EDX_ERR_NOMEM, "%s", edx_last_error());
        EDX_HIP(edx::fill_sync(flag, 0, sizeof(int)));
        hipLaunchKernelGGL(k_check_complete, dim3(grid_for(n)), dim3(256), 0, edx::stream(), n,
                           nbits, nocc, d, flag);
        EDX_LAUNCH_CHECK();
        int h = 1;
        EDX_HIP(edx::copy_sync(&h, flag, sizeof(int), hipMemcpyDeviceToHost));
        edx::dfree(flag);
        complete = (h == 0);
    }
    if (!complete && n > 1) {   // binary search needs strictly ascending words
        int *flag = (int *)edx::dmalloc(sizeof(int), "flag");
        EDX_CHECK(flag, EDX_ERR_NOMEM, "%s", edx_last_error());
        EDX_HIP(edx::fill_sync(flag, 0, sizeof(int)));
        hipLaunchKernelGGL(k_check_sorted, dim3(grid_for(n)), dim3(256), 0, edx::stream(), n, d,
                           flag);
        EDX_LAUNCH_CHECK();
        int h = 0;
        EDX_HIP(edx::copy_sync(&h, flag, sizeof(int), hipMemcpyDeviceToHost));
        edx::dfree(flag);
        if (h) {
            edx::dfree(d);
            return edx::set_error(EDX_ERR_ARG, "Fock words are not strictly ascending");
        }
    }
    *out = new edx_basis{n, nbits, nocc, complete, d};
    return EDX_OK;
}

int edx_basis_size(const edx_basis *b, int64_t *n, int *is_complete) {
    EDX_CHECK(b, EDX_ERR_ARG, "null basis");
    if (n) *n = b->n;
    if (is_complete) *is_complete = b->complete;
    return EDX_OK;
}

int edx_basis_download(const edx_basis *b, int64_t *words) {
    EDX_CHECK(b && words, EDX_ERR_ARG, "null argument");
    if (b->n) EDX_HIP(edx::copy_sync(words, b->d_words, b->n * sizeof(int64_t), hipMemcpyDeviceToHost));
    return EDX_OK;
}

int edx_basis_lookup(const edx_basis *b, const int64_t *words, int64_t n, int64_t *index) {
    EDX_CHECK(b && (n == 0 || (words && index)), EDX_ERR_ARG, "null argument");
    if (n == 0) return EDX_OK;
    int64_t *d_in = (int64_t *)edx::dmalloc(n * sizeof(int64_t), "lookup in");
    int64_t *d_out = (int64_t *)edx::dmalloc(n * sizeof(int64_t), "lookup out");
    if (!d_in || !d_out) {
        edx::dfree(d_in);
        edx::dfree(d_out);
        return EDX_ERR_NOMEM;
    }
    EDX_HIP(edx::copy_sync(d_in, words, n * sizeof(int64_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_lookup, dim3(grid_for(n)), dim3(256), 0, edx::stream(), view_of(b), d_in,
                       n, d_out);
    EDX_LAUNCH_CHECK();
    EDX_HIP(edx::copy_sync(index, d_out, n * sizeof(int64_t), hipMemcpyDeviceToHost));
    edx::dfree(d_in);
    edx::dfree(d_out);
    return EDX_OK;
}

void edx_basis_free(edx_basis *b) {
    if (!b) return;
    edx::dfree(b->d_words);
    delete b;
}

}  // extern "C"
