// Device-side helpers for the libedx HIP translation unit (edx_kernels.hip).
#pragma once

#include "common.h"

namespace edx {
extern __constant__ int64_t c_binom[64 * 64];
}  // namespace edx

// ----------------------------------------------------------- device helpers

__device__ __forceinline__ int64_t binom_dev(int n, int k) {
    return (k < 0 || n < 0 || k > n) ? 0 : edx::c_binom[n * 64 + k];
}

// Colex rank of w among nbits-bit words with popcount nocc, or -1.
__device__ __forceinline__ int64_t colex_rank(uint64_t w, int nbits, int nocc) {
    if (nbits < 64 && (w >> nbits) != 0) return -1;
    if (__popcll(w) != nocc) return -1;
    int64_t r = 0;
    int k = 0;
    while (w) {
        int p = __ffsll((unsigned long long)w) - 1;
        ++k;
        r += binom_dev(p, k);
        w &= w - 1;
    }
    return r;
}

__device__ __forceinline__ int64_t search_words(const int64_t *words, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        int64_t v = words[mid];
        if (v < key) lo = mid + 1;
        else if (v > key) hi = mid - 1;
        else return mid;
    }
    return -1;
}

__device__ __forceinline__ int64_t basis_index(const BasisView &b, uint64_t w) {
    if (b.complete) return colex_rank(w, b.nbits, b.nocc);
    return search_words(b.words, b.n, (int64_t)w);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {   // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) {
    return make_double2(a.x * s, a.y * s);
}
