// Density matrix and continued-fraction spectra (part of edx_kernels.hip).
//
//   edx_denmat       <- src/ed_driver.f90:219-248
//   edx_cf_spectrum  <- edrixs/plot_spectrum.py:52-59 (twin: src/lanczos.f90:306-336)

namespace {

// One block per (i <= j) orbital pair and row chunk; partial sums per block are
// reduced in a fixed order afterwards.
__global__ __launch_bounds__(256) void k_denmat_pairs(BasisView b, int norb, const double2 *v,
                                                      const int2 *pairs, int nchunk,
                                                      double2 *partials) {
    const int pid = blockIdx.x;
    const int i = pairs[pid].x, j = pairs[pid].y;
    const int64_t chunk = (b.n + nchunk - 1) / nchunk;
    const int64_t lo = blockIdx.y * chunk;
    const int64_t hi = lo + chunk < b.n ? lo + chunk : b.n;
    double2 sij = make_double2(0, 0), sji = make_double2(0, 0);
    for (int64_t r = lo + threadIdx.x; r < hi; r += 256) {
        const uint64_t w = (uint64_t)b.words[r];
        const double2 vr = v[r];
        if (i == j) {
            if ((w >> j) & 1ull) sij = cadd(sij, cconjmul(vr, vr));
            continue;
        }
        if (hypot(vr.x, vr.y) < 1e-10) continue;   // src/ed_driver.f90:226
        if (!((w >> i) & 1ull)) continue;
        int par = __popcll(w & ((1ull << i) - 1ull));
        uint64_t m = w ^ (1ull << i);
        if ((m >> j) & 1ull) continue;
        par += __popcll(m & ((1ull << j) - 1ull));
        m |= (1ull << j);
        const int64_t c = basis_index(b, m);
        if (c < 0) continue;
        const double sg = (par & 1) ? -1.0 : 1.0;
        const double2 vc = v[c];
        sij = cadd(sij, cscale(cconjmul(vr, vc), sg));
        sji = cadd(sji, cscale(cconjmul(vc, vr), sg));
    }
    double2 a = block_sum_256(sij);
    double2 c2 = block_sum_256(sji);
    if (threadIdx.x == 0) {
        partials[((int64_t)pid * gridDim.y + blockIdx.y) * 2] = a;
        partials[((int64_t)pid * gridDim.y + blockIdx.y) * 2 + 1] = c2;
    }
}

__global__ void k_cf(int neff, const double *alpha, const double *beta, double norm, double e_gs,
                     double weight, int nw, const double *omega, const double *gamma,
                     double *spec) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const double2 z = make_double2(omega[w] + e_gs, gamma[w]);
    double2 t = make_double2(0, 0);
    for (int j = neff - 1; j >= 1; --j) {
        const double b2 = beta[j - 1] * beta[j - 1];
        t = cdiv(make_double2(b2, 0.0), csub(make_double2(z.x - alpha[j], z.y), t));
    }
    t = cdiv(make_double2(1.0, 0.0), csub(make_double2(z.x - alpha[0], z.y), t));
    spec[w] += -1.0 / 3.14159265358979323846 * t.y * norm * weight;
}

}  // namespace

extern "C" {

int edx_denmat(const edx_basis *fock, int num_val_orbs, const void *d_vec, double *denmat) {
    EDX_CHECK(fock && d_vec && denmat && num_val_orbs >= 1 && num_val_orbs <= 63, EDX_ERR_ARG,
              "bad argument");
    EDX_TRY(edx::ensure_binomials());
    const int no = num_val_orbs;
    std::vector<int2> pairs;
    for (int i = 0; i < no; ++i)
        for (int j = i; j < no; ++j) pairs.push_back(make_int2(i, j));
    const int np = (int)pairs.size();
    int nchunk = (int)std::min<int64_t>(64, std::max<int64_t>(1, fock->n / 4096));
    DevBuf dp, part;
    EDX_TRY(dp.alloc(np * sizeof(int2), "pairs"));
    EDX_TRY(part.alloc((size_t)np * nchunk * 2 * sizeof(double2), "denmat partials"));
    EDX_HIP(edx::copy_sync(dp.p, pairs.data(), np * sizeof(int2), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_denmat_pairs, dim3(np, nchunk), dim3(256), 0, edx::stream(), view_of(fock),
                       no, (const double2 *)d_vec, dp.as<int2>(), nchunk, part.as<double2>());
    EDX_LAUNCH_CHECK();
    std::vector<double2> hp((size_t)np * nchunk * 2);
    EDX_HIP(edx::copy_sync(hp.data(), part.p, hp.size() * sizeof(double2), hipMemcpyDeviceToHost));
    double *dm = denmat;
    std::memset(dm, 0, sizeof(double) * 2 * no * no);
    for (int p = 0; p < np; ++p) {
        double2 a = make_double2(0, 0), c = make_double2(0, 0);
        for (int k = 0; k < nchunk; ++k) {
            a.x += hp[((size_t)p * nchunk + k) * 2].x;
            a.y += hp[((size_t)p * nchunk + k) * 2].y;
            c.x += hp[((size_t)p * nchunk + k) * 2 + 1].x;
            c.y += hp[((size_t)p * nchunk + k) * 2 + 1].y;
        }
        const int i = pairs[p].x, j = pairs[p].y;
        dm[2 * (i * no + j)] += a.x;
        dm[2 * (i * no + j) + 1] += a.y;
        if (i != j) {
            dm[2 * (j * no + i)] += c.x;
            dm[2 * (j * no + i) + 1] += c.y;
        }
    }
    return EDX_OK;
}

int edx_cf_spectrum(int neff, const double *alpha, const double *beta, double norm, double e_gs,
                    double weight, int nw, const double *omega, const double *gamma, double *spec) {
    EDX_CHECK(neff >= 1 && nw >= 0 && alpha && beta && (nw == 0 || (omega && gamma && spec)),
              EDX_ERR_ARG, "bad argument");
    if (nw == 0) return EDX_OK;
    DevBuf buf;
    const size_t bytes = sizeof(double) * (2 * (size_t)neff + 3 * (size_t)nw);
    EDX_TRY(buf.alloc(bytes, "cf buffers"));
    double *da = buf.as<double>(), *db = da + neff, *dom = db + neff, *dg = dom + nw, *ds = dg + nw;
    EDX_HIP(edx::copy_sync(da, alpha, neff * sizeof(double), hipMemcpyHostToDevice));
    EDX_HIP(edx::copy_sync(db, beta, neff * sizeof(double), hipMemcpyHostToDevice));
    EDX_HIP(edx::copy_sync(dom, omega, nw * sizeof(double), hipMemcpyHostToDevice));
    EDX_HIP(edx::copy_sync(dg, gamma, nw * sizeof(double), hipMemcpyHostToDevice));
    EDX_HIP(edx::copy_sync(ds, spec, nw * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cf, dim3((nw + 255) / 256), dim3(256), 0, edx::stream(), neff, da, db, norm,
                       e_gs, weight, nw, dom, dg, ds);
    EDX_LAUNCH_CHECK();
    EDX_HIP(edx::copy_sync(spec, ds, nw * sizeof(double), hipMemcpyDeviceToHost));
    return EDX_OK;
}

}  // extern "C"
