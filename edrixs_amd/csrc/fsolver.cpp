// Drop-in drivers behind edx_{ed,xas,rixs,opavg}_fsolver.
//
// They reproduce the f2py entry points of src/pyapi.f90 and the drivers
// src/ed_driver.f90, src/xas_driver.f90, src/rixs_driver.f90 and
// src/opavg_driver.f90: the same files are read from the current directory
// (config.in namelist, hopping/coulomb/transop term lists, fock lists,
// eigvec.N) and the same files are written (eigvals.dat, denmat.dat,
// eigvec.N, xas_poles.N, rixs_poles.N, opavg.dat) with the reference formats
// (src/io.f90:515-659).  All numerical work runs on the device through the
// C ABI in include/edx.h.  One GPU serves the whole communicator: rank 0
// computes, other ranks return immediately.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"

namespace edx {
int herm_eig_lowest(int n, std::complex<double> *A, double *w, int nvec, std::complex<double> *Z);
}

namespace {

// ------------------------------------------------------------ config.in

// &control namelist with the defaults of src/io.f90:17-32
struct Config {
    int ed_solver = 1;
    int num_val_orbs = 2;
    int num_core_orbs = 2;
    int neval = 1;
    int nvector = 1;
    bool idump = false;
    int num_gs = 1;
    int maxiter = 500;
    int linsys_max = 500;
    int min_ndim = 1000;
    int ncv = 3;
    double eigval_tol = 1e-8;
    double linsys_tol = 1e-8;
    int nkryl = 500;
    double omega_in = 0.0;
    double gamma_in = 0.1;
};

std::string lower(std::string s) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

bool parse_real(const std::string &tok, double &v) {
    std::string t = tok;
    for (auto &c : t)
        if (c == 'd' || c == 'D') c = 'e';
    char *end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0' && !t.empty();
}

bool parse_int(const std::string &tok, int &v) {
    char *end = nullptr;
    long x = std::strtol(tok.c_str(), &end, 10);
    if (!end || *end != '\0' || tok.empty()) return false;
    v = (int)x;
    return true;
}

bool parse_logical(const std::string &tok, bool &v) {
    std::string t = lower(tok);
    if (!t.empty() && t[0] == '.') t = t.substr(1);
    if (t.empty()) return false;
    if (t[0] == 't') v = true;
    else if (t[0] == 'f') v = false;
    else return false;
    return true;
}

int read_config(Config &cfg) {
    std::ifstream f("config.in");
    EDX_CHECK(f.good(), EDX_ERR_IO, " fedrixs >>> ERROR: config.in doesn't exist");
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    std::string lt = lower(text);
    size_t start = lt.find("&control");
    EDX_CHECK(start != std::string::npos, EDX_ERR_IO, "config.in has no &control namelist");
    start += 8;
    size_t stop = lt.find("&end", start);
    size_t slash = lt.find('/', start);
    if (slash != std::string::npos && (stop == std::string::npos || slash < stop)) stop = slash;
    EDX_CHECK(stop != std::string::npos, EDX_ERR_IO, "config.in namelist is not terminated");
    std::string body = text.substr(start, stop - start);
    cfg.ncv = cfg.neval + 2;
    // split on newlines and commas into name=value items
    std::vector<std::string> items;
    std::string cur;
    for (char c : body) {
        if (c == '\n' || c == ',') {
            items.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    items.push_back(cur);
    for (auto &raw : items) {
        std::string it = trim(raw);
        if (it.empty() || it[0] == '!') continue;
        size_t eq = it.find('=');
        EDX_CHECK(eq != std::string::npos, EDX_ERR_IO, "config.in: cannot parse '%s'", it.c_str());
        std::string key = lower(trim(it.substr(0, eq)));
        std::string val = trim(it.substr(eq + 1));
        size_t bang = val.find('!');
        if (bang != std::string::npos) val = trim(val.substr(0, bang));
        bool ok = true;
        if (key == "ed_solver") ok = parse_int(val, cfg.ed_solver);
        else if (key == "num_val_orbs") ok = parse_int(val, cfg.num_val_orbs);
        else if (key == "num_core_orbs") ok = parse_int(val, cfg.num_core_orbs);
        else if (key == "neval") ok = parse_int(val, cfg.neval);
        else if (key == "nvector") ok = parse_int(val, cfg.nvector);
        else if (key == "idump") ok = parse_logical(val, cfg.idump);
        else if (key == "num_gs") ok = parse_int(val, cfg.num_gs);
        else if (key == "maxiter") ok = parse_int(val, cfg.maxiter);
        else if (key == "linsys_max") ok = parse_int(val, cfg.linsys_max);
        else if (key == "min_ndim") ok = parse_int(val, cfg.min_ndim);
        else if (key == "ncv") ok = parse_int(val, cfg.ncv);
        else if (key == "eigval_tol") ok = parse_real(val, cfg.eigval_tol);
        else if (key == "linsys_tol") ok = parse_real(val, cfg.linsys_tol);
        else if (key == "nkryl") ok = parse_int(val, cfg.nkryl);
        else if (key == "omega_in") ok = parse_real(val, cfg.omega_in);
        else if (key == "gamma_in") ok = parse_real(val, cfg.gamma_in);
        else return edx::set_error(EDX_ERR_IO, "config.in: unknown namelist item '%s'", key.c_str());
        EDX_CHECK(ok, EDX_ERR_IO, "config.in: bad value '%s' for %s", val.c_str(), key.c_str());
    }
    return EDX_OK;
}

// ------------------------------------------------------------ term files

// Whitespace / comma separated tokens of a list-directed Fortran file.
int read_tokens(const char *fname, std::vector<std::string> &tok) {
    std::ifstream f(fname);
    EDX_CHECK(f.good(), EDX_ERR_IO, " fedrixs >>> ERROR: %s doesn't exist", fname);
    std::string line;
    while (std::getline(f, line)) {
        std::string cur;
        for (char c : line) {
            if (std::isspace((unsigned char)c) || c == ',') {
                if (!cur.empty()) tok.push_back(cur);
                cur.clear();
            } else {
                cur += c;
            }
        }
        if (!cur.empty()) tok.push_back(cur);
    }
    return EDX_OK;
}

int read_t2(const char *fname, std::vector<edx_tensor2> &out) {   // src/io.f90:70-114
    std::vector<std::string> t;
    EDX_TRY(read_tokens(fname, t));
    int n = 0;
    EDX_CHECK(!t.empty() && parse_int(t[0], n) && n >= 0, EDX_ERR_IO, "%s: bad term count", fname);
    EDX_CHECK((int64_t)t.size() >= 1 + 4 * (int64_t)n, EDX_ERR_IO, "%s: truncated", fname);
    out.resize(n);
    for (int k = 0; k < n; ++k) {
        const std::string *p = &t[1 + 4 * (size_t)k];
        bool ok = parse_int(p[0], out[k].ind1) && parse_int(p[1], out[k].ind2) &&
                  parse_real(p[2], out[k].re) && parse_real(p[3], out[k].im);
        EDX_CHECK(ok, EDX_ERR_IO, "%s: bad term %d", fname, k + 1);
    }
    return EDX_OK;
}

int read_t4(const char *fname, std::vector<edx_tensor4> &out) {   // src/io.f90:117-171
    std::vector<std::string> t;
    EDX_TRY(read_tokens(fname, t));
    int n = 0;
    EDX_CHECK(!t.empty() && parse_int(t[0], n) && n >= 0, EDX_ERR_IO, "%s: bad term count", fname);
    EDX_CHECK((int64_t)t.size() >= 1 + 6 * (int64_t)n, EDX_ERR_IO, "%s: truncated", fname);
    out.resize(n);
    for (int k = 0; k < n; ++k) {
        const std::string *p = &t[1 + 6 * (size_t)k];
        bool ok = parse_int(p[0], out[k].ind1) && parse_int(p[1], out[k].ind2) &&
                  parse_int(p[2], out[k].ind3) && parse_int(p[3], out[k].ind4) &&
                  parse_real(p[4], out[k].re) && parse_real(p[5], out[k].im);
        EDX_CHECK(ok, EDX_ERR_IO, "%s: bad term %d", fname, k + 1);
    }
    return EDX_OK;
}

// fock_*.in: count, then one decimal word per line (src/io.f90:174-199)
int read_fock(const char *fname, std::vector<int64_t> &out) {
    FILE *f = std::fopen(fname, "r");
    EDX_CHECK(f, EDX_ERR_IO, " fedrixs >>> ERROR: %s doesn't exist", fname);
    long long n = 0;
    if (std::fscanf(f, "%lld", &n) != 1 || n < 0) {
        std::fclose(f);
        return edx::set_error(EDX_ERR_IO, "%s: bad count", fname);
    }
    out.resize((size_t)n);
    for (long long k = 0; k < n; ++k) {
        long long w;
        if (std::fscanf(f, "%lld", &w) != 1) {
            std::fclose(f);
            return edx::set_error(EDX_ERR_IO, "%s: truncated at entry %lld", fname, k + 1);
        }
        out[(size_t)k] = w;
    }
    std::fclose(f);
    return EDX_OK;
}

// ------------------------------------------ Fortran unformatted eigvec.N

int write_eigvec(const std::string &fname, double eigval, const std::vector<double2> &v) {
    FILE *f = std::fopen(fname.c_str(), "wb");
    EDX_CHECK(f, EDX_ERR_IO, "cannot write %s", fname.c_str());
    int32_t m1 = 8;
    std::fwrite(&m1, 4, 1, f);
    std::fwrite(&eigval, 8, 1, f);
    std::fwrite(&m1, 4, 1, f);
    int64_t bytes = (int64_t)v.size() * 16;
    EDX_CHECK(bytes < ((int64_t)1 << 31), EDX_ERR_IO, "eigenvector too long for one record");
    int32_t m2 = (int32_t)bytes;
    std::fwrite(&m2, 4, 1, f);
    std::fwrite(v.data(), 16, v.size(), f);
    std::fwrite(&m2, 4, 1, f);
    std::fclose(f);
    return EDX_OK;
}

int read_eigvec(const std::string &fname, int64_t n, double &eigval, std::vector<double2> &v) {
    FILE *f = std::fopen(fname.c_str(), "rb");
    EDX_CHECK(f, EDX_ERR_IO, " fedrixs >>> ERROR: %s doesn't exist", fname.c_str());
    int32_t m1 = 0, m1b = 0, m2 = 0;
    bool ok = std::fread(&m1, 4, 1, f) == 1 && m1 == 8 && std::fread(&eigval, 8, 1, f) == 1 &&
              std::fread(&m1b, 4, 1, f) == 1 && std::fread(&m2, 4, 1, f) == 1;
    if (!ok || (int64_t)m2 < n * 16) {
        std::fclose(f);
        return edx::set_error(EDX_ERR_IO, "%s: not an eigenvector record of length %lld",
                              fname.c_str(), (long long)n);
    }
    v.resize((size_t)n);
    ok = std::fread(v.data(), 16, (size_t)n, f) == (size_t)n;
    std::fclose(f);
    EDX_CHECK(ok, EDX_ERR_IO, "%s: truncated", fname.c_str());
    return EDX_OK;
}

int write_krylov(const std::string &fname, int neff, const double *alpha, const double *beta,
                 double norm, double e_gs) {   // src/io.f90:515-541
    FILE *f = std::fopen(fname.c_str(), "w");
    EDX_CHECK(f, EDX_ERR_IO, "cannot write %s", fname.c_str());
    std::fprintf(f, "%-20s%10d\n", "#number_of_poles:", neff);
    std::fprintf(f, "%-20s%20.10f\n", "#enegry:", e_gs);
    std::fprintf(f, "%-20s%20.10f\n", "#normalization:", norm);
    for (int i = 0; i < neff; ++i) std::fprintf(f, "%5d%20.10f%20.10f\n", i + 1, alpha[i], beta[i]);
    std::fclose(f);
    return EDX_OK;
}

// --------------------------------------------------------- RAII handles

struct Basis {
    edx_basis *b = nullptr;
    ~Basis() { edx_basis_free(b); }
};
struct Csr {
    edx_csr *a = nullptr;
    ~Csr() { edx_csr_free(a); }
};
struct Dev {
    void *p = nullptr;
    ~Dev() { edx::dfree(p); }
    double2 *z() const { return (double2 *)p; }
};

int dev_alloc(Dev &d, int64_t n, const char *what) {
    d.p = edx::dmalloc((size_t)(n > 0 ? n : 1) * sizeof(double2), what);
    return d.p ? EDX_OK : EDX_ERR_NOMEM;
}

int load_basis(const char *fname, int nbits, Basis &b, int64_t *n_out) {
    std::vector<int64_t> w;
    EDX_TRY(read_fock(fname, w));
    EDX_TRY(edx_basis_from_words(w.data(), (int64_t)w.size(), nbits, &b.b));
    if (n_out) *n_out = (int64_t)w.size();
    return EDX_OK;
}

uint64_t lanczos_seed() {
    const char *s = std::getenv("EDX_SEED");
    return s ? std::strtoull(s, nullptr, 10) : 20190131ull;
}

int lanczos_deflate() {
    const char *s = std::getenv("EDX_LANCZOS_DEFLATE");
    return s ? std::atoi(s) : 1;
}

// Form of the products with a square operator: the exact indexed form where
// its distinct off-diagonal values fit the table (the same FP64 operator,
// diagonal added apart, faster products), else complex128 CSR;
// EDX_FSOLVER_C128=1 keeps complex128 CSR throughout.
int exact_form(edx_csr *a) {
    const char *s = std::getenv("EDX_FSOLVER_C128");
    if (s && *s && *s != '0') return EDX_C128;
    return edx_csr_prepare(a, EDX_C128I) == EDX_OK ? EDX_C128I : EDX_C128;
}

// ------------------------------------------------------------ ED driver

int ed_driver() {   // src/ed_driver.f90:1-281
    Config cfg;
    EDX_TRY(read_config(cfg));
    std::vector<edx_tensor2> hop;
    std::vector<edx_tensor4> coul;
    EDX_TRY(read_t2("hopping_i.in", hop));
    EDX_TRY(read_t4("coulomb_i.in", coul));
    std::vector<int64_t> fock;
    EDX_TRY(read_fock("fock_i.in", fock));
    const int64_t ndim = (int64_t)fock.size();
    EDX_CHECK(ndim >= 1, EDX_ERR_IO, "fock_i.in is empty");
    int nbits = 1;
    for (int64_t w : fock)
        while (nbits < 62 && (w >> nbits)) ++nbits;
    nbits = std::max(nbits, cfg.num_val_orbs);
    Basis B;
    EDX_TRY(edx_basis_from_words(fock.data(), ndim, std::min(nbits, 62), &B.b));
    int solver = cfg.ed_solver;
    if (ndim < cfg.min_ndim) solver = 0;
    int neval = (int)std::min<int64_t>(cfg.neval, ndim);
    int nvector = (int)std::min<int64_t>(cfg.nvector, ndim);
    Csr H;
    EDX_TRY(edx_build_ham_i(B.b, hop.data(), (int64_t)hop.size(), coul.data(), (int64_t)coul.size(),
                            0.0, 0.0, 1.0, EDX_C128, &H.a));
    std::vector<double> evals(std::max(neval, 1));
    Dev vecs;
    EDX_TRY(dev_alloc(vecs, ndim * std::max(nvector, 1), "eigenvectors"));
    if (solver == 0) {   // full diagonalization (src/ed_driver.f90:127-152)
        EDX_CHECK(ndim <= 20000, EDX_ERR_ARG, "full diagonalization limited to dim <= 20000");
        int64_t nnz = 0;
        EDX_TRY(edx_csr_info(H.a, nullptr, nullptr, &nnz, nullptr));
        std::vector<int64_t> rp(ndim + 1);
        std::vector<int32_t> cols(nnz);
        std::vector<std::complex<double>> vals(nnz);
        EDX_TRY(edx_csr_download(H.a, rp.data(), cols.data(), vals.data()));
        // only the nvector lowest eigenvectors are kept (src/ed_driver.f90:147-148)
        const int nz = std::max(nvector, 1);
        std::vector<std::complex<double>> A((size_t)ndim * ndim, 0.0), Z((size_t)ndim * nz);
        for (int64_t r = 0; r < ndim; ++r)
            for (int64_t k = rp[r]; k < rp[r + 1]; ++k) A[(size_t)r * ndim + cols[k]] = vals[k];
        std::vector<double> w(ndim);
        EDX_CHECK(edx::herm_eig_lowest((int)ndim, A.data(), w.data(), nvector, Z.data()) == 0,
                  EDX_ERR_NUMERIC, "dense eigensolver did not converge");
        for (int k = 0; k < neval; ++k) evals[k] = w[k];
        std::vector<double2> hv((size_t)ndim * nvector);
        for (int j = 0; j < nvector; ++j)
            for (int64_t r = 0; r < ndim; ++r)
                hv[(size_t)j * ndim + r] = make_double2(Z[(size_t)r * nvector + j].real(),
                                                        Z[(size_t)r * nvector + j].imag());
        EDX_HIP(edx::copy_sync(vecs.p, hv.data(), hv.size() * sizeof(double2), hipMemcpyHostToDevice));
    } else {   // Lanczos (ed_solver=1) or, for ed_solver=2, thick-restart Lanczos with
               // ARPACK's ncv / maxiter / tol meaning (src/ed_driver.f90:154-172)
        int nev = std::max(neval, nvector);
        std::vector<double> ev(nev);
        Dev tmp;
        EDX_TRY(dev_alloc(tmp, ndim * nev, "eigenvectors"));
        int iters = 0, nconv = 0;
        if (solver == 2)
            EDX_TRY(edx_trlan_eigs(H.a, exact_form(H.a), nev, std::max(cfg.ncv, 1), cfg.maxiter,
                                   cfg.eigval_tol, lanczos_seed(), lanczos_deflate(), nev, ev.data(),
                                   tmp.p, &iters, &nconv));
        else
            EDX_TRY(edx_lanczos_eigs(H.a, exact_form(H.a), nev, cfg.maxiter, cfg.eigval_tol, lanczos_seed(),
                                     lanczos_deflate(), nev, ev.data(), tmp.p, &iters));
        for (int k = 0; k < neval; ++k) evals[k] = ev[k];
        if (nvector)
            EDX_HIP(edx::copy_sync(vecs.p, tmp.p, (size_t)ndim * nvector * sizeof(double2),
                              hipMemcpyDeviceToDevice));
    }
    {   // eigvals.dat, (i5, f20.10)   src/io.f90:587-607
        FILE *f = std::fopen("eigvals.dat", "w");
        EDX_CHECK(f, EDX_ERR_IO, "cannot write eigvals.dat");
        for (int k = 0; k < neval; ++k) std::fprintf(f, "%5d%20.10f\n", k + 1, evals[k]);
        std::fclose(f);
    }
    const int no = cfg.num_val_orbs;
    std::vector<double> dm((size_t)std::max(nvector, 1) * no * no * 2, 0.0);
    for (int k = 0; k < nvector; ++k) {
        const double2 *v = vecs.z() + (size_t)k * ndim;
        EDX_TRY(edx_denmat(B.b, no, v, dm.data() + (size_t)k * no * no * 2));
        if (cfg.idump) {
            std::vector<double2> hv(ndim);
            EDX_HIP(edx::copy_sync(hv.data(), v, ndim * sizeof(double2), hipMemcpyDeviceToHost));
            EDX_TRY(write_eigvec("eigvec." + std::to_string(k + 1), evals[std::min(k, neval - 1)], hv));
        }
    }
    {   // denmat.dat, (3i5, 2f20.10) looping vector, j, k   src/io.f90:610-635
        FILE *f = std::fopen("denmat.dat", "w");
        EDX_CHECK(f, EDX_ERR_IO, "cannot write denmat.dat");
        for (int i = 0; i < nvector; ++i)
            for (int j = 0; j < no; ++j)
                for (int k = 0; k < no; ++k) {
                    const double *e = dm.data() + ((size_t)i * no * no + (size_t)j * no + k) * 2;
                    std::fprintf(f, "%5d%5d%5d%20.10f%20.10f\n", i + 1, j + 1, k + 1, e[0], e[1]);
                }
        std::fclose(f);
    }
    return EDX_OK;
}

// ----------------------------------------------------------- XAS driver

int xas_driver() {   // src/xas_driver.f90:1-166
    Config cfg;
    EDX_TRY(read_config(cfg));
    std::vector<edx_tensor2> hop_n, trans;
    std::vector<edx_tensor4> coul_n;
    EDX_TRY(read_t2("hopping_n.in", hop_n));
    EDX_TRY(read_t4("coulomb_n.in", coul_n));
    EDX_TRY(read_t2("transop_xas.in", trans));
    Basis Bi, Bn;
    int64_t ndim_i = 0, ndim_nv = 0;
    EDX_TRY(load_basis("fock_i.in", cfg.num_val_orbs, Bi, &ndim_i));
    EDX_TRY(load_basis("fock_n.in", cfg.num_val_orbs, Bn, &ndim_nv));
    const int nc = cfg.num_core_orbs, nv = cfg.num_val_orbs;
    const int64_t ndim_n = ndim_nv * nc;
    Csr D, Hn;
    EDX_TRY(edx_build_transop_i(Bn.b, Bi.b, nv, nc, trans.data(), (int64_t)trans.size(), EDX_C128,
                                &D.a));
    EDX_TRY(edx_build_ham_n(Bn.b, nv, nc, hop_n.data(), (int64_t)hop_n.size(), coul_n.data(),
                            (int64_t)coul_n.size(), 0.0, 0.0, 1.0, EDX_C128, &Hn.a));
    Dev gs, phi;
    EDX_TRY(dev_alloc(gs, ndim_i, "ground state"));
    EDX_TRY(dev_alloc(phi, ndim_n, "phi"));
    std::vector<double> alpha(cfg.nkryl), beta(cfg.nkryl);
    for (int igs = 1; igs <= cfg.num_gs; ++igs) {
        std::vector<double2> hv;
        double eig = 0.0;
        EDX_TRY(read_eigvec("eigvec." + std::to_string(igs), ndim_i, eig, hv));
        EDX_HIP(edx::copy_sync(gs.p, hv.data(), ndim_i * sizeof(double2), hipMemcpyHostToDevice));
        EDX_HIP(edx::fill_sync(phi.p, 0, ndim_n * sizeof(double2)));
        EDX_TRY(edx_spmv(D.a, EDX_C128, gs.p, phi.p));
        int neff = 0;
        double norm = 0.0;
        EDX_TRY(edx_build_krylov(Hn.a, exact_form(Hn.a), phi.p, cfg.nkryl, &neff, alpha.data(), beta.data(),
                                 &norm));
        EDX_TRY(write_krylov("xas_poles." + std::to_string(igs), neff, alpha.data(), beta.data(),
                             norm, eig));
    }
    return EDX_OK;
}

// ---------------------------------------------------------- RIXS driver

int rixs_driver() {   // src/rixs_driver.f90:1-246
    Config cfg;
    EDX_TRY(read_config(cfg));
    std::vector<edx_tensor2> hop_i, hop_n, ti, tf;
    std::vector<edx_tensor4> coul_i, coul_n;
    EDX_TRY(read_t2("hopping_i.in", hop_i));
    EDX_TRY(read_t4("coulomb_i.in", coul_i));
    EDX_TRY(read_t2("hopping_n.in", hop_n));
    EDX_TRY(read_t4("coulomb_n.in", coul_n));
    EDX_TRY(read_t2("transop_rixs_i.in", ti));
    EDX_TRY(read_t2("transop_rixs_f.in", tf));
    const int nc = cfg.num_core_orbs, nv = cfg.num_val_orbs;
    Basis Bi, Bn, Bf;
    int64_t ndim_i = 0, ndim_nv = 0, ndim_f = 0;
    EDX_TRY(load_basis("fock_i.in", nv, Bi, &ndim_i));
    EDX_TRY(load_basis("fock_n.in", nv, Bn, &ndim_nv));
    EDX_TRY(load_basis("fock_f.in", nv, Bf, &ndim_f));
    const int64_t ndim_n = ndim_nv * nc;
    Csr Di, Df, Hf;
    EDX_TRY(edx_build_transop_i(Bn.b, Bi.b, nv, nc, ti.data(), (int64_t)ti.size(), EDX_C128, &Di.a));
    EDX_TRY(edx_build_transop_f(Bf.b, Bn.b, nv, nc, tf.data(), (int64_t)tf.size(), EDX_C128, &Df.a));
    EDX_TRY(edx_build_ham_i(Bf.b, hop_i.data(), (int64_t)hop_i.size(), coul_i.data(),
                            (int64_t)coul_i.size(), 0.0, 0.0, 1.0, EDX_C128, &Hf.a));
    Dev gs, phi, x, phif;
    EDX_TRY(dev_alloc(gs, ndim_i, "ground state"));
    EDX_TRY(dev_alloc(phi, ndim_n, "phi"));
    EDX_TRY(dev_alloc(x, ndim_n, "x"));
    EDX_TRY(dev_alloc(phif, ndim_f, "phi_f"));
    std::vector<double> alpha(cfg.nkryl), beta(cfg.nkryl);
    for (int igs = 1; igs <= cfg.num_gs; ++igs) {
        std::vector<double2> hv;
        double eig = 0.0;
        EDX_TRY(read_eigvec("eigvec." + std::to_string(igs), ndim_i, eig, hv));
        EDX_HIP(edx::copy_sync(gs.p, hv.data(), ndim_i * sizeof(double2), hipMemcpyHostToDevice));
        EDX_HIP(edx::fill_sync(phi.p, 0, ndim_n * sizeof(double2)));
        EDX_TRY(edx_spmv(Di.a, EDX_C128, gs.p, phi.p));
        // A = (omega_in + E_gs + i gamma_in) - H_n, omega last on the diagonal
        Csr A;
        EDX_TRY(edx_build_ham_n(Bn.b, nv, nc, hop_n.data(), (int64_t)hop_n.size(), coul_n.data(),
                                (int64_t)coul_n.size(), cfg.omega_in + eig, cfg.gamma_in, -1.0,
                                EDX_C128, &A.a));
        int iters = 0;
        double res = 0.0;
        EDX_TRY(edx_minres(A.a, exact_form(A.a), phi.p, x.p, cfg.linsys_tol, cfg.linsys_max, &iters, &res));
        EDX_HIP(edx::fill_sync(phif.p, 0, ndim_f * sizeof(double2)));
        EDX_TRY(edx_spmv(Df.a, EDX_C128, x.p, phif.p));
        int neff = 0;
        double norm = 0.0;
        EDX_TRY(edx_build_krylov(Hf.a, exact_form(Hf.a), phif.p, cfg.nkryl, &neff, alpha.data(), beta.data(),
                                 &norm));
        EDX_TRY(write_krylov("rixs_poles." + std::to_string(igs), neff, alpha.data(), beta.data(),
                             norm, eig));
    }
    return EDX_OK;
}

// --------------------------------------------------------- opavg driver

int opavg_driver() {   // src/opavg_driver.f90:1-126
    Config cfg;
    EDX_TRY(read_config(cfg));
    std::vector<edx_tensor2> hop;
    std::vector<edx_tensor4> coul;
    EDX_TRY(read_t2("hopping_i.in", hop));
    EDX_TRY(read_t4("coulomb_i.in", coul));
    Basis B;
    int64_t ndim = 0;
    EDX_TRY(load_basis("fock_i.in", cfg.num_val_orbs, B, &ndim));
    Csr H;
    EDX_TRY(edx_build_ham_i(B.b, hop.data(), (int64_t)hop.size(), coul.data(), (int64_t)coul.size(),
                            0.0, 0.0, 1.0, EDX_C128, &H.a));
    Dev v, w;
    EDX_TRY(dev_alloc(v, ndim, "v"));
    EDX_TRY(dev_alloc(w, ndim, "w"));
    std::vector<double> evals(cfg.num_gs);
    std::vector<std::complex<double>> avg(cfg.num_gs);
    for (int igs = 1; igs <= cfg.num_gs; ++igs) {
        std::vector<double2> hv;
        EDX_TRY(read_eigvec("eigvec." + std::to_string(igs), ndim, evals[igs - 1], hv));
        EDX_HIP(edx::copy_sync(v.p, hv.data(), ndim * sizeof(double2), hipMemcpyHostToDevice));
        EDX_HIP(edx::fill_sync(w.p, 0, ndim * sizeof(double2)));
        EDX_TRY(edx_spmv(H.a, EDX_C128, v.p, w.p));
        std::vector<double2> hw(ndim);
        EDX_HIP(edx::copy_sync(hw.data(), w.p, ndim * sizeof(double2), hipMemcpyDeviceToHost));
        std::complex<double> s = 0.0;
        for (int64_t k = 0; k < ndim; ++k)
            s += std::conj(std::complex<double>(hv[k].x, hv[k].y)) *
                 std::complex<double>(hw[k].x, hw[k].y);
        avg[igs - 1] = s;
    }
    FILE *f = std::fopen("opavg.dat", "w");   // (1i5, 4f20.10)  src/io.f90:638-659
    EDX_CHECK(f, EDX_ERR_IO, "cannot write opavg.dat");
    for (int i = 0; i < cfg.num_gs; ++i)
        std::fprintf(f, "%5d%20.10f%20.10f%20.10f\n", i + 1, evals[i], avg[i].real(), avg[i].imag());
    std::fclose(f);
    return EDX_OK;
}

int run_rank0(int my_id, int (*fn)()) {
    if (my_id != 0) return EDX_OK;
    int rc = fn();
    hipError_t e = hipStreamSynchronize(edx::stream());
    if (!rc && e != hipSuccess) rc = edx::hip_fail(e, "fsolver", __FILE__, __LINE__);
    return rc;
}

}  // namespace

extern "C" {

int edx_ed_fsolver(int comm, int my_id, int num_procs) {
    (void)comm;
    (void)num_procs;
    return run_rank0(my_id, ed_driver);
}

int edx_xas_fsolver(int comm, int my_id, int num_procs) {
    (void)comm;
    (void)num_procs;
    return run_rank0(my_id, xas_driver);
}

int edx_rixs_fsolver(int comm, int my_id, int num_procs) {
    (void)comm;
    (void)num_procs;
    return run_rank0(my_id, rixs_driver);
}

int edx_opavg_fsolver(int comm, int my_id, int num_procs) {
    (void)comm;
    (void)num_procs;
    return run_rank0(my_id, opavg_driver);
}

}  // extern "C"
