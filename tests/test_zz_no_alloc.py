"""The no-allocation Krylov entry point against edx_build_krylov.

edx_build_krylov_noalloc runs the same recurrence with its device memory in
the library's static scratch array, so on a prepared matrix it must return
the same (alpha, beta, norm) bit for bit.  Two calls in a row, the second on
a smaller input (the array keeps the first call's contents), then a request
beyond the array's size, which must fail with a clear error.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _krylov(core, entry, H, precision, v0, maxn):
    alpha = np.zeros(maxn)
    beta = np.zeros(maxn)
    neff = ctypes.c_int()
    norm = ctypes.c_double()
    core._lib.call(entry, H._h, core._prec(precision), core._dptr(v0), maxn, ctypes.byref(neff),
                   core._ptr(alpha), core._ptr(beta), ctypes.byref(norm))
    n = neff.value
    return alpha[:n], beta[:n], norm.value


def test_noalloc_krylov_equals_allocating_entry_point():
    import torch
    from edrixs_amd import core, siam
    from tests import helpers as H
    cases = [(6, "c128i", 300), (4, "c128", 120)]     # norb 16 then norb 14: smaller
    for nbath, precision, maxn in cases:
        m = H.small_siam(nbath=nbath, seed=5)
        bn = core.Basis.combinations(m.norb, m.nocc + 1)
        Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout="interleaved")
        Hn.prepare(precision)
        g = torch.Generator(device="cuda").manual_seed(nbath)
        v0 = torch.randn(Hn.n, dtype=torch.complex128, device="cuda", generator=g)
        used = core.mem_in_use()
        a1, b1, n1 = _krylov(core, "edx_build_krylov_noalloc", Hn, precision, v0, maxn)
        assert core.mem_in_use() == used          # no device allocation by the library
        a2, b2, n2 = _krylov(core, "edx_build_krylov", Hn, precision, v0, maxn)
        assert len(a1) == len(a2) == maxn
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and n1 == n2
    # an unprepared form is refused (preparing it would allocate)
    m = H.small_siam(nbath=2, seed=5)
    bn = core.Basis.combinations(m.norb, m.nocc + 1)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    v0 = torch.ones(Hn.n, dtype=torch.complex128, device="cuda")
    with pytest.raises(Exception, match="not prepared"):
        _krylov(core, "edx_build_krylov_noalloc", Hn, "c64d", v0, 10)
    # more than the static array holds: a clear error, nothing run
    with pytest.raises(Exception, match="bytes of scratch"):
        _krylov(core, "edx_build_krylov_noalloc", Hn, "c128", v0, 1 << 23)
