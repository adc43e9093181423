"""ctypes binding of libedx.so (the C ABI declared in include/edx.h).

The library is built in-tree (``edrixs_amd/libedx.so``).  There is no CPU
fallback: if the library or a HIP device is missing, calls raise.
"""
import ctypes
import os
import subprocess
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedx.so")

EDX_C128 = 0
EDX_C64 = 1
EDX_C64D = 2
EDX_C128I = 3

_STATUS = {0: "OK", 1: "invalid argument", 2: "HIP error", 3: "I/O error",
           4: "out of device memory", 5: "numerical failure", 6: "capacity exceeded"}


class EdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libedx error %d (%s): %s" % (code, _STATUS.get(code, "?"), msg))
        self.code = code


class Tensor2(ctypes.Structure):
    _fields_ = [("ind1", ctypes.c_int32), ("ind2", ctypes.c_int32),
                ("re", ctypes.c_double), ("im", ctypes.c_double)]


class Tensor4(ctypes.Structure):
    _fields_ = [("ind1", ctypes.c_int32), ("ind2", ctypes.c_int32), ("ind3", ctypes.c_int32),
                ("ind4", ctypes.c_int32), ("re", ctypes.c_double), ("im", ctypes.c_double)]


# (name, restype, argtypes); every int-returning entry point is status-checked
_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)
_SIGS = {
    "edx_last_error": (ctypes.c_char_p, []),
    "edx_init": (_I, [_I]),
    "edx_mem_in_use": (_I, [ctypes.POINTER(_I64)]),
    "edx_set_stream": (_I, [_P]),
    "edx_basis_combinations": (_I, [_I, _I, _PP]),
    "edx_basis_from_words": (_I, [_P, _I64, _I, _PP]),
    "edx_basis_size": (_I, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I)]),
    "edx_basis_download": (_I, [_P, _P]),
    "edx_basis_lookup": (_I, [_P, _P, _I64, _P]),
    "edx_basis_free": (None, [_P]),
    "edx_build_ham_i": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _D, _I, _PP]),
    "edx_build_ham_n": (_I, [_P, _I, _I, _P, _I64, _P, _I64, _D, _D, _D, _I, _PP]),
    "edx_build_transop_i": (_I, [_P, _P, _I, _I, _P, _I64, _I, _PP]),
    "edx_build_transop_f": (_I, [_P, _P, _I, _I, _P, _I64, _I, _PP]),
    "edx_csr_info": (_I, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                          ctypes.POINTER(_I64), ctypes.POINTER(_I)]),
    "edx_csr_download": (_I, [_P, _P, _P, _P]),
    "edx_csr_download_values": (_I, [_P, _I, _P]),
    "edx_csr_add_c64": (_I, [_P]),
    "edx_csr_prepare": (_I, [_P, _I]),
    "edx_build_krylov_batch": (_I, [_P, _I, _I, _P, _I, _P, _P, _P, _P]),
    "edx_build_krylov_reorth": (_I, [_P, _I, _P, _I, ctypes.POINTER(_I), _P, _P,
                                     ctypes.POINTER(_D)]),
    "edx_minres_shifted_batch": (_I, [_P, _I, _I, _P, _P, _P, _D, _I, _P, _P]),
    "edx_krylov_batch_start": (_I, [_P, _I, _I, _P, _I, _PP]),
    "edx_krylov_batch_advance": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "edx_krylov_batch_read": (_I, [_P, _I, _P, _P, _P, _P]),
    "edx_krylov_batch_keep": (_I, [_P, _I, _P]),
    "edx_krylov_batch_free": (None, [_P]),
    "edx_mf_ham_i": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _D, _PP]),
    "edx_mf_ham_n": (_I, [_P, _I, _I, _P, _I64, _P, _I64, _D, _D, _D, _PP]),
    "edx_mf_size": (_I, [_P, ctypes.POINTER(_I64)]),
    "edx_mf_apply": (_I, [_P, _P, _P]),
    "edx_build_krylov_mf": (_I, [_P, _P, _I, ctypes.POINTER(_I), _P, _P, ctypes.POINTER(_D)]),
    "edx_mf_free": (None, [_P]),
    "edx_group_create": (_I, [_I, _I, _PP]),
    "edx_group_add": (_I, [_P, _P, _P, ctypes.POINTER(_I)]),
    "edx_group_remove": (_I, [_P, _I]),
    "edx_group_advance": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "edx_group_read": (_I, [_P, _I, ctypes.POINTER(_I), _P, _P, ctypes.POINTER(_D),
                            ctypes.POINTER(_I)]),
    "edx_group_free": (None, [_P]),
    "edx_shifted_fom_project": (_I, [_P, _I, _P, _I, _P, _I, _PP, _PP, _D, _I, _P, _P]),
    "edx_shifted_fom_project_batch": (_I, [_P, _I, _I, _P, _I, _P, _PP, _PP, _D, _I, _P, _P]),
    "edx_csr_free": (None, [_P]),
    "edx_csr_spmv_lanes": (_I, [_P, _I, _I, ctypes.POINTER(_I)]),
    "edx_csr_factored_info": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I64),
                                   ctypes.POINTER(_I64)]),
    "edx_csr_factored_blocks": (_I, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I)]),
    "edx_spmv": (_I, [_P, _I, _P, _P]),
    "edx_spmm": (_I, [_P, _I, _I, _P, _P]),
    "edx_profile": (_I, [_I]),
    "edx_profile_read": (_I, [_I, _I, ctypes.POINTER(_I64), ctypes.POINTER(_D),
                              ctypes.POINTER(_D)]),
    "edx_profile_read_rhs": (_I, [_I, _I, _I, ctypes.POINTER(_I64), ctypes.POINTER(_D),
                                  ctypes.POINTER(_D)]),
    "edx_profile_read_alg": (_I, [_I, _I, _I, ctypes.POINTER(_D)]),
    "edx_profile_read_flops": (_I, [_I, _I, ctypes.POINTER(_D)]),
    "edx_set_corehole_layout": (_I, [_I, ctypes.POINTER(_I)]),
    "edx_lanczos_eigs": (_I, [_P, _I, _I, _I, _D, ctypes.c_uint64, _I, _I, _P, _P,
                              ctypes.POINTER(_I)]),
    "edx_lanczos_ground": (_I, [_P, _I, _I, _I, _D, ctypes.c_uint64, _I, _P, _P,
                                ctypes.POINTER(_I)]),
    "edx_trlan_eigs": (_I, [_P, _I, _I, _I, _I, _D, ctypes.c_uint64, _I, _I, _P, _P,
                            ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "edx_lanczos_refine": (_I, [_P, _I, _I, _I, _D, _I, _P, _I, _P, _P, ctypes.POINTER(_I)]),
    "edx_block_lanczos_eigs": (_I, [_P, _I, _I, _I, _D, ctypes.c_uint64, _I, _P, _P,
                                    ctypes.POINTER(_I), _P]),
    "edx_build_krylov": (_I, [_P, _I, _P, _I, ctypes.POINTER(_I), _P, _P,
                              ctypes.POINTER(_D)]),
    "edx_build_krylov_noalloc": (_I, [_P, _I, _P, _I, ctypes.POINTER(_I), _P, _P,
                                      ctypes.POINTER(_D)]),
    "edx_minres": (_I, [_P, _I, _P, _P, _D, _I, ctypes.POINTER(_I), ctypes.POINTER(_D)]),
    "edx_cf_spectrum": (_I, [_I, _P, _P, _D, _D, _D, _I, _P, _P, _P]),
    "edx_denmat": (_I, [_P, _I, _P, _P]),
    "edx_ed_fsolver": (_I, [_I, _I, _I]),
    "edx_xas_fsolver": (_I, [_I, _I, _I]),
    "edx_rixs_fsolver": (_I, [_I, _I, _I]),
    "edx_opavg_fsolver": (_I, [_I, _I, _I]),
}

EXPORTED = tuple(_SIGS)

_lock = threading.Lock()
_lib = None
_initialised = False


def build(force=False):
    """Compile libedx.so for gfx950 with hipcc (csrc/Makefile)."""
    args = ["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")]
    if force:
        subprocess.run(args + ["clean"], check=True)
    subprocess.run(args, check=True)


def load():
    """The raw ctypes library (loads, does not initialise a device)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise EdxError(3, "%s is not built; run __graft_entry__.build()" % LIB_PATH)
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def last_error():
    return load().edx_last_error().decode(errors="replace")


def check(rc):
    if rc != 0:
        raise EdxError(rc, last_error())
    return rc


def lib(device=0):
    """The library with a HIP device selected; raises if none is present."""
    global _initialised
    L = load()
    if not _initialised:
        check(L.edx_init(device))
        _initialised = True
    return L


def call(name, *args):
    return check(getattr(lib(), name)(*args))
