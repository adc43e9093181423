"""Drop-in replacement of the f2py module ``edrixs.fedrixs`` (src/pyapi.f90).

Same four callables, same three integer arguments, same input and output
files in the current directory; the work runs on the MI355X through libedx.
Where the reference would ``STOP`` the interpreter on bad input, these raise
:class:`edrixs_amd.EdxError`.  Install under the reference's module name with
:func:`edrixs_amd.install_as_fedrixs`.
"""
from . import _lib


def _run(name, comm, my_id, num_procs):
    _lib.call(name, int(comm), int(my_id), int(num_procs))


def ed_fsolver(comm, my_id, num_procs):
    """src/pyapi.f90:1-56 -> eigvals.dat, denmat.dat, eigvec.N (idump)."""
    _run("edx_ed_fsolver", comm, my_id, num_procs)


def xas_fsolver(comm, my_id, num_procs):
    """src/pyapi.f90:58-118 -> xas_poles.N."""
    _run("edx_xas_fsolver", comm, my_id, num_procs)


def rixs_fsolver(comm, my_id, num_procs):
    """src/pyapi.f90:120-182 -> rixs_poles.N."""
    _run("edx_rixs_fsolver", comm, my_id, num_procs)


def opavg_fsolver(comm, my_id, num_procs):
    """src/pyapi.f90:184-240 -> opavg.dat."""
    _run("edx_opavg_fsolver", comm, my_id, num_procs)
