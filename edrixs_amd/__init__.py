"""edrixs_amd: the EDRIXS exact-diagonalization core on AMD MI355X (gfx950).

* ``edrixs_amd.fedrixs``: drop-in for the f2py module ``edrixs.fedrixs``
  (``ed_fsolver``, ``xas_fsolver``, ``rixs_fsolver``, ``opavg_fsolver``).
* ``edrixs_amd.core``: device-resident bases, CSR operators and solvers.
* ``edrixs_amd.iostream`` / ``fock_basis`` / ``spectra``: the file formats and
  post-processing the reference's Python side wraps around the solvers.
* ``edrixs_amd.siam`` / ``solvers``: synthetic impurity models and the
  orchestration that drives the drop-in the way ``edrixs.solvers`` does.
"""
import sys

from ._lib import EdxError  # noqa: F401


def install_as_fedrixs():
    """Make ``from edrixs.fedrixs import ed_fsolver`` resolve to this package."""
    from . import fedrixs
    sys.modules["edrixs.fedrixs"] = fedrixs
    parent = sys.modules.get("edrixs")
    if parent is not None:
        setattr(parent, "fedrixs", fedrixs)
    return fedrixs
