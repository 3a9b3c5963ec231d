"""tncontract_amd -- MI355X (gfx950) contraction-plus-truncation path of
tncontract behind the same Python API.

Import order matters: torch is imported first so the HIP library binds to the
HIP runtime torch loaded (one runtime per process).
"""
import torch  # noqa: F401  (device buffers / streams; must precede the HIP library)

from tncontract_amd.version import __version__
from tncontract_amd.tensor import *  # noqa: F401,F403
from tncontract_amd.tensor import tensor_qr, tensor_lq, conjugate  # noqa: F401
from tncontract_amd.label import *  # noqa: F401,F403
from tncontract_amd.tncon import con  # noqa: F401
from tncontract_amd import onedim  # noqa: F401
from tncontract_amd import twodim  # noqa: F401
from tncontract_amd import _lib  # noqa: F401


def native_library():
    """Path of the loaded HIP library (loads it)."""
    _lib.lib()
    return _lib.LIB_PATH
