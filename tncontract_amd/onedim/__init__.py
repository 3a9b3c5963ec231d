"""One-dimensional tensor networks on the GPU (tncontract/onedim)."""
from tncontract_amd.onedim.onedim_core import *  # noqa: F401,F403
from tncontract_amd.onedim.onedim_utils import *  # noqa: F401,F403
