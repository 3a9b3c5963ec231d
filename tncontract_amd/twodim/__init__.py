"""Two-dimensional (square-lattice) tensor networks on the GPU
(tncontract/twodim)."""
from tncontract_amd.twodim.square_lattice import *  # noqa: F401,F403
