"""mr_aha_amd — MI355X-native Toeplitz A^H A (drop-in for the reference's
``method="toeplitz"`` path on one GPU).

The compute lives in ``libmr_aha_c.so`` (hand-written HIP kernels for
gfx950 behind the reference's C ABI); this package only binds it.
"""

from ._lib import MRError, TrajType, Triangulation, backend, LIB_PATH  # noqa: F401
from .api import (AhAOp, DEFAULT_MASK_RADIUS, MRConfig, aha_toeplitz, calc_mask,  # noqa: F401
                  calc_toep, calc_toep_from_psf, create_op, make_config, num_tri,
                  phi2_toep_shape)

__version__ = "1.0.0"
