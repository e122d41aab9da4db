"""MI355X-native AMUN per-block finite-volume update.

The compute path is the HIP library ``libamun_hip.so`` built from
``amun_amd/csrc`` (hand-written gfx950 kernels behind the C-ABI of
``include/amun_dev.h``).  This package holds the Python side of that
boundary: the parameter mapping (``config``), seeded initial states
(``problems``), and the ctypes binding of the C-ABI (``solver``).
"""
from .config import Params, ConfigError, c1_params, c2_params, c4_params  # noqa: F401
