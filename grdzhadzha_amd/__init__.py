"""MI355X-native fixed-background GRHD operator (GRDzhadzha hot path).

The product is the HIP library grdzhadzha_amd/libgrhd_hip.so behind the
C-ABI of include/grhd_abi.h.  This package holds its host-side mirror of the
reference's level-operator interface (operator.py), the option records
(options.py), the ctypes view of the ABI (abi.py), the benchmark workloads
(workloads.py) and the build recipe (build.py).
"""
from . import abi  # noqa: F401
from .options import (AtmosphereOptions, BoxSpec, CircularBinaryPNParams,  # noqa: F401
                      KerrSchildParams, OperatorOptions, RecoveryOptions,
                      cube_box)
