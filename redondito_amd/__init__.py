"""MI355X-native pseudospectral transform path of REDONDITO (sphp + pseudo).

The compute lives in the HIP library redondito_amd/_lib/librdsphp.so
(built by redondito_amd/build.py); redondito_amd.sphp mirrors the
reference's Fortran interface over its C-ABI (include/redondito.h).
"""
from . import sphp  # noqa: F401
from .sphp import FCPlan, RdError, grid_nlat  # noqa: F401
