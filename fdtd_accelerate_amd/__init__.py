"""fdtd_accelerate_amd — MI355X-native FDTD time-step engine.

The Yee update_e/update_h curl stencils, CPML psi updates on the boundary slabs and
source injection (plus the 4th-order symplectic variant), as hand-written HIP kernels
for gfx950 behind the KEMP/kuFDTD call surface of the reference
(wbkifun/fdtd_accelerate).  Python drives them through torch tensors and a ctypes
C ABI (``include/mifdtd.h``).
"""
from .util import check_type, check_value
from .fields import Fields
from .core import Core, CoreSplit
from .cpml import Cpml, cpml_coefficients, sigma_max
from .incident import IncidentDirect, GetFields, SetFields
from .pbc import Pbc
from .tfsf import Tfsf
from .drude import Drude
from .sources import gaussian_pulse
from . import geometry

__all__ = [
    "Fields", "Core", "CoreSplit", "Cpml", "IncidentDirect", "GetFields", "SetFields", "Pbc", "Tfsf", "Drude",
    "cpml_coefficients", "sigma_max", "gaussian_pulse", "geometry",
    "check_type", "check_value",
]
