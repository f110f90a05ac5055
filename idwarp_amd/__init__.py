"""MI355X-native volume warp with IDWarp's USMesh API.

``USMesh`` (idwarp_amd.usmesh) mirrors idwarp.USMesh; the numerics run in the
HIP library libidwarp_hip.so (include/idwarp_hip.h) built for gfx950.
"""
__version__ = "0.1.0"

from .usmesh import USMesh, Error  # noqa: E402,F401
from .multiusmesh import MultiUSMesh  # noqa: E402,F401
