"""paintop_amd -- the MI355X (gfx950) ``hip`` backend for paintop's Poisson /
screened-Poisson solver ops (repair.poisson_blend@1, repair.screened_poisson@1).

Layers:
  csrc/        hand-written HIP kernels and the C ABI (libpaintop_hip.so)
  hip.py       ctypes binding of the C ABI (no CPU fallback)
  ops.py       the two ops' OpImplementation on the GPU
  executor.py  the registry / backend-policy boundary the backend plugs into
  inputs.py    synthetic workloads generated on the device
"""
from .errors import ErrorClass, PaintopError  # noqa: F401
from .ops import (  # noqa: F401
    HIP_IMPL_ID,
    POISSON_BLEND_OP_ID,
    SCREENED_POISSON_OP_ID,
    PoissonBlend,
    ScreenedPoisson,
)
from .values import ResourceValue, image, mask  # noqa: F401
