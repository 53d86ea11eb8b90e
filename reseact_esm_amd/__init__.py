"""MI355X-native operator-split step of ReSEACT (EarthSciML/reseact.esm):
SSPRK43 flux-form transport, per-cell Rosenbrock23 chemistry and their
discrete adjoint, as hand-written HIP kernels for gfx950 behind a C ABI
(include/reseact_hip.h, libreseact_hip.so)."""
from . import problems  # noqa: F401
from .host import (Context, Library, ReseactError, SplitOpts, hip_library,  # noqa: F401
                   pictrl_ros23, pictrl_ssprk43, split_opts)
