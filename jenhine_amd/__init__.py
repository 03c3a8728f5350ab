"""jenhine_amd -- MI355X (gfx950) FDK reconstruction back end.

Native parts: ``libjhip.so`` (HIP kernels behind the C-ABI of include/jhip.h),
``libjhsynth.so`` (synthetic scan generator) and the C++ agent/pipe host under
``host/``.  Build them with ``python -m jenhine_amd.build``.
"""
from . import configs, geometry  # noqa: F401

__all__ = ["configs", "geometry"]
