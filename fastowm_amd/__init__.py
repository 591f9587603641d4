"""MI355X-native OWM (Overlapping Window Minimum) ground-seed pipeline.

Host interface: fastowm_amd.owm (ctypes over the C ABI in include/owm.h).
Synthetic workloads: fastowm_amd.synth.  Build: fastowm_amd.build.
"""
from .owm import NONE, Context, OWMError, Tree, derive, lib, make_params  # noqa: F401

__all__ = ["Context", "Tree", "OWMError", "derive", "make_params", "lib", "NONE"]
