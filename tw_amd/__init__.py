"""tw_amd -- MI355X-native paDIC hot path (FFT-CC + first-order IC-GN).

The product is the HIP/C++ library tw_amd/lib/libtwdic.so behind the C ABI in
include/twdic.h and the TW_Engine-compatible C++ classes in include/tw/.
This Python package only binds that ABI (tw_amd.api) for tests and
benchmarks.
"""
from . import api  # noqa: F401

__all__ = ["api"]
