"""MI355X flip-graph search engine: Python access to libflipgraph.so.

The search itself runs in HIP kernels behind the C ABI in include/fg.h;
this package only binds that ABI for tests and benchmarks.
"""
from ._lib import (RECORD_DTYPE, RING_Z2, RING_ZT, Engine, FgConfig, FgError, Reducer, Scorer, device_info,
                   load_library, valu_probe, verify, verify_host, complexity_host)

__all__ = ["Engine", "FgConfig", "FgError", "RECORD_DTYPE", "RING_Z2", "RING_ZT", "Reducer", "Scorer", "device_info",
           "load_library", "valu_probe", "verify", "verify_host", "complexity_host"]
