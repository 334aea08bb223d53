"""Host-side Python bindings of the MI355X wavefront path tracer.

The product is the C-ABI library ``liblotus_hip.so`` (``include/lotus_hip.h``)
built from ``hydra-lotus_amd/csrc``; this package only loads it through ctypes
and builds scene fixtures (``scenes``) for tests, smoke and bench.
"""
