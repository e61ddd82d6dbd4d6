"""pdds_amd: MI355X-native PDDSparse path engine.

The compute path is libpdds.so (HIP kernels for gfx950 behind the C ABI of
include/pdds_engine.h).  This package is the host-side mirror used from Python.
"""
from ._capi import lib, PddsError  # noqa: F401
