"""MI355X (gfx950) backend for the PCM->DSD hot path of foo_dsd_trellis.

Half-band FIR upsampling chain, Viterbi trellis SDM and Density-Aligned
Stitching as hand-written HIP kernels behind a C ABI (include/dsd_hip.h,
libdsdhip.so).  This package is the host-side mirror of the reference's
interfaces used by tests and benchmarks.
"""
from ._lib import DsdHipError, lib  # noqa: F401
