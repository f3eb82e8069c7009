"""MI355X-native SAXS scattering engine (drop-in for pyCuSaxs's per-frame
path).  Hand-written HIP kernels for gfx950 behind a C ABI
(include/saxs_engine.h); Python host code mirrors pycusaxs_cuda.run."""

__version__ = "0.1.0"
