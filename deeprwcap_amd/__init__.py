"""MI355X-native deepRWCap transition-kernel path.

The product is the C-ABI library ``libdnnsolver.so`` built from ``csrc/``
(HIP kernels for gfx950 + C++ host code); this package holds the ctypes host
mirror of the reference interface and the synthetic-structure generators.
"""
