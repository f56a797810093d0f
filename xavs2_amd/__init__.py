"""MI355X-native motion-estimation cost path for the AVS2 encoder xavs2.

The product is the C-ABI library ``libxavs2_amd.so`` (declared in
``include/xavs2_amd.h``).  This package holds its HIP sources (``csrc/``) and a
thin ctypes mirror of the reference's primitive interface used by the tests
and the benchmark.
"""
