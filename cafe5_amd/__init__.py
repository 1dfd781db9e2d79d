"""MI355X-native likelihood hot path of CAFE5 (matrix build + Felsenstein pruning).

The compute path is the HIP library built from cafe5_amd/csrc (libcafe_gpu.so,
C ABI declared in include/cafe_gpu.h).  cafe5_amd.native loads it with ctypes
and fails loudly when it is missing.
"""
