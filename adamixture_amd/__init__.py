"""MI355X-native backend for ADAMIXTURE's `--device gpu` likelihood path.

Host code mirrors the reference's model/entry API (adamixture/model/*_gpu.py,
adamixture/src/utils.py, adamixture/src/adamixture.py); the work runs in
hand-written gfx950 kernels of libadx.so (C ABI: include/adx.h).
"""
__version__ = "0.1.0"
