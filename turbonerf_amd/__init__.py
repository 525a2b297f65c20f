"""TurboNeRF training/rendering hot path, native to AMD Instinct MI355X (gfx950).

Layers:
  * ``lib/libturbonerf_kernels.so`` -- hand-written HIP kernels behind the
    ``tn_*`` C ABI (include/turbonerf/tn_c_api.h), bound here by ``_capi``.
  * ``lib/libturbonerf.so`` + ``PyTurboNeRF`` -- the C++ host layer and the
    pybind11 module mirroring the reference's Python API.
  * ``ops`` -- torch.autograd.Function wrappers over the C ABI.
"""
__version__ = "0.1.0"
