"""saclaymocks_amd: MI355X-native kernels for the SaclayMocks hot path.

- box:      GRF box stage (bin/make_boxes.py) - fused R2C / sqrt(P/Vcell) / C2R
- util:     reference helpers used on the path (fgpa, fgrowth, cosmo, ...)

Compute runs in libsaclay_hip.so (hand-written HIP for gfx950) through a
ctypes C ABI; torch tensors are the device buffers.
"""
__version__ = "0.1.0"
