"""pcbqs_amd: MI355X-native (gfx950) static fill and loop/port solve for pcbqs.

The `mi355x` pcbqs.static_backends plugin (pcbqs_amd.backend.make_backend) and the
device-resident pipeline behind it. Hand-written HIP kernels in pcbqs_amd/csrc are reached
through the C ABI of libpcbqs_hip.so (include/pcbqs_amd.h) with ctypes.
"""
__version__ = '0.1.0'
