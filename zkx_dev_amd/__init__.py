"""gfx950-native radar hot path: polyphase filtering, FFT pulse compression,
Doppler FFT and CA-CFAR behind the reference's cusignal_cpp operator API.

The compute path is libzkx_radar.so (hand-written HIP for gfx950) reached
through its C ABI (include/zkx_radar.h); this package is the host-side mirror.
"""
from ._lib import ZkxError, lib  # noqa: F401

__version__ = "0.1.0"
