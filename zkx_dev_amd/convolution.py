"""Host-side mirror of the reference's correlate API on torch device tensors.

Follows ZKX/cusignal_cpp/include/cusignal/operators/convolution/
convolution_typed.h (correlate_device(in1, in2, out, mode), the direct
method) in cusignal.correlate's call form: 1-D inputs of one element type
(float32, float16, int32, int16, int8), output in that type.  The direct
method keeps the reference's arithmetic bit for bit (an FP32 fma chain for
floating types, sums modulo 2^bits for integers).  All compute runs in
libzkx_radar.so.
"""
from __future__ import annotations

import torch

from ._lib import check, lib
from .radar import _ptr, _stream

_DTYPES = {torch.float32: 0, torch.float16: 1, torch.int32: 2, torch.int16: 3, torch.int8: 4}
_MODES = {"full": 0, "same": 1, "valid": 2}


def correlate(in1, in2, mode="full", method="auto", *, out=None, stream=None):
    """Cross-correlation of two 1-D device tensors of the same dtype.
    method: "direct" or "auto" (the reference's choice for 1-D is direct)."""
    if mode not in _MODES:
        raise ValueError("typed correlate mode must be full, same, or valid")
    if method not in ("direct", "auto", "fft"):
        raise ValueError("typed correlate method must be direct, fft, or auto")
    if method == "fft":
        raise NotImplementedError("correlate: only the direct method is built")
    for t, name in ((in1, "in1"), (in2, "in2")):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
    if in1.dtype != in2.dtype or in1.dtype not in _DTYPES:
        raise ValueError("correlate inputs must share one of float32, float16, int32, int16, int8")
    if in1.dim() != 1 or in2.dim() != 1:
        raise ValueError("correlate direct method is only implemented for 1D")
    a, b = in1.contiguous(), in2.contiguous()
    n = lib().zkx_correlate_out_len(a.numel(), b.numel(), _MODES[mode])
    if out is None:
        out = torch.empty(max(n, 0), dtype=a.dtype, device=a.device)
    elif out.numel() != n or out.dtype != a.dtype:
        raise ValueError("typed correlate output shape mismatch")
    code = lib().zkx_correlate_direct(_ptr(a), a.numel(), _ptr(b), b.numel(), _DTYPES[a.dtype], _MODES[mode],
                                      _ptr(out), _stream(stream))
    if code in (1001, 1002):
        raise ValueError(lib().zkx_last_error_message().decode())
    check(code)
    return out


__all__ = ["correlate"]
