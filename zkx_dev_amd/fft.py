"""The reference's FFT backend API (fft_* C ABI, include/zkx_fft.h) on torch
device tensors: plan / execute / destroy, shifts and sample frequencies.

Semantics follow ZKX/cusignal_cpp/include/cusignal/backends/fft/fft_interface.h:
forward unnormalised, inverse scaled by 1/N, row-major complex64 batches.
"""
from __future__ import annotations

import ctypes

import torch

from ._check import out_tensor

from ._lib import lib

FFT_FORWARD, FFT_INVERSE = 1, -1
FFT_TYPE_C2C, FFT_TYPE_R2C, FFT_TYPE_C2R = 0, 1, 2
_FFT_SIGS = {
    "fft_plan_create_1d": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int]),
    "fft_plan_create_2d": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int]),
    "fft_execute": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fft_execute_on_stream": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "fft_execute_inplace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fft_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fft_shift": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "fft_inverse_shift": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "fft_shift_2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "fft_inverse_shift_2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "fft_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "fft_version": (ctypes.c_int, []),
    "fft_freq": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_float]),
    "fft_rfftfreq": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_float]),
}
_bound = False


def _L():
    global _bound
    L = lib()
    if not _bound:
        for name, (res, args) in _FFT_SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


class FFTError(RuntimeError):
    def __init__(self, code):
        self.code = code
        super().__init__(_L().fft_error_string(code).decode())


def _check(code):
    if code != 0:
        raise FFTError(code)


class Plan:
    """fft_plan_create_1d / fft_plan_create_2d, fft_execute, fft_plan_destroy.

    1-D: `nx` points, `batch` rows, type C2C (complex64 rows), R2C (float32
    rows in, nx/2+1 complex64 bins out) or C2R (nx/2+1 bins in, float32 rows
    out).  2-D (ny given): a complex64 [ny, nx] array, as the reference's
    2-D plans transform complex data whatever their type."""

    def __init__(self, nx, batch=1, direction=FFT_FORWARD, type=FFT_TYPE_C2C, ny=None):
        h = ctypes.c_void_p()
        if ny is None:
            _check(_L().fft_plan_create_1d(ctypes.byref(h), int(nx), int(batch), int(type), int(direction)))
        else:
            _check(_L().fft_plan_create_2d(ctypes.byref(h), int(nx), int(ny), int(type), int(direction)))
        self.handle, self.nx, self.batch, self.type, self.ny = h, int(nx), int(batch), int(type), ny

    def _shapes(self):
        if self.ny is not None:
            return (torch.complex64, self.nx * self.ny), (torch.complex64, self.nx * self.ny)
        half = self.nx // 2 + 1
        if self.type == FFT_TYPE_R2C:
            return (torch.float32, self.nx * self.batch), (torch.complex64, half * self.batch)
        if self.type == FFT_TYPE_C2R:
            return (torch.complex64, half * self.batch), (torch.float32, self.nx * self.batch)
        return (torch.complex64, self.nx * self.batch), (torch.complex64, self.nx * self.batch)

    def execute(self, x, out=None, stream=None):
        (din, nin), (dout, nout) = self._shapes()
        if x.dtype != din or not x.is_cuda or x.numel() != nin:
            raise ValueError(f"input must be {din} with {nin} elements on the device")
        x = x.contiguous()
        out = torch.empty(nout, dtype=dout, device=x.device) if out is None else out
        out_tensor(out, "out", x, dout, nout)
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(_L().fft_execute_on_stream(self.handle, out.data_ptr(), x.data_ptr(), s.cuda_stream))
        return out

    def destroy(self):
        if self.handle:
            _check(_L().fft_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def fft(x, inverse=False):
    """Batched C2C transform along the last axis (any length the backend takes)."""
    n = x.shape[-1]
    p = Plan(n, x.numel() // n, FFT_INVERSE if inverse else FFT_FORWARD)
    try:
        return p.execute(x).reshape(x.shape)
    finally:
        p.destroy()


def rfft(x):
    """R2C along the last axis of float32 rows: n/2+1 bins."""
    n = x.shape[-1]
    p = Plan(n, x.numel() // n, FFT_FORWARD, FFT_TYPE_R2C)
    try:
        return p.execute(x).reshape(tuple(x.shape[:-1]) + (n // 2 + 1,))
    finally:
        p.destroy()


def irfft(X, n):
    """C2R of n/2+1 bins per row into n real samples (reference semantics)."""
    rows = X.numel() // (n // 2 + 1)
    p = Plan(n, rows, FFT_INVERSE, FFT_TYPE_C2R)
    try:
        return p.execute(X).reshape(tuple(X.shape[:-1]) + (n,))
    finally:
        p.destroy()


def fft2(x, inverse=False):
    """2-D C2C transform of a complex64 [ny, nx] array."""
    ny, nx = x.shape
    p = Plan(nx, 1, FFT_INVERSE if inverse else FFT_FORWARD, ny=ny)
    try:
        return p.execute(x).reshape(ny, nx)
    finally:
        p.destroy()


def fftshift(x, inverse=False):
    """fft_shift / fft_inverse_shift over the last axis of complex64 rows."""
    x = x.contiguous()
    out = torch.empty_like(x)
    n = x.shape[-1]
    fn = _L().fft_inverse_shift if inverse else _L().fft_shift
    _check(fn(out.data_ptr(), x.data_ptr(), n, x.numel() // n))
    return out


def fftshift2d(x, inverse=False):
    """fft_shift_2d / fft_inverse_shift_2d of a complex64 [ny, nx] array."""
    x = x.contiguous()
    ny, nx = x.shape
    out = torch.empty_like(x)
    fn = _L().fft_inverse_shift_2d if inverse else _L().fft_shift_2d
    _check(fn(out.data_ptr(), x.data_ptr(), nx, ny))
    return out


def fftfreq(n, d=1.0, rfft=False, device=None):
    """fft_freq / fft_rfftfreq into a host (device=None) or device tensor."""
    size = n // 2 + 1 if rfft else n
    out = torch.empty(size, dtype=torch.float32, device=device)
    fn = _L().fft_rfftfreq if rfft else _L().fft_freq
    _check(fn(out.data_ptr(), int(n), float(d)))
    return out
