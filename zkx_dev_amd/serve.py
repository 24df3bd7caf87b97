"""Continuous-batching dwell server (C ABI zkx_server_*).

Dwells stay on the device: the caller owns a flat complex64 input pool and
submits requests (offset, P, R, template id).  The server admits them in
arrival order under an HBM budget, batches whatever is queued when a lane
frees up, runs each batch through the batch engine on its own stream and
reports per-request latency.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .radar import _ptr, _stream


class Server:
    def __init__(self, engine, pool, budget_bytes, lanes=2, max_batch=64, keep_outputs=False):
        if not isinstance(pool, torch.Tensor) or not pool.is_cuda or pool.dtype != torch.complex64:
            raise ValueError("pool must be a complex64 device tensor")
        if lanes > engine.lanes:
            raise ValueError("the engine needs at least as many lanes as the server")
        self.engine = engine
        self.pool = pool.contiguous()
        h = ctypes.c_void_p()
        check(lib().zkx_server_create(ctypes.byref(h), engine._h, _ptr(self.pool), self.pool.numel(),
                                      int(budget_bytes), int(lanes), int(max_batch), int(bool(keep_outputs))))
        self._h = h
        self.keep = keep_outputs

    def close(self):
        if self._h:
            check(lib().zkx_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, offset, P, R, template_id):
        t = ctypes.c_longlong()
        check(lib().zkx_server_submit(self._h, int(offset), int(P), int(R), int(template_id), ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        done = ctypes.c_int()
        check(lib().zkx_server_wait(self._h, int(ticket), int(timeout_ms), ctypes.byref(done)))
        return bool(done.value)

    def result(self, ticket, P=None, R=None, stream=None):
        """(latency_us, threshold, detections); with keep_outputs and a shape
        the outputs are copied into new [P, R] device tensors."""
        lat = ctypes.c_double()
        check(lib().zkx_server_result(self._h, int(ticket), ctypes.byref(lat), None, None))
        thr = det = None
        if self.keep and P is not None:
            thr = torch.empty((P, R), dtype=torch.float32, device=self.pool.device)
            det = torch.empty((P, R), dtype=torch.uint8, device=self.pool.device)
            check(lib().zkx_server_copy_result(self._h, int(ticket), _ptr(thr), _ptr(det), _stream(stream)))
        return lat.value, thr, det

    def hold(self, on=True):
        """Hold (or resume) batch formation: requests submitted while held
        are launched together when released."""
        check(lib().zkx_server_hold(self._h, int(bool(on))))

    def release(self, ticket):
        check(lib().zkx_server_release(self._h, int(ticket)))

    def stats(self):
        s = np.zeros(5, np.int64)
        check(lib().zkx_server_stats(self._h, ctypes.c_void_p(s.ctypes.data)))
        return dict(zip(["batches", "dwells_done", "queued", "reserved_bytes", "peak_reserved_bytes"],
                        map(int, s)))

    def replay(self, arrival_us, offsets, pulses, ranges, templates):
        """Open-loop load: request i submitted at arrival_us[i]; returns the
        submit-to-completion latency of each request in microseconds."""
        a = np.ascontiguousarray(arrival_us, np.float64)
        o = np.ascontiguousarray(offsets, np.int64)
        p = np.ascontiguousarray(pulses, np.int32)
        r = np.ascontiguousarray(ranges, np.int32)
        t = np.ascontiguousarray(templates, np.int32)
        lat = np.zeros(len(a), np.float64)
        vp = lambda v: ctypes.c_void_p(v.ctypes.data)  # noqa: E731
        check(lib().zkx_server_replay(self._h, vp(a), vp(o), vp(p), vp(r), vp(t), len(a), vp(lat)))
        return lat


__all__ = ["Server"]
