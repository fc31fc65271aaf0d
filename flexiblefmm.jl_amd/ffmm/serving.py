"""Serving: continuous batching of FMM solves on one device (libffmm_hip's server,
include/ffmm.h ``ffmm_server_*``).  No reference counterpart; the requests are the
reference's ``FMM(method, sources)`` + ``build_expansions`` + ``evaluate_at_sources`` /
``evaluate`` on host arrays, with a geometry key under which the plan (tree, lists,
buffers) stays resident so that a repeated geometry with new densities runs the
passes only (``build_expansions!``, src/fmm.jl:264-298).

    srv = Server(nslots=4, hbm_budget=64 << 30)
    t = srv.submit(method, cache, points, tau, targets=None, geometry=7)
    u, v, latency_ms = srv.wait(t)
"""
from __future__ import annotations

import ctypes
import struct
import math

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, call
from .fmm import (OperatorCache, Potential, StokesHarmonics, Yukawa3D, _is_harmonic, _output, context, square_bbox,
                  _M2L, _PREC)


class Request(ctypes.Structure):
    """``ffmm_request`` (include/ffmm.h)."""
    _fields_ = [("cache", ctypes.c_void_p), ("geometry", ctypes.c_int64), ("points", ctypes.c_void_p),
                ("n", ctypes.c_int64), ("ppl", ctypes.c_int), ("origin", ctypes.c_double * 3),
                ("width", ctypes.c_double), ("e0", ctypes.c_int), ("max_levels", ctypes.c_int),
                ("precision", ctypes.c_int), ("m2l", ctypes.c_int), ("dens", ctypes.c_void_p), ("nd", ctypes.c_int),
                ("targets", ctypes.c_void_p), ("nt", ctypes.c_int64), ("output", ctypes.c_int),
                ("out_sources", ctypes.c_void_p), ("out_targets", ctypes.c_void_p)]


def _host(a, cols=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if cols is not None and (a.ndim != 2 or a.shape[1] != cols):
        raise DimensionMismatch(f"expected an N × {cols} array, got {a.shape}")
    return a


class Server:
    """``nslots`` worker slots (each with its own HIP stream) sharing one device
    under an ``hbm_budget`` of device bytes."""

    def __init__(self, nslots: int = 4, hbm_budget: int = 64 << 30, device: int = 0, yukawa_levels: int = 16):
        context(device)
        self.device = device
        self.yukawa_levels = int(yukawa_levels)
        h = ctypes.c_void_p()
        call("ffmm_server_create", device, int(nslots), int(hbm_budget), ctypes.byref(h))
        self._h = h
        self._live = {}

    def _prepare(self, cache: OperatorCache, lo, hi, bbox):
        origin, W, e0 = cache.fit_box(lo, hi, bbox)
        oc = cache.device(self.device)
        m = cache.method
        if _is_harmonic(m) and isinstance(m.kernel, Yukawa3D):
            # every level a request can reach: the server does not call back into the host
            cache.ensure_widths(self.device, [math.ldexp(W, -l) for l in range(2, self.yukawa_levels)])
        return oc, origin, W, e0

    def submit(self, cache: OperatorCache, points, dens, targets=None, *, geometry: int = -1, ppl: int = 64,
               bbox=None, output=Potential, precision: str = "f64", m2l: str = "dense", max_levels: int = 0) -> int:
        m = cache.method
        if _is_harmonic(m) and m2l != "dense":
            raise ArgumentError("the harmonic methods take the dense class M2L")
        pts = _host(points, 3)
        d = _host(dens)
        n = len(pts)
        if d.shape[0] != n:
            raise DimensionMismatch(f"expected {n} densities")
        nd = 1 if d.ndim == 1 else d.shape[1]
        o = _output(output)
        nc = (1 if o.code == _lib.PRESSURE else 3) if isinstance(m, StokesHarmonics) else o.ncomp
        # per coordinate: numpy's axis-0 reduction of an N x 3 array is several times slower
        lo = np.array([pts[:, d].min() for d in range(3)])
        hi = np.array([pts[:, d].max() for d in range(3)])
        if bbox is None:
            bbox = square_bbox(lo, hi)
        oc, origin, W, e0 = self._prepare(cache, lo, hi, bbox)
        out_s = np.empty((n, nc) if nc > 1 else n)
        tg = None if targets is None else _host(targets, 3)
        out_t = None if tg is None else np.empty((len(tg), nc) if nc > 1 else len(tg))
        P = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        r = Request(cache=oc, geometry=int(geometry), points=P(pts), n=n, ppl=int(ppl),
                    origin=(ctypes.c_double * 3)(*origin), width=W, e0=e0, max_levels=int(max_levels),
                    precision=_PREC[precision], m2l=_M2L[m2l], dens=P(d), nd=nd, targets=P(tg),
                    nt=0 if tg is None else len(tg), output=o.code, out_sources=P(out_s), out_targets=P(out_t))
        t = ctypes.c_int64()
        call("ffmm_server_submit", self._h, ctypes.addressof(r), ctypes.byref(t))
        self._live[t.value] = (pts, d, tg, out_s, out_t, r)   # the buffers stay alive until wait
        return t.value

    def wait(self, ticket: int):
        lat = ctypes.c_double()
        rc = _lib.lib().ffmm_server_wait(self._h, ticket, ctypes.byref(lat))
        keep = self._live.pop(ticket, None)
        _lib.check(rc)
        _, _, _, out_s, out_t, _ = keep
        return out_s, out_t, lat.value

    def export_plan(self, cache: OperatorCache, geometry: int) -> bytes:
        """The plan data of the resident plan of (cache, geometry) (``ffmm_server_export``)."""
        oc = cache.device(self.device)
        size = ctypes.c_int64()
        call("ffmm_server_export", self._h, oc, int(geometry), None, 0, ctypes.byref(size))
        buf = (ctypes.c_uint8 * size.value)()
        call("ffmm_server_export", self._h, oc, int(geometry), buf, size.value, ctypes.byref(size))
        return bytes(buf)

    def import_plan(self, cache: OperatorCache, geometry: int, data: bytes) -> None:
        """Make plan data (``export_plan``, ``DevicePlan.save``'s payload) the resident plan
        of (cache, geometry) (``ffmm_server_import``): a warm restart of the resident set."""
        w_ref = struct.unpack_from("<d", data, 36)[0]   # the cache identity's reference width
        if not cache.has_base_width():
            cache.W0 = w_ref
        elif cache.W0 != w_ref:
            raise ArgumentError("the plan data was made with another reference width")
        oc = cache.device(self.device)
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        call("ffmm_server_import", self._h, oc, int(geometry), buf, len(data))

    def stats(self) -> dict:
        v = (ctypes.c_int64 * 12)()
        call("ffmm_server_stats", self._h, v, 12)
        keys = ["completed", "plans_built", "plans_reused", "resident", "bytes_in_use", "peak_bytes", "evictions",
                "queued", "idle_cached_bytes", "worst_plan_over_estimate_permille", "requeued_out_of_memory",
                "allocation_flushes"]
        return dict(zip(keys, v[:]))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("ffmm_server_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
