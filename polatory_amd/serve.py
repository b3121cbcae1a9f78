"""Request loop over fitted interpolants kept resident in HBM (continuous batching).

polatory has no server: the closest reference call is ``Interpolant::evaluate``
per request (interpolant.hpp:54-67), which rebuilds its evaluator each time.
Here every model's source tree, weights and multipole expansions stay in HBM;
requests are admitted in arrival order whenever a slot is free and the HBM
budget holds their working set, and the admitted requests run concurrently,
one stream per slot.
"""
import ctypes

import numpy as np

from . import _capi
from .fmm import FmmConfig, _points


class Server:
    def __init__(self, n_slots=4, hbm_budget=0, rel_accuracy=1e-6):
        o = _capi.PamServerOptions()
        _capi.call("pam_server_options_default", ctypes.byref(o))
        o.n_slots = int(n_slots)
        o.hbm_budget = int(hbm_budget)
        o.rel_accuracy = float(rel_accuracy)
        h = ctypes.c_void_p()
        _capi.call("pam_server_create", ctypes.byref(o), ctypes.byref(h))
        self._h = h
        self._pending = {}  # request id -> (points, out) kept alive until completion

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _capi.call("pam_server_destroy", h)
        self._h = None
        self._pending = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_model(self, interpolant, bbox=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), config: FmmConfig = None):
        lo = (ctypes.c_double * 3)(*bbox[0])
        hi = (ctypes.c_double * 3)(*bbox[1])
        mid = ctypes.c_int()
        cfg = None if config is None else ctypes.byref(config.to_c())
        _capi.call("pam_server_add_model", self._h, interpolant._h, lo, hi, cfg, ctypes.byref(mid))
        return mid.value

    def model_config(self, model_id):
        c = _capi.PamFmmConfig()
        b = ctypes.c_int64()
        _capi.call("pam_server_model_config", self._h, model_id, ctypes.byref(c), ctypes.byref(b))
        return FmmConfig.from_c(c), b.value

    def submit(self, model_id, points):
        p = _points(points)
        out = np.empty(p.shape[0])
        rid = ctypes.c_int64()
        _capi.call("pam_server_submit", self._h, int(model_id), p.ctypes.data_as(ctypes.c_void_p),
                   p.shape[0], out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(rid))
        self._pending[rid.value] = (p, out)
        return rid.value

    def wait(self, request_id):
        """Blocks until the request completes; returns its output array."""
        try:
            _capi.call("pam_server_wait", self._h, int(request_id))
        finally:
            p, out = self._pending.pop(request_id, (None, None))
        return out

    def done(self, request_id):
        d = ctypes.c_int()
        _capi.call("pam_server_test", self._h, int(request_id), ctypes.byref(d))
        return bool(d.value)

    def drain(self):
        _capi.call("pam_server_drain", self._h)

    def request_times(self, request_id):
        q, t = ctypes.c_double(), ctypes.c_double()
        _capi.call("pam_server_request_times", self._h, int(request_id), ctypes.byref(q), ctypes.byref(t))
        return q.value, t.value

    def stats(self):
        s = _capi.PamServerStats()
        _capi.call("pam_server_get_stats", self._h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def reset_stats(self):
        _capi.call("pam_server_reset_stats", self._h)
