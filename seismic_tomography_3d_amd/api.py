"""Pythonic wrapper over the C ABI.  Every call goes to libst_hip.so."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import FimOpts, FimStats


class StError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_lib.ST_STATUS.get(code, code)}: {msg}")
        self.code = code


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return a.ctypes.data_as(_lib.c_float_p) if a is not None else None


class Context:
    """One device context: model, FIM state and tomography buffers."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        self._check(self.lib.st_ctx_create(device, ctypes.byref(h)), None)
        self.h = h
        self.shape = None

    def _check(self, rc, h="self"):
        if rc != 0:
            handle = self.h if h == "self" else None
            msg = self.lib.st_last_error(handle).decode() if handle else ""
            raise StError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.st_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- model ----
    def set_model(self, slowness, dx, dy=None, dz=None):
        """slowness: (ny, nx, nz) float32 (z fastest in memory)."""
        s = _f32(slowness)
        ny, nx, nz = s.shape
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        self._check(self.lib.st_set_model(self.h, _ptr(s), nx, ny, nz, dx, dy, dz))
        self.shape = (ny, nx, nz)
        self.spacing = (dx, dy, dz)

    def get_model(self):
        out = np.empty(self.shape, np.float32)
        self._check(self.lib.st_get_model(self.h, _ptr(out)))
        return out

    # ---- eikonal ----
    @staticmethod
    def options(**kw):
        o = FimOpts()
        _lib.load().st_fim_default_opts(ctypes.byref(o))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        return o

    def fim_solve(self, src, rcv, return_T=False, **opts):
        """Solve each source; returns (picks[nshot, nrcv], T or None, stats)."""
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3)) if rcv is not None else np.zeros((0, 3), np.float32)
        ns, nr = src.shape[0], rcv.shape[0]
        picks = np.empty((ns, nr), np.float32)
        T = np.empty((ns,) + self.shape, np.float32) if return_T else None
        o = self.options(**opts)
        st = FimStats()
        self._check(self.lib.st_fim_solve(self.h, _ptr(src), ns, _ptr(rcv) if nr else None, nr,
                                          _ptr(picks) if nr else None, _ptr(T),
                                          ctypes.byref(o), ctypes.byref(st)))
        return picks, T, st.as_dict()

    def picks_volume(self, T, spacing, rcv):
        T = _f32(T)
        ny, nx, nz = T.shape
        rcv = _f32(rcv, (-1, 3))
        out = np.empty(rcv.shape[0], np.float32)
        self._check(self.lib.st_picks_volume(self.h, _ptr(T), nx, ny, nz, *spacing, _ptr(rcv),
                                             rcv.shape[0], _ptr(out)))
        return out
