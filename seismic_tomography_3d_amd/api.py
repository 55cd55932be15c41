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
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array passed across the C ABI must be C-contiguous")
    return a.ctypes.data_as(_lib.c_float_p)


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

    def fim_solve_noalloc(self, src, rcv, **opts):
        """Cold-start solve whose device buffers all come from the library's
        static scratch array (no device allocation); returns (picks, stats)."""
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3)) if rcv is not None else np.zeros((0, 3), np.float32)
        ns, nr = src.shape[0], rcv.shape[0]
        picks = np.empty((ns, nr), np.float32)
        o = self.options(**opts)
        st = FimStats()
        self._check(self.lib.st_fim_solve_noalloc(self.h, _ptr(src), ns, _ptr(rcv) if nr else None,
                                                  nr, _ptr(picks) if nr else None,
                                                  ctypes.byref(o), ctypes.byref(st)))
        return picks, st.as_dict()

    def fim_solve_warm(self, src, rcv, T_pred, return_T=False, **opts):
        """Warm-started solve from caller predictions T_pred[nshot, ny, nx, nz]."""
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3)) if rcv is not None else np.zeros((0, 3), np.float32)
        ns, nr = src.shape[0], rcv.shape[0]
        Tp = _f32(T_pred, (ns,) + self.shape)
        picks = np.empty((ns, nr), np.float32)
        T = np.empty((ns,) + self.shape, np.float32) if return_T else None
        o = self.options(**opts)
        st = FimStats()
        self._check(self.lib.st_fim_solve_warm(self.h, _ptr(src), ns, _ptr(rcv) if nr else None,
                                               nr, _ptr(picks) if nr else None, _ptr(T), _ptr(Tp),
                                               ctypes.byref(o), ctypes.byref(st)))
        return picks, T, st.as_dict()

    def fsm_solve(self, src, rcv, return_T=False, batch=0):
        """The reference's Accurate_FSM on the device, bit-identical to its CPU
        code.  Returns (picks[nshot, nrcv], T or None, device ms)."""
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3)) if rcv is not None else np.zeros((0, 3), np.float32)
        ns, nr = src.shape[0], rcv.shape[0]
        picks = np.empty((ns, nr), np.float32)
        T = np.empty((ns,) + self.shape, np.float32) if return_T else None
        ms = ctypes.c_double()
        self._check(self.lib.st_fsm_solve(self.h, _ptr(src), ns, _ptr(rcv) if nr else None, nr,
                                          _ptr(picks) if nr else None, _ptr(T), int(batch),
                                          ctypes.byref(ms)))
        return picks, T, ms.value

    def podvin_solve(self, src, rcv, return_T=False, batch=0):
        """The reference's Podvin & Lecomte eikonal on the device, bit-identical
        to its CPU code.  Returns (picks[nshot, nrcv], T or None, device ms)."""
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3)) if rcv is not None else np.zeros((0, 3), np.float32)
        ns, nr = src.shape[0], rcv.shape[0]
        picks = np.empty((ns, nr), np.float32)
        T = np.empty((ns,) + self.shape, np.float32) if return_T else None
        ms = ctypes.c_double()
        self._check(self.lib.st_podvin_solve(self.h, _ptr(src), ns, _ptr(rcv) if nr else None, nr,
                                             _ptr(picks) if nr else None, _ptr(T), int(batch),
                                             ctypes.byref(ms)))
        return picks, T, ms.value

    def picks_volume(self, T, spacing, rcv):
        T = _f32(T)
        ny, nx, nz = T.shape
        rcv = _f32(rcv, (-1, 3))
        out = np.empty(rcv.shape[0], np.float32)
        self._check(self.lib.st_picks_volume(self.h, _ptr(T), nx, ny, nz, *spacing, _ptr(rcv),
                                             rcv.shape[0], _ptr(out)))
        return out

    # ---- data synthesis (src/utils/wave, src/utils/auto_picking) ----
    def acoustic_set_model(self, vel, spacing, nb=50, factor=0.0045, dt=0.001):
        """Velocity model vel[ny, nx, nz] (m/s) for the acoustic modelling;
        spacing = h or (dx, dy, dz); nb absorbing cells per side."""
        vel = _f32(vel)
        ny, nx, nz = vel.shape
        dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
        self._check(self.lib.st_acoustic_set_model(self.h, _ptr(vel), nx, ny, nz, dx, dy, dz,
                                                   int(nb), factor, dt))
        self._acoustic = (nx, ny, nz, int(nb))

    def acoustic_shots(self, wavelet, src, rcv, group=0, return_ufut=False, return_seis=True):
        """Seismograms [nshot, nrec, nt] of the acoustic modelling (and the
        last shot's final padded U_fut [nyy, nxx, nzz]); returns (seis or
        None, ufut or None, device ms)."""
        w = _f32(wavelet, (-1,))
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3))
        ns, nr, nt = src.shape[0], rcv.shape[0], w.shape[0]
        seis = np.empty((ns, nr, nt), np.float32) if return_seis else None
        uf = None
        if return_ufut:
            nx, ny, nz, nb = self._acoustic
            uf = np.empty((ny + 2 * nb, nx + 2 * nb, nz + 2 * nb), np.float32)
        ms = ctypes.c_double()
        self._check(self.lib.st_acoustic_shots(self.h, _ptr(w), nt, _ptr(src), ns, _ptr(rcv), nr,
                                               int(group), _ptr(seis), _ptr(uf),
                                               ctypes.byref(ms)))
        return seis, uf, ms.value

    def acoustic_picks(self, wavelet, src, rcv, window=0.01, amp_cut=1e-15, shift=0.045,
                       group=0):
        """Model and auto-pick on the device: (picks[nshot, nrec], device ms)."""
        w = _f32(wavelet, (-1,))
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3))
        picks = np.empty((src.shape[0], rcv.shape[0]), np.float32)
        ms = ctypes.c_double()
        self._check(self.lib.st_acoustic_picks(self.h, _ptr(w), w.shape[0], _ptr(src),
                                               src.shape[0], _ptr(rcv), rcv.shape[0], int(group),
                                               window, amp_cut, shift, _ptr(picks),
                                               ctypes.byref(ms)))
        return picks, ms.value

    def auto_pick(self, seis, dt, window=0.01, amp_cut=1e-15, shift=0.045):
        """The auto-picker over one file seis[ntraces, nt]."""
        seis = _f32(seis)
        ntr, nt = seis.shape
        picks = np.empty(ntr, np.float32)
        self._check(self.lib.st_auto_pick(self.h, _ptr(seis), ntr, nt, dt, window, amp_cut,
                                          shift, _ptr(picks)))
        return picks

    # ---- tomography ----
    def tomo_set_grid(self, dxt, dyt=None, dzt=None, max_steps=0):
        dyt = dxt if dyt is None else dyt
        dzt = dxt if dzt is None else dzt
        self._check(self.lib.st_tomo_set_grid(self.h, dxt, dyt, dzt, max_steps))
        return self.tomo_dims()

    def tomo_dims(self):
        n = [ctypes.c_int() for _ in range(3)]
        self._check(self.lib.st_tomo_dims(self.h, *[ctypes.byref(v) for v in n]))
        return tuple(v.value for v in n)  # (nx_tomo, ny_tomo, nz_tomo)

    def tomo_set_eikonal(self, eikonal_type):
        """The forward pass's eikonal: 0 Podvin, 1 block FIM, 2 Accurate_FSM
        (other values: Accurate_FSM, as the reference's default branch)."""
        self._check(self.lib.st_tomo_set_eikonal(self.h, int(eikonal_type)))

    def forward(self, src, rcv, illum=False, return_T=False, **opts):
        """Eikonal + picks + ray tracing into the device G.  Returns (dcal,
        illum, stats), and the travel times T[nshot, ny, nx, nz] the rays were
        traced on as a fourth item when return_T."""
        src = _f32(src, (-1, 3))
        rcv = _f32(rcv, (-1, 3))
        dcal = np.empty((src.shape[0], rcv.shape[0]), np.float32)
        il = np.empty(self.shape, np.float32) if illum else None
        o = self.options(**opts)
        st = FimStats()
        if return_T:
            T = np.empty((src.shape[0],) + self.shape, np.float32)
            self._check(self.lib.st_forward_times(self.h, _ptr(src), src.shape[0], _ptr(rcv),
                                                  rcv.shape[0], _ptr(dcal), _ptr(il), _ptr(T),
                                                  ctypes.byref(o), ctypes.byref(st)))
            return dcal, il, st.as_dict(), T
        self._check(self.lib.st_forward(self.h, _ptr(src), src.shape[0], _ptr(rcv), rcv.shape[0],
                                        _ptr(dcal), _ptr(il), ctypes.byref(o), ctypes.byref(st)))
        return dcal, il, st.as_dict()

    def csr_reset(self):
        self._check(self.lib.st_csr_reset(self.h))

    def raytrace_volume(self, T, src, rcv, illum=False):
        T = _f32(T)
        src = _f32(src, (3,))
        rcv = _f32(rcv, (-1, 3))
        il = np.empty(self.shape, np.float32) if illum else None
        self._check(self.lib.st_raytrace_volume(self.h, _ptr(T), _ptr(src), _ptr(rcv),
                                                rcv.shape[0], _ptr(il)))
        return il

    def csr_size(self):
        """(rows, non-zeros) of the device sensitivity matrix."""
        rows, nnz = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.lib.st_csr_size(self.h, ctypes.byref(rows), ctypes.byref(nnz)))
        return rows.value, nnz.value

    def csr(self):
        rows, nnz = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.lib.st_csr_size(self.h, ctypes.byref(rows), ctypes.byref(nnz)))
        rp = np.empty(rows.value + 1, np.int64)
        c = np.empty(max(1, nnz.value), np.int32)
        v = np.empty(max(1, nnz.value), np.float32)
        self._check(self.lib.st_csr_copy(self.h, rp.ctypes.data_as(_lib.c_ll_p),
                                         c.ctypes.data_as(_lib.c_int_p), _ptr(v)))
        return rp, c[:nnz.value], v[:nnz.value]

    def residual_norm(self, dobs, dcal):
        dobs = _f32(dobs).ravel()
        dcal = _f32(dcal).ravel()
        out = ctypes.c_double()
        self._check(self.lib.st_residual_norm(self.h, _ptr(dobs), _ptr(dcal), dobs.size,
                                              ctypes.byref(out)))
        return out.value

    def optimize(self, dobs, dcal, tk_order=2, lam=1e-2, iters=10):
        dobs = _f32(dobs).ravel()
        dcal = _f32(dcal).ravel()
        nt = self.tomo_dims()
        x = np.empty(nt[0] * nt[1] * nt[2], np.float32)
        self._check(self.lib.st_optimize(self.h, _ptr(dobs), _ptr(dcal), dobs.size, tk_order,
                                         lam, iters, _ptr(x)))
        return x

    def cgls_csr(self, rowptr, cols, vals, M, b, iters):
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        cols = np.ascontiguousarray(cols, np.int32)
        vals = _f32(vals)
        b = _f32(b)
        x = np.empty(M, np.float32)
        self._check(self.lib.st_cgls_csr(self.h, rowptr.ctypes.data_as(_lib.c_ll_p),
                                         cols.ctypes.data_as(_lib.c_int_p), _ptr(vals),
                                         len(rowptr) - 1, M, _ptr(b), iters, _ptr(x)))
        return x

    def optimize_update(self, dobs, dcal, tk_order, lam, iters, smooth=True, samples=5,
                        stdv=2.0, graph=True, fetch=True):
        """CGLS and the model update in one step (one captured HIP graph when
        graph=True).  Returns (x, dm, velocity), or (None, None, None) when
        fetch=False (the updated model stays on the device)."""
        dobs = _f32(dobs, (-1,))
        dcal = _f32(dcal, (-1,))
        nxt, nyt, nzt = self.tomo_dims()
        x = dm = vel = None
        if fetch:
            x = np.empty(nxt * nyt * nzt, np.float32)
            dm = np.empty(self.shape, np.float32)
            vel = np.empty(self.shape, np.float32)
        self._check(self.lib.st_optimize_update(self.h, _ptr(dobs), _ptr(dcal), dobs.shape[0],
                                                int(tk_order), lam, int(iters), int(bool(smooth)),
                                                int(samples), stdv, int(bool(graph)), _ptr(x),
                                                _ptr(dm), _ptr(vel)))
        return x, dm, vel

    def model_update(self, xdm=None, smooth=False, samples=5, stdv=2.0):
        dm = np.empty(self.shape, np.float32)
        vel = np.empty(self.shape, np.float32)
        x = _f32(xdm) if xdm is not None else None
        self._check(self.lib.st_model_update(self.h, _ptr(x), int(bool(smooth)), samples, stdv,
                                             _ptr(dm), _ptr(vel)))
        return dm, vel

    # ---- learned warm start ----
    def surrogate_init(self, nc=32, seed=0):
        self._check(self.lib.st_surrogate_init(self.h, nc, seed))
        n = ctypes.c_longlong()
        self._check(self.lib.st_surrogate_nparams(self.h, ctypes.byref(n)))
        self.nc = nc
        return n.value

    def surrogate_params(self, set_to=None):
        n = ctypes.c_longlong()
        self._check(self.lib.st_surrogate_nparams(self.h, ctypes.byref(n)))
        out = np.empty(n.value, np.float32)
        s = _f32(set_to) if set_to is not None else None
        self._check(self.lib.st_surrogate_params(self.h, _ptr(out), _ptr(s)))
        return out

    def surrogate_eval(self, X, target=None, mask=None):
        """X[ns, nc^3, 6] -> (Y[ns, nc^3], loss, grad)."""
        X = _f32(X)
        ns = X.shape[0]
        V = X.size // (ns * 6)
        Y = np.empty((ns, V), np.float32)
        loss = ctypes.c_double()
        grad = None
        t = m = None
        if target is not None:
            t, m = _f32(target), _f32(mask)
            n = ctypes.c_longlong()
            self._check(self.lib.st_surrogate_nparams(self.h, ctypes.byref(n)))
            grad = np.empty(n.value, np.float32)
        self._check(self.lib.st_surrogate_eval(self.h, _ptr(X), ns, _ptr(t), _ptr(m), _ptr(Y),
                                               ctypes.byref(loss), _ptr(grad)))
        return Y, (loss.value if target is not None else None), grad

    def surrogate_make_data(self, n_models, shots_per_model=4, zmax_src=2000.0, seed=5,
                            grad_fraction=0.5):
        self._check(self.lib.st_surrogate_make_data(self.h, n_models, shots_per_model, zmax_src,
                                                    seed, grad_fraction))

    def surrogate_loss(self, under_weight):
        """Weight of under-predictions in the training loss (1 = plain MSE)."""
        self._check(self.lib.st_surrogate_loss(self.h, float(under_weight)))

    def surrogate_train(self, epochs, batch=8, lr=1e-3, n_val=64, seed=0):
        hist = np.zeros((epochs, 2), np.float32)
        self._check(self.lib.st_surrogate_train(self.h, epochs, batch, lr, n_val, seed, _ptr(hist)))
        return hist

    def surrogate_save(self, path):
        self._check(self.lib.st_surrogate_save(self.h, str(path).encode()))

    def surrogate_load(self, path):
        self._check(self.lib.st_surrogate_load(self.h, str(path).encode()))

    def surrogate_predict(self, src):
        src = _f32(src, (-1, 3))
        nc = getattr(self, "nc", 32)
        y = np.empty((src.shape[0], nc ** 3), np.float32)
        self._check(self.lib.st_surrogate_predict(self.h, _ptr(src), src.shape[0], _ptr(y)))
        return y

    def nn_gemm(self, A, B):
        A = _f32(A)
        B = _f32(B)
        M, K = A.shape
        N = B.shape[1]
        C = np.empty((M, N), np.float32)
        self._check(self.lib.st_nn_gemm(self.h, _ptr(A), _ptr(B), _ptr(C), M, N, K))
        return C


def acoustic_dampers(factor, nb):
    """Absorbing-layer profiles (d1[nb], d2[nb, nb], d3[nb, nb, nb]) made on
    the host (no device needed)."""
    lib = _lib.load()
    d1 = np.zeros(nb, np.float32)
    d2 = np.zeros((nb, nb), np.float32)
    d3 = np.zeros((nb, nb, nb), np.float32)
    rc = lib.st_acoustic_dampers(_ptr(d1), _ptr(d2), _ptr(d3), factor, int(nb))
    if rc != 0:
        raise ValueError(f"st_acoustic_dampers: status {rc}")
    return d1, d2, d3


def ricker(nt, dt, fmax, delay):
    """The reference's Ricker wavelet (made on the host)."""
    out = np.zeros(int(nt), np.float32)
    rc = _lib.load().st_ricker(delay, fmax, dt, int(nt), _ptr(out))
    if rc != 0:
        raise ValueError(f"st_ricker: status {rc}")
    return out
