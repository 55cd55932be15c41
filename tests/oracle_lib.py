"""ctypes access to the checker: oracle/liboracle.so (C restatement) and,
when built, oracle/_ref/libstref.so (the reference's own CPU code)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libstref.so")

fp = ctypes.POINTER(ctypes.c_float)
dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)
lp = ctypes.POINTER(ctypes.c_long)
I, F, D, L = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_long

_O = None
_R = None


def P(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("oracle arguments must be C-contiguous")
    if a.dtype == np.float32:
        return a.ctypes.data_as(fp)
    if a.dtype == np.float64:
        return a.ctypes.data_as(dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(ip)
    if a.dtype == np.int64:
        return a.ctypes.data_as(lp)
    raise TypeError(a.dtype)


def oracle():
    global _O
    if _O is None:
        o = ctypes.CDLL(ORACLE)
        sig = {
            "or_pad_extent": (I, [I, I]),
            "or_pad_replicate": (None, [fp, I, I, I, fp, I, I, I]),
            "or_godunov_f32": (F, [F, F, F, F, F]),
            "or_godunov_f64": (D, [D, D, D, D, D]),
            "or_source_seed": (None, [fp, I, I, I, F, F, F, F, F, F, lp, fp]),
            "or_eikonal_fp64": (I, [fp, I, I, I, D, L, D, D, I, dp]),
            "or_block_fim_ref": (I, [fp, I, I, I, F, L, F, I, fp, lp]),
            "or_crop_f32": (None, [fp, I, I, I, fp, I, I, I]),
            "or_crop_f64": (None, [dp, I, I, I, dp, I, I, I]),
            "or_trilinear": (F, [F] * 17),
            "or_picks": (None, [fp, I, I, I, F, F, F, fp, fp, fp, I, fp]),
            "or_fsm": (None, [fp, I, I, I, F, F, F, F, F, F, fp]),
            "or_podvin": (None, [fp, I, I, I, F, F, F, F, F, F, fp]),
            "or_fsm_batch": (None, [fp, I, I, I, F, F, F, fp, I, fp, fp, fp, I, fp, I]),
            "or_expand_fdm": (None, [fp, I, I, I, fp]),
            "or_ray_trace_shot": (L, [fp, I, I, I, F, F, F, I, I, I, F, F, F, F, F, F,
                                      fp, fp, fp, I, I, ip, ip, fp, L, fp]),
            "or_tk_stencil": (None, [I, ip]),
            "or_cgls_f64": (None, [lp, ip, fp, I, I, fp, I, dp]),
            "or_cgls_ref_f32": (None, [lp, ip, fp, I, I, fp, I, fp]),
            "or_residual_norm": (F, [fp, fp, I]),
            "or_gaussian": (None, [fp, fp, I, I, I, I, F]),
            "or_gaussian_kernel": (None, [I, F, fp]),
            "or_model_update": (None, [fp, fp, I, I, I, F, F, F, I, I, I, F, F, F, I, I, F, fp, fp]),
            "or_acoustic_dampers": (None, [fp, fp, fp, F, I]),
            "or_acoustic_shot": (None, [fp, I, I, I, F, F, F, I, F, fp, I, F, F, F, F,
                                        fp, fp, fp, I, fp, fp]),
            "or_auto_pick": (None, [fp, I, I, F, F, F, F, fp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(o, name)
            f.restype = res
            f.argtypes = args
        _O = o
    return _O


def ref_available():
    return os.path.exists(REF)


def ref():
    global _R
    if _R is None:
        r = ctypes.CDLL(REF)
        r.ref_eikonal.argtypes = [I, fp, I, I, I, F, F, F, fp, I, fp, fp, fp, I, fp, fp,
                                  ctypes.c_char_p]
        r.ref_eikonal.restype = I
        r.ref_fsm_threads.argtypes = [fp, I, I, I, F, F, F, fp, I, fp, fp, fp, I, fp, I]
        r.ref_fsm_threads.restype = I
        r.ref_run_params.argtypes = [ctypes.c_char_p, I, ip, ip]
        r.ref_run_params.restype = I
        r.ref_trilinear.argtypes = [F] * 17
        r.ref_trilinear.restype = F
        r.ref_gaussian.argtypes = [fp, fp, I, I, I, I, F]
        r.ref_gaussian.restype = None
        r.ref_catch_parameter.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, I]
        r.ref_catch_parameter.restype = I
        r.ref_regular.argtypes = [ctypes.c_char_p, ctypes.c_char_p, fp, fp, fp, I]
        r.ref_regular.restype = I
        r.ref_acoustic_shot.argtypes = [fp, I, I, I, F, F, F, I, F, fp, I, F, F, F, F,
                                        fp, fp, fp, I, fp, fp]
        r.ref_acoustic_shot.restype = I
        r.ref_dampers.argtypes = [fp, fp, fp, F, I]
        r.ref_ricker.argtypes = [F, F, F, I, fp]
        r.ref_ricker.restype = None
        r.ref_dampers.restype = None
        _R = r
    return _R


# ---------------------------------------------------------------- helpers --

def padded_model(vel, B=4):
    """Slowness padded to the FIM domain (pad-to-4, clamp replication)."""
    o = oracle()
    s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
    ny, nx, nz = s.shape
    nxx, nyy, nzz = (o.or_pad_extent(n, B) for n in (nx, ny, nz))
    S = np.zeros((nyy, nxx, nzz), np.float32)
    o.or_pad_replicate(P(s), nx, ny, nz, P(S), nxx, nyy, nzz)
    return s, S


def eikonal_fp64(vel, h, src, tol=1e-12):
    """FP64 fixed point of the block-FIM discretisation, cropped (ny,nx,nz)."""
    o = oracle()
    s, S = padded_model(vel)
    ny, nx, nz = s.shape
    nyy, nxx, nzz = S.shape
    sid = ctypes.c_long()
    t0 = ctypes.c_float()
    o.or_source_seed(P(S), nxx, nyy, nzz, h, h, h, float(src[0]), float(src[1]), float(src[2]),
                     ctypes.byref(sid), ctypes.byref(t0))
    Tp = np.zeros(S.shape, np.float64)
    npass = o.or_eikonal_fp64(P(S), nxx, nyy, nzz, h, sid.value, t0.value, tol, 1000, P(Tp))
    assert npass > 0, "FP64 oracle did not converge"
    T = np.zeros(s.shape, np.float64)
    o.or_crop_f64(P(Tp), nxx, nyy, nzz, P(T), nx, ny, nz)
    return T


def block_fim_ref(vel, h, src, n_iter=100):
    o = oracle()
    s, S = padded_model(vel)
    ny, nx, nz = s.shape
    nyy, nxx, nzz = S.shape
    sid = ctypes.c_long()
    t0 = ctypes.c_float()
    o.or_source_seed(P(S), nxx, nyy, nzz, h, h, h, float(src[0]), float(src[1]), float(src[2]),
                     ctypes.byref(sid), ctypes.byref(t0))
    Tp = np.zeros(S.shape, np.float32)
    st = np.zeros(2, np.int64)
    o.or_block_fim_ref(P(S), nxx, nyy, nzz, h, sid.value, t0.value, n_iter, P(Tp), P(st))
    T = np.zeros(s.shape, np.float32)
    o.or_crop_f32(P(Tp), nxx, nyy, nzz, P(T), nx, ny, nz)
    return T, st


def picks(T, spacing, rcv):
    o = oracle()
    T = np.ascontiguousarray(T, dtype=np.float32)
    ny, nx, nz = T.shape
    rcv = np.ascontiguousarray(rcv, np.float32).reshape(-1, 3)
    rx, ry, rz = (np.ascontiguousarray(rcv[:, i]) for i in range(3))
    out = np.zeros(rcv.shape[0], np.float32)
    o.or_picks(P(T), nx, ny, nz, *spacing, P(rx), P(ry), P(rz), rcv.shape[0], P(out))
    return out


def fsm(vel, spacing, src):
    o = oracle()
    s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
    ny, nx, nz = s.shape
    T = np.zeros(s.shape, np.float32)
    o.or_fsm(P(s), nx, ny, nz, *spacing, *(float(v) for v in src), P(T))
    return T


def acoustic_case(nx=22, ny=18, nz=16, h=10.0, nb=8, nt=120, dt=0.001, seed=3, nrec=9,
                  spacing=None):
    """A small layered model with +-3 % noise, a 40 Hz Ricker wavelet
    delayed 20 ms, one source and nrec random receivers (inside the model).
    spacing = (dx, dy, dz) overrides h."""
    rng = np.random.default_rng(seed)
    dx, dy, dz = spacing if spacing is not None else (h, h, h)
    vel = np.empty((ny, nx, nz), np.float32)
    vel[:] = np.linspace(1500.0, 2600.0, nz, dtype=np.float32)[None, None, :]
    vel *= rng.uniform(0.97, 1.03, vel.shape).astype(np.float32)
    t = np.arange(nt, dtype=np.float32) * dt - 0.02
    f = 40.0
    w = ((1 - 2 * (np.pi * f * t) ** 2) * np.exp(-(np.pi * f * t) ** 2)).astype(np.float32)
    ext = np.array([(nx - 1) * dx, (ny - 1) * dy, (nz - 1) * dz], np.float32)
    src = (np.array([0.45, 0.55, 0.3], np.float32) * ext).astype(np.float32)
    rcv = (rng.uniform(0, 1, (nrec, 3)).astype(np.float32) * ext).astype(np.float32)
    return dict(vel=vel, nx=nx, ny=ny, nz=nz, d=(dx, dy, dz), nb=nb, nt=nt, dt=dt, wavelet=w,
                src=src, rcv=rcv, factor=0.0045 * 50 / nb)


def acoustic_shot(c, lib=None, src=None, want_ufut=False):
    """Seismogram [nrec, nt] (and the final padded U_fut) of one shot."""
    lib = lib or oracle()
    rcv = c["rcv"]
    src = c["src"] if src is None else src
    rx, ry, rz = (np.ascontiguousarray(rcv[:, i]) for i in range(3))
    seis = np.zeros((rcv.shape[0], c["nt"]), np.float32)
    nb = c["nb"]
    uf = np.zeros((c["ny"] + 2 * nb, c["nx"] + 2 * nb, c["nz"] + 2 * nb), np.float32) \
        if want_ufut else None
    fn = lib.or_acoustic_shot if hasattr(lib, "or_acoustic_shot") else lib.ref_acoustic_shot
    fn(P(np.ascontiguousarray(c["vel"])), c["nx"], c["ny"], c["nz"], *c["d"], nb, c["factor"],
       P(c["wavelet"]), c["nt"], c["dt"], *(float(v) for v in src), P(rx), P(ry), P(rz),
       rcv.shape[0], P(seis), P(uf) if uf is not None else None)
    return (seis, uf) if want_ufut else seis
