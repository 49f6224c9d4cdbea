"""CPU oracle for the gHCM hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
library (``ghcm_amd``) never loads it.

Thin ctypes bindings over ``oracle/liboracle.so`` (built from ``oracle.c`` by
``oracle/Makefile``).  Each C function cites the reference lines it restates;
see ``oracle.h`` for how each one is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    """Compile oracle.c into liboracle.so (gcc, CPU only)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f32, f64 = ctypes.c_float, ctypes.c_double
        vp, u32 = ctypes.c_void_p, ctypes.c_uint32
        L.orc_godunov_cpu.restype = f32
        L.orc_godunov_cpu.argtypes = [f32, f32, f32, f32]
        L.orc_godunov_gpu.restype = f32
        L.orc_godunov_gpu.argtypes = [f32, f32, f32, f32]
        L.orc_godunov_safe.restype = f32
        L.orc_godunov_safe.argtypes = [f32, f32, f32, f32]
        L.orc_godunov_f64.restype = f64
        L.orc_godunov_f64.argtypes = [f64, f64, f64, f64]
        L.orc_fmm_f32.restype = ctypes.c_int
        L.orc_fmm_f32.argtypes = [vp, u32, u32, vp, u32, vp]
        L.orc_fmm_f64.restype = ctypes.c_int
        L.orc_fmm_f64.argtypes = [vp, u32, u32, vp, u32, vp]
        L.orc_fim_f32.restype = ctypes.c_int64
        L.orc_fim_f32.argtypes = [f32, vp, u32, u32, vp, u32, vp]
        L.orc_sweep_f64.restype = ctypes.c_int
        L.orc_sweep_f64.argtypes = [vp, u32, u32, vp, u32, vp, ctypes.c_int]
        L.orc_godunov_k_f32.restype = f32
        L.orc_godunov_k_f32.argtypes = [vp, ctypes.c_int, f32, f32]
        L.orc_godunov_k_f64.restype = f64
        L.orc_godunov_k_f64.argtypes = [vp, ctypes.c_int, f64, f64]
        L.orc_fmm3d_f32.restype = ctypes.c_int
        L.orc_fmm3d_f32.argtypes = [vp, u32, u32, u32, vp, u32, vp]
        L.orc_fmm3d_f64.restype = ctypes.c_int
        L.orc_fmm3d_f64.argtypes = [vp, u32, u32, u32, vp, u32, vp]
        L.orc_sweep3d_f64.restype = ctypes.c_int
        L.orc_sweep3d_f64.argtypes = [vp, u32, u32, u32, vp, u32, vp, ctypes.c_int]
        L.orc_error_metrics.restype = None
        L.orc_error_metrics.argtypes = [vp, vp, ctypes.c_size_t, vp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _seeds(seeds) -> np.ndarray:
    s = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint32).reshape(-1, 2))
    return s


def godunov_cpu(ax, ay, cost, t_curr=np.inf) -> float:
    return lib().orc_godunov_cpu(ax, ay, cost, t_curr)


def godunov_gpu(ax, ay, cost, t_curr=np.inf) -> float:
    return lib().orc_godunov_gpu(ax, ay, cost, t_curr)


def godunov_safe(ax, ay, cost, t_curr=np.inf) -> float:
    return lib().orc_godunov_safe(ax, ay, cost, t_curr)


def godunov_f64(ax, ay, cost, t_curr=np.inf) -> float:
    return lib().orc_godunov_f64(ax, ay, cost, t_curr)


def fmm_f32(cost: np.ndarray, seeds) -> np.ndarray:
    """Reference-restated sequential FMM in FP32; cost is (H, W)."""
    c = np.ascontiguousarray(cost, dtype=np.float32)
    h, w = c.shape
    s = _seeds(seeds)
    out = np.empty_like(c)
    rc = lib().orc_fmm_f32(_ptr(c), w, h, _ptr(s), s.shape[0], _ptr(out))
    if rc != 0:
        raise ValueError("orc_fmm_f32 failed")
    return out


def fmm_f64(cost: np.ndarray, seeds) -> np.ndarray:
    """FMM in FP64 -- the accuracy oracle."""
    c = np.ascontiguousarray(cost, dtype=np.float64)
    h, w = c.shape
    s = _seeds(seeds)
    out = np.empty_like(c)
    rc = lib().orc_fmm_f64(_ptr(c), w, h, _ptr(s), s.shape[0], _ptr(out))
    if rc != 0:
        raise ValueError("orc_fmm_f64 failed")
    return out


def fim_f32(cost: np.ndarray, seeds, minimum_change: float = 1e-5):
    """Reference-restated sequential FIM; returns (times, number of solves)."""
    c = np.ascontiguousarray(cost, dtype=np.float32)
    h, w = c.shape
    s = _seeds(seeds)
    out = np.empty_like(c)
    n = lib().orc_fim_f32(minimum_change, _ptr(c), w, h, _ptr(s), s.shape[0], _ptr(out))
    if n < 0:
        raise ValueError("orc_fim_f32 failed")
    return out, int(n)


def sweep_f64(cost: np.ndarray, seeds, max_rounds: int = 10000):
    """Brute-force Gauss-Seidel fixed point in FP64; returns (times, rounds)."""
    c = np.ascontiguousarray(cost, dtype=np.float64)
    h, w = c.shape
    s = _seeds(seeds)
    out = np.empty_like(c)
    r = lib().orc_sweep_f64(_ptr(c), w, h, _ptr(s), s.shape[0], _ptr(out), max_rounds)
    if r < 0:
        raise ValueError("orc_sweep_f64 failed")
    return out, int(r)


def godunov_k_f32(dims, cost, best_prev=np.inf) -> float:
    """eikonal_solve<K> in FP32 (K = len(dims), 1..3), larger root."""
    d = np.ascontiguousarray(dims, dtype=np.float32)
    return lib().orc_godunov_k_f32(_ptr(d), d.size, cost, best_prev)


def godunov_k_f64(dims, cost, best_prev=np.inf) -> float:
    """eikonal_solve<K> in FP64 (K = len(dims), 1..3), larger root."""
    d = np.ascontiguousarray(dims, dtype=np.float64)
    return lib().orc_godunov_k_f64(_ptr(d), d.size, cost, best_prev)


def _seeds3(seeds) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(seeds, dtype=np.uint32).reshape(-1, 3))


def fmm3d_f32(cost: np.ndarray, seeds) -> np.ndarray:
    """Three-dimensional FMM in FP32; cost is (D, H, W), seeds (x, y, z)."""
    c = np.ascontiguousarray(cost, dtype=np.float32)
    d, h, w = c.shape
    s = _seeds3(seeds)
    out = np.empty_like(c)
    if lib().orc_fmm3d_f32(_ptr(c), w, h, d, _ptr(s), s.shape[0], _ptr(out)) != 0:
        raise ValueError("orc_fmm3d_f32 failed")
    return out


def fmm3d_f64(cost: np.ndarray, seeds) -> np.ndarray:
    """Three-dimensional FMM in FP64 -- the accuracy oracle in 3-D."""
    c = np.ascontiguousarray(cost, dtype=np.float64)
    d, h, w = c.shape
    s = _seeds3(seeds)
    out = np.empty_like(c)
    if lib().orc_fmm3d_f64(_ptr(c), w, h, d, _ptr(s), s.shape[0], _ptr(out)) != 0:
        raise ValueError("orc_fmm3d_f64 failed")
    return out


def sweep3d_f64(cost: np.ndarray, seeds, max_rounds: int = 10000):
    """Three-dimensional Gauss-Seidel fixed point in FP64; returns (times, rounds)."""
    c = np.ascontiguousarray(cost, dtype=np.float64)
    d, h, w = c.shape
    s = _seeds3(seeds)
    out = np.empty_like(c)
    r = lib().orc_sweep3d_f64(_ptr(c), w, h, d, _ptr(s), s.shape[0], _ptr(out), max_rounds)
    if r < 0:
        raise ValueError("orc_sweep3d_f64 failed")
    return out, int(r)


def error_metrics(got: np.ndarray, ref: np.ndarray) -> dict:
    """L_inf / L1 of an FP32 result against an FP64 reference (finite cells)."""
    g = np.ascontiguousarray(got, dtype=np.float32).ravel()
    r = np.ascontiguousarray(ref, dtype=np.float64).ravel()
    if g.size != r.size:
        raise ValueError("size mismatch")
    out = np.zeros(5, dtype=np.float64)
    lib().orc_error_metrics(_ptr(g), _ptr(r), g.size, _ptr(out))
    return {"linf": out[0], "l1": out[1], "max_ref": out[2],
            "finiteness_mismatch": int(out[3]), "compared": int(out[4])}
