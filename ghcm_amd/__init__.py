"""ghcm_amd -- MI355X-native gHCM eikonal solver (Python host bindings).

The solver itself lives in ``libghcm.so`` (gfx950 HIP kernels + C-ABI + C++ API,
built from ``ghcm_amd/csrc``).  This module is a thin ctypes layer over the C-ABI of
``include/ghcm.h`` and ``include/eikonal_c.h``; it mirrors the reference's solver
surface (src/eikonal.h:12-67):

    HcmConfig fields          -> Config(...) / Context(**config)
    eikonal_hcm_opencl_setup  -> Context(ordinal, **config)
    eikonal_hcm_opencl_ex     -> Context.solve(cost, seeds)
    eikonal_fmm_sequential    -> fmm_sequential(cost, seeds)
    eikonal_fim_sequential    -> fim_sequential(minimum_change, cost, seeds)

There is no CPU fallback for the GPU path: if ``libghcm.so`` is missing, or no HIP
device is present, the GPU entry points raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libghcm.so")

# ghcm_status
OK, ERR_INVALID_ARG, ERR_NO_DEVICE, ERR_HIP, ERR_OUT_OF_MEMORY, ERR_NOT_CONVERGED, \
    ERR_UNSUPPORTED, ERR_OVER_BUDGET, ERR_NOT_FOUND, ERR_STOPPED, ERR_TIMEOUT = range(11)
# ghcm_kernel / ghcm_scheduler / ghcm_field_kind
KERNEL_SYSTOLIC, KERNEL_FIM, KERNEL_SWEEP = 0, 1, 2
SCHED_DEVICE_TOPN, SCHED_HOST_PQ = 0, 1
FIELD_COST_F32, FIELD_SPEED_F32, FIELD_SPEED_F16, FIELD_SPEED_BF16 = 0, 1, 2, 3

KINDS = {"cost": FIELD_COST_F32, "speed_f32": FIELD_SPEED_F32,
         "speed_f16": FIELD_SPEED_F16, "speed_bf16": FIELD_SPEED_BF16}


class Config(ctypes.Structure):
    """ghcm_config (include/ghcm.h); mirrors HcmConfig, src/eikonal.h:12-27."""
    _fields_ = [("tile_w", ctypes.c_uint32), ("tile_h", ctypes.c_uint32),
                ("workers_overload", ctypes.c_float),
                ("compute_unit_overload", ctypes.c_float),
                ("n_active", ctypes.c_uint32), ("min_change", ctypes.c_float),
                ("cache_mode", ctypes.c_int32), ("kernel", ctypes.c_int32),
                ("scheduler", ctypes.c_int32), ("inner_iter_cap", ctypes.c_uint32),
                ("max_waves", ctypes.c_uint32), ("waves_per_poll", ctypes.c_uint32),
                ("profile_kernels", ctypes.c_uint32), ("launch_mode", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("duration_total_us", ctypes.c_double),
                ("duration_device_us", ctypes.c_double),
                ("waves", ctypes.c_uint64), ("tile_visits", ctypes.c_uint64),
                ("tiles_total", ctypes.c_uint64), ("cells", ctypes.c_uint64),
                ("inner_iterations", ctypes.c_uint64), ("cells_improved", ctypes.c_uint64),
                ("converged", ctypes.c_uint32), ("tile_w", ctypes.c_uint32),
                ("tile_h", ctypes.c_uint32), ("n_active", ctypes.c_uint32),
                ("bytes_per_visit", ctypes.c_uint32), ("tile_d", ctypes.c_uint32),
                ("solve_kernel_us", ctypes.c_double), ("solve_launches", ctypes.c_uint64)]

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64),
                ("ordinal", ctypes.c_int32), ("compute_units", ctypes.c_int32),
                ("clock_khz", ctypes.c_int32), ("max_workgroup", ctypes.c_int32),
                ("lds_bytes", ctypes.c_uint64), ("global_mem_bytes", ctypes.c_uint64),
                ("pci_bus_id", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class TuneCandidate(ctypes.Structure):
    _fields_ = [("tile_w", ctypes.c_uint32), ("tile_h", ctypes.c_uint32),
                ("kernel", ctypes.c_int32), ("n_active", ctypes.c_uint32),
                ("inner_iter_cap", ctypes.c_uint32), ("launch_mode", ctypes.c_uint32),
                ("device_us", ctypes.c_double), ("tile_visits", ctypes.c_uint64),
                ("converged", ctypes.c_uint32), ("status", ctypes.c_int32)]

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Instance(ctypes.Structure):
    _fields_ = [("d_field", ctypes.c_void_p), ("d_time", ctypes.c_void_p),
                ("w", ctypes.c_uint32), ("h", ctypes.c_uint32),
                ("seeds_xy", ctypes.c_void_p), ("n_seeds", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_tiles", ctypes.c_uint32),
                ("max_slots", ctypes.c_uint32), ("field_kind", ctypes.c_int32),
                ("reserved", ctypes.c_uint32 * 3)]


class ServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("w", ctypes.c_uint32), ("h", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("waves", ctypes.c_uint64),
                ("queue_us", ctypes.c_double), ("latency_us", ctypes.c_double)]

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServeStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "submitted", "admitted", "completed", "pending", "waves_enqueued", "busy_waves",
        "loop_wave", "peak_live", "budget_bytes", "bytes_in_use", "peak_bytes", "max_tiles")] + \
        [(k, ctypes.c_uint32) for k in ("slots", "n_active", "waves_per_chunk", "waiters")]

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TileDesc(ctypes.Structure):
    _fields_ = [("tiles_x", ctypes.c_uint32), ("tiles_y", ctypes.c_uint32),
                ("tx", ctypes.c_uint32), ("ty", ctypes.c_uint32),
                ("x0", ctypes.c_uint32), ("y0", ctypes.c_uint32),
                ("w", ctypes.c_uint32), ("h", ctypes.c_uint32),
                ("neighbour", ctypes.c_uint32 * 4), ("border_bit", ctypes.c_uint32 * 4),
                ("seed_tile", ctypes.c_uint32)]

    def to_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["neighbour"] = list(self.neighbour)
        d["border_bit"] = list(self.border_bit)
        return d


class Image(ctypes.Structure):
    _fields_ = [("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("comp", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("data", ctypes.POINTER(ctypes.c_float))]


# Every symbol declared by include/ghcm.h, include/eikonal_c.h and include/ghcm_io.h.
EXPORTED_SYMBOLS = (
    "ghcm_abi_version", "ghcm_status_string", "ghcm_last_error", "ghcm_config_default",
    "ghcm_device_count", "ghcm_device_info_get", "ghcm_ctx_create", "ghcm_ctx_destroy",
    "ghcm_ctx_get_config", "ghcm_ctx_set_config", "ghcm_solve", "ghcm_solve_device",
    "ghcm_solve3d", "ghcm_solve3d_device",
    "ghcm_dev_alloc", "ghcm_dev_free", "ghcm_dev_upload", "ghcm_dev_download", "ghcm_dev_copy",
    "ghcm_solve_batch", "ghcm_tune_default_space", "ghcm_autotune",
    "ghcm_serve_config_default", "ghcm_serve_start", "ghcm_serve_submit",
    "ghcm_serve_submit_device", "ghcm_serve_wait",
    "ghcm_serve_get_stats", "ghcm_serve_stop",
    "ghcm_gen_ring_speed", "ghcm_gen_smooth_speed", "ghcm_gen_points",
    "ghcm_cpu_fmm", "ghcm_cpu_fim", "ghcm_hcm_solve_cxx",
    "ghcm_image_free", "ghcm_image_load_f32", "ghcm_png_decode8", "ghcm_png_write8",
    "ghcm_field_load_mono", "ghcm_field_load_rgb_luminosity", "ghcm_rgb_luminosity_device",
    "ghcm_write_time_field_png", "ghcm_write_field_hdr", "ghcm_blob_write", "ghcm_blob_read",
    "ghcm_test_run", "ghcm_field_cost_sin_checkerboard", "ghcm_field_cost_random",
    "ghcm_tile_info", "ghcm_debug_select", "ghcm_solve_device_noalloc",
)

_lib = None


class PreparedBatch:
    """A batch's ghcm_instance array and the seed arrays it points into."""

    def __init__(self, arr, keep, n: int):
        self.arr, self.keep, self.n = arr, keep, n


class GhcmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"ghcm status {status}: {msg}")
        self.status = status


def lib() -> ctypes.CDLL:
    """Load libghcm.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    P = ctypes.POINTER
    L.ghcm_abi_version.restype = i32
    L.ghcm_status_string.restype = ctypes.c_char_p
    L.ghcm_status_string.argtypes = [i32]
    L.ghcm_last_error.restype = ctypes.c_char_p
    L.ghcm_config_default.argtypes = [P(Config)]
    L.ghcm_config_default.restype = None
    L.ghcm_device_count.argtypes = [P(ctypes.c_int)]
    L.ghcm_device_info_get.argtypes = [i32, P(DeviceInfo)]
    L.ghcm_ctx_create.argtypes = [i32, P(Config), P(vp)]
    L.ghcm_ctx_destroy.argtypes = [vp]
    L.ghcm_ctx_get_config.argtypes = [vp, P(Config)]
    L.ghcm_ctx_set_config.argtypes = [vp, P(Config)]
    L.ghcm_solve.argtypes = [vp, vp, i32, u32, u32, vp, u32, vp, P(Stats)]
    L.ghcm_solve_device.argtypes = [vp, vp, i32, u32, u32, vp, u32, vp, P(Stats)]
    L.ghcm_solve_device_noalloc.argtypes = [vp, vp, i32, u32, u32, vp, u32, vp, P(Stats)]
    L.ghcm_solve3d.argtypes = [vp, vp, i32, u32, u32, u32, vp, u32, vp, P(Stats)]
    L.ghcm_solve3d_device.argtypes = [vp, vp, i32, u32, u32, u32, vp, u32, vp, P(Stats)]
    L.ghcm_solve_batch.argtypes = [vp, i32, P(Instance), u32, P(Stats)]
    L.ghcm_tune_default_space.argtypes = [i32, P(TuneCandidate), u32, P(u32)]
    L.ghcm_autotune.argtypes = [vp, i32, P(Instance), u32, P(TuneCandidate), u32, u32, P(u32)]
    L.ghcm_serve_config_default.argtypes = [P(ServeConfig)]
    L.ghcm_serve_config_default.restype = None
    L.ghcm_serve_start.argtypes = [i32, P(Config), P(ServeConfig), P(vp)]
    L.ghcm_serve_submit.argtypes = [vp, vp, u32, u32, vp, u32, vp, P(ctypes.c_uint64)]
    L.ghcm_serve_submit_device.argtypes = [vp, vp, u32, u32, vp, u32, vp, P(ctypes.c_uint64)]
    L.ghcm_serve_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_double, P(ServeResult)]
    L.ghcm_serve_get_stats.argtypes = [vp, P(ServeStats)]
    L.ghcm_serve_stop.argtypes = [vp, i32]
    L.ghcm_dev_alloc.argtypes = [vp, ctypes.c_size_t, P(vp)]
    L.ghcm_dev_free.argtypes = [vp, vp]
    L.ghcm_dev_upload.argtypes = [vp, vp, vp, ctypes.c_size_t]
    L.ghcm_dev_download.argtypes = [vp, vp, vp, ctypes.c_size_t]
    L.ghcm_dev_copy.argtypes = [vp, vp, vp, ctypes.c_size_t]
    L.ghcm_gen_ring_speed.argtypes = [u32, u32, ctypes.c_uint64, ctypes.c_double,
                                      ctypes.c_double, vp]
    L.ghcm_gen_smooth_speed.argtypes = [u32, u32, ctypes.c_uint64, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, vp]
    L.ghcm_gen_points.argtypes = [u32, u32, ctypes.c_uint64, u32, vp]
    L.ghcm_cpu_fmm.argtypes = [vp, u32, u32, vp, u32, vp, P(ctypes.c_double)]
    L.ghcm_cpu_fim.argtypes = [ctypes.c_float, vp, u32, u32, vp, u32, vp, P(ctypes.c_double)]
    L.ghcm_hcm_solve_cxx.argtypes = [i32, u32, ctypes.c_float, ctypes.c_float, i32, vp,
                                     u32, u32, vp, u32, vp, P(Stats)]
    PI = P(Image)
    L.ghcm_image_free.argtypes = [PI]
    L.ghcm_image_free.restype = None
    L.ghcm_image_load_f32.argtypes = [ctypes.c_char_p, i32, PI]
    L.ghcm_png_decode8.argtypes = [vp, ctypes.c_size_t, i32, P(u32), P(u32), P(u32), vp, ctypes.c_size_t]
    L.ghcm_png_write8.argtypes = [ctypes.c_char_p, vp, u32, u32, u32]
    L.ghcm_field_load_mono.argtypes = [ctypes.c_char_p, PI]
    L.ghcm_field_load_rgb_luminosity.argtypes = [vp, ctypes.c_char_p, PI]
    L.ghcm_rgb_luminosity_device.argtypes = [vp, vp, u32, u32, vp]
    L.ghcm_write_time_field_png.argtypes = [ctypes.c_char_p, vp, u32, u32, ctypes.c_float,
                                            ctypes.c_float, i32]
    L.ghcm_write_field_hdr.argtypes = [ctypes.c_char_p, vp, u32, u32]
    L.ghcm_blob_write.argtypes = [ctypes.c_char_p, vp, u32, u32]
    L.ghcm_blob_read.argtypes = [ctypes.c_char_p, PI]
    L.ghcm_test_run.argtypes = [ctypes.c_char_p, vp, ctypes.c_char_p, ctypes.c_char_p, vp, u32, u32,
                                vp, u32, ctypes.c_char_p, u32, ctypes.c_float]
    L.ghcm_field_cost_sin_checkerboard.argtypes = [u32, u32, ctypes.c_float, ctypes.c_float, vp]
    L.ghcm_field_cost_random.argtypes = [u32, u32, ctypes.c_uint64, vp]
    L.ghcm_tile_info.argtypes = [u32, u32, u32, u32, u32, u32, P(TileDesc)]
    L.ghcm_debug_select.argtypes = [vp, u32, vp, vp, vp, u32, u32, i32, vp, P(u32), vp, P(u32),
                                    vp, vp, vp, P(u32)]
    for name in EXPORTED_SYMBOLS:
        getattr(L, name).restype = getattr(L, name).restype or i32
    _lib = L
    return L


def _check(status: int) -> None:
    if status != OK:
        L = lib()
        raise GhcmError(status, (L.ghcm_last_error() or b"").decode() or
                        L.ghcm_status_string(status).decode())


def status_string(status: int) -> str:
    return lib().ghcm_status_string(status).decode()


def default_config(**overrides) -> Config:
    c = Config()
    lib().ghcm_config_default(ctypes.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise TypeError(f"unknown config field {k}")
        setattr(c, k, v)
    return c


def device_count() -> int:
    n = ctypes.c_int(0)
    st = lib().ghcm_device_count(ctypes.byref(n))
    return n.value if st == OK else 0


def device_info(ordinal: int = 0) -> dict:
    d = DeviceInfo()
    _check(lib().ghcm_device_info_get(ordinal, ctypes.byref(d)))
    return {"name": d.name.decode(), "arch": d.arch.decode(), "ordinal": d.ordinal,
            "compute_units": d.compute_units, "clock_khz": d.clock_khz,
            "lds_bytes": d.lds_bytes, "global_mem_bytes": d.global_mem_bytes}


def _seeds_array(seeds) -> np.ndarray:
    s = np.asarray(seeds, dtype=np.uint32).reshape(-1, 2)
    return np.ascontiguousarray(s)


def _field_array(field: np.ndarray, kind: int) -> np.ndarray:
    if kind in (FIELD_COST_F32, FIELD_SPEED_F32):
        return np.ascontiguousarray(field, dtype=np.float32)
    if kind == FIELD_SPEED_F16:
        return np.ascontiguousarray(field, dtype=np.float16)
    if kind == FIELD_SPEED_BF16:
        a = np.ascontiguousarray(field)
        if a.dtype != np.uint16:
            raise TypeError("bf16 fields are passed as uint16 bit patterns (see to_bf16)")
        return a
    raise ValueError(f"unknown field kind {kind}")


def to_bf16(x: np.ndarray) -> np.ndarray:
    """FP32 -> bfloat16 bit patterns (uint16), round to nearest even."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    rounded = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    nan = np.isnan(f)
    out = rounded.astype(np.uint16)
    out[nan] = 0x7FC0
    return out


def from_bf16(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


class Context:
    """A solver context on one HIP device (HcmSetup of src/eikonal.h:53-63)."""

    def __init__(self, ordinal: int = 0, **config):
        self._L = lib()
        self.config = default_config(**config)
        h = ctypes.c_void_p()
        _check(self._L.ghcm_ctx_create(ordinal, ctypes.byref(self.config), ctypes.byref(h)))
        self._h = h
        self.ordinal = ordinal

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.ghcm_ctx_destroy(self._h)
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

    def set_config(self, **changes) -> None:
        c = Config()
        _check(self._L.ghcm_ctx_get_config(self._h, ctypes.byref(c)))
        for k, v in changes.items():
            if not hasattr(c, k):
                raise TypeError(f"unknown config field {k}")
            setattr(c, k, v)
        _check(self._L.ghcm_ctx_set_config(self._h, ctypes.byref(c)))
        self.config = c

    def solve(self, field: np.ndarray, seeds, kind: int = FIELD_COST_F32,
              allow_not_converged: bool = False):
        """Solve one field held in host memory; returns (times (H, W) float32, stats)."""
        f = _field_array(field, kind)
        if f.ndim != 2:
            raise ValueError("field must be 2-D (H, W)")
        h, w = f.shape
        s = _seeds_array(seeds)
        out = np.empty((h, w), dtype=np.float32)
        st = Stats()
        rc = self._L.ghcm_solve(self._h, f.ctypes.data, kind, w, h, s.ctypes.data, s.shape[0],
                                out.ctypes.data, ctypes.byref(st))
        if not (allow_not_converged and rc == ERR_NOT_CONVERGED):
            _check(rc)
        return out, st.to_dict()

    def solve3d(self, field: np.ndarray, seeds, kind: int = FIELD_COST_F32,
                allow_not_converged: bool = False):
        """Solve one 3-D field (D, H, W) held in host memory; seeds are (x, y, z).
        Returns (times (D, H, W) float32, stats)."""
        f = _field_array(field, kind)
        if f.ndim != 3:
            raise ValueError("field must be 3-D (D, H, W)")
        d, h, w = f.shape
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint32).reshape(-1, 3))
        out = np.empty((d, h, w), dtype=np.float32)
        st = Stats()
        rc = self._L.ghcm_solve3d(self._h, f.ctypes.data, kind, w, h, d, s.ctypes.data, s.shape[0],
                                  out.ctypes.data, ctypes.byref(st))
        if not (allow_not_converged and rc == ERR_NOT_CONVERGED):
            _check(rc)
        return out, st.to_dict()

    def solve3d_device(self, d_field: int, kind: int, w: int, h: int, d: int, seeds,
                       d_time: int) -> dict:
        """Solve a 3-D field already resident in device memory (raw device pointers)."""
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint32).reshape(-1, 3))
        st = Stats()
        _check(self._L.ghcm_solve3d_device(self._h, ctypes.c_void_p(d_field), kind, w, h, d,
                                           s.ctypes.data, s.shape[0], ctypes.c_void_p(d_time),
                                           ctypes.byref(st)))
        return st.to_dict()

    def alloc(self, shape, dtype) -> "DeviceArray":
        """Device buffer on this context's device (freed with the object)."""
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    def solve_device(self, d_field: int, kind: int, w: int, h: int, seeds, d_time: int) -> dict:
        """Solve a field already resident in device memory (raw device pointers)."""
        s = _seeds_array(seeds)
        st = Stats()
        _check(self._L.ghcm_solve_device(self._h, ctypes.c_void_p(d_field), kind, w, h,
                                         s.ctypes.data, s.shape[0], ctypes.c_void_p(d_time),
                                         ctypes.byref(st)))
        return st.to_dict()

    def solve_device_noalloc(self, d_field: int, kind: int, w: int, h: int, seeds,
                             d_time: int) -> dict:
        """ghcm_solve_device_noalloc: the same solve with the scheduler state in the device's
        static scratch array (no allocation in the call)."""
        s = _seeds_array(seeds)
        st = Stats()
        _check(self._L.ghcm_solve_device_noalloc(self._h, d_field, kind, w, h, s.ctypes.data,
                                                 s.shape[0], d_time, ctypes.byref(st)))
        return st.to_dict()

    @staticmethod
    def _instances(instances):
        n = len(instances)
        arr = (Instance * n)()
        keep = []
        for i, (df, dt, w, h, seeds) in enumerate(instances):
            s = _seeds_array(seeds)
            keep.append(s)
            arr[i] = Instance(df, dt, w, h, s.ctypes.data, s.shape[0], 0)
        return arr, keep

    def solve_batch(self, kind: int, instances) -> dict:
        """instances: (d_field_ptr, d_time_ptr, w, h, seeds) tuples (device pointers), or a
        PreparedBatch from prepare_batch()."""
        if not isinstance(instances, PreparedBatch):
            instances = self.prepare_batch(instances)
        st = Stats()
        _check(self._L.ghcm_solve_batch(self._h, kind, instances.arr, instances.n,
                                        ctypes.byref(st)))
        return st.to_dict()

    def prepare_batch(self, instances: Sequence[tuple]) -> "PreparedBatch":
        """Build the ghcm_instance array of a batch once (a C caller keeps its array too)."""
        arr, keep = self._instances(instances)
        return PreparedBatch(arr, keep, len(instances))

    def autotune(self, kind: int, instances: Sequence[tuple], candidates=None, reps: int = 2):
        """Time candidate configurations on the instances; the fastest becomes the config.

        candidates: list of dicts (tile_w, tile_h, kernel, n_active, inner_iter_cap,
        launch_mode) or None for the library's default space.  Returns (best, results).
        """
        if candidates is None:
            cap = 1024
            buf = (TuneCandidate * cap)()
            n = ctypes.c_uint32(0)
            cus = device_info(self.ordinal)["compute_units"]
            _check(self._L.ghcm_tune_default_space(cus, buf, cap, ctypes.byref(n)))
            cands = (TuneCandidate * n.value)(*buf[:n.value])
        else:
            cands = (TuneCandidate * len(candidates))()
            for i, c in enumerate(candidates):
                cands[i] = TuneCandidate(c["tile_w"], c["tile_h"], c["kernel"], c["n_active"],
                                         c.get("inner_iter_cap", 0), c.get("launch_mode", 0),
                                         0.0, 0, 0, 0)
        arr, keep = self._instances(instances)
        best = ctypes.c_uint32(0)
        _check(self._L.ghcm_autotune(self._h, kind, arr, len(instances), cands, len(cands),
                                     reps, ctypes.byref(best)))
        c = Config()
        _check(self._L.ghcm_ctx_get_config(self._h, ctypes.byref(c)))
        self.config = c
        return cands[best.value].to_dict(), [x.to_dict() for x in cands]


    def debug_select(self, tiles_total: int, tiles, keys, masks, n: int, handoff: bool = False):
        """One device wave boundary on a given list (test support, ghcm_debug_select):
        returns (jobs, handed, next_tiles, next_keys, next_masks)."""
        t = np.ascontiguousarray(np.asarray(tiles, dtype=np.uint32))
        k = np.ascontiguousarray(np.asarray(keys, dtype=np.float32))
        m = np.ascontiguousarray(np.asarray(masks, dtype=np.uint32))
        c = t.shape[0]
        assert k.shape == (c,) and m.shape == (c,)
        cap = max(c, 1)
        jobs, handed, nxt = (np.zeros(cap, np.uint32) for _ in range(3))
        nkeys, nmasks = np.zeros(cap, np.float32), np.zeros(cap, np.uint32)
        nj, nh, nn = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
        _check(self._L.ghcm_debug_select(self._h, tiles_total, t.ctypes.data, k.ctypes.data,
                                         m.ctypes.data, c, n, 1 if handoff else 0,
                                         jobs.ctypes.data, ctypes.byref(nj), handed.ctypes.data,
                                         ctypes.byref(nh), nxt.ctypes.data, nkeys.ctypes.data,
                                         nmasks.ctypes.data, ctypes.byref(nn)))
        return (jobs[:nj.value], handed[:nh.value], nxt[:nn.value], nkeys[:nn.value],
                nmasks[:nn.value])


def tile_info(w: int, h: int, tile_w: int, tile_h: int, tile: int, tile_base: int = 0) -> dict:
    """Geometry of one tile as the scheduler sees it (ghcm_tile_info, host only)."""
    d = TileDesc()
    _check(lib().ghcm_tile_info(w, h, tile_w, tile_h, tile_base, tile, ctypes.byref(d)))
    return d.to_dict()


class Server:
    """Continuous-batching solve server (ghcm_serve_*): one device wave loop that admits
    requests at wave boundaries while earlier ones are still being solved.

    ``submit`` (host arrays; the field is copied before it returns) and ``submit_device``
    (device pointers, results written in place) return a ticket; ``wait`` collects it."""

    def __init__(self, ordinal: int = 0, kind: int = FIELD_COST_F32, hbm_budget_bytes: int = 0,
                 max_slots: int = 0, max_tiles: int = 0, **config):
        self._L = lib()
        self.kind = kind
        self.config = default_config(**config)
        sc = ServeConfig()
        self._L.ghcm_serve_config_default(ctypes.byref(sc))
        if hbm_budget_bytes:
            sc.hbm_budget_bytes = hbm_budget_bytes
        if max_slots:
            sc.max_slots = max_slots
        if max_tiles:
            sc.max_tiles = max_tiles
        sc.field_kind = kind
        self.serve_config = sc
        h = ctypes.c_void_p()
        _check(self._L.ghcm_serve_start(ordinal, ctypes.byref(self.config), ctypes.byref(sc),
                                        ctypes.byref(h)))
        self._h = h
        self._held = {}

    def submit(self, field: np.ndarray, seeds) -> int:
        f = _field_array(field, self.kind)
        if f.ndim != 2:
            raise ValueError("field must be 2-D (H, W)")
        h, w = f.shape
        s = _seeds_array(seeds)
        out = np.empty((h, w), dtype=np.float32)
        t = ctypes.c_uint64(0)
        _check(self._L.ghcm_serve_submit(self._h, f.ctypes.data, w, h, s.ctypes.data, s.shape[0],
                                         out.ctypes.data, ctypes.byref(t)))
        self._held[t.value] = out
        return t.value

    def submit_device(self, d_field: int, w: int, h: int, seeds, d_time: int) -> int:
        """Device-resident request: raw device pointers on the server's device."""
        s = _seeds_array(seeds)
        t = ctypes.c_uint64(0)
        _check(self._L.ghcm_serve_submit_device(self._h, ctypes.c_void_p(d_field), w, h,
                                                s.ctypes.data, s.shape[0], ctypes.c_void_p(d_time),
                                                ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_s: float = -1.0):
        """(times (H, W) float32 or None for a device request, result dict); raises
        GhcmError on failure or time-out."""
        r = ServeResult()
        r.status = -1
        rc = self._L.ghcm_serve_wait(self._h, ticket, timeout_s, ctypes.byref(r))
        # The output buffer is released only by the call that collected the ticket: a
        # refused second waiter must not free it while the collecting call copies into it.
        out = self._held.pop(ticket, None) if r.status >= 0 else None
        _check(rc)
        return out, r.to_dict()

    def stats(self) -> dict:
        st = ServeStats()
        _check(self._L.ghcm_serve_get_stats(self._h, ctypes.byref(st)))
        return st.to_dict()

    def stop(self, drain: bool = True) -> None:
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            rc = self._L.ghcm_serve_stop(h, 1 if drain else 0)
            self._held.clear()
            _check(rc)

    def __del__(self):
        try:
            self.stop(drain=False)
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop(drain=exc[0] is None)


class DeviceArray:
    """A device allocation made through the C-ABI (no other GPU runtime needed)."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(ctx._L.ghcm_dev_alloc(ctx._h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(self.ctx._L.ghcm_dev_upload(self.ctx._h, ctypes.c_void_p(self.ptr),
                                           a.ctypes.data, self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.ctx._L.ghcm_dev_download(self.ctx._h, out.ctypes.data,
                                             ctypes.c_void_p(self.ptr), self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr and getattr(self.ctx, "_h", None):
            self.ctx._L.ghcm_dev_free(self.ctx._h, ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- CPU solvers of the C++ API (eikonal_fmm_sequential / eikonal_fim_sequential) ----

def fmm_sequential(cost: np.ndarray, seeds):
    c = np.ascontiguousarray(cost, dtype=np.float32)
    h, w = c.shape
    s = _seeds_array(seeds)
    out = np.empty_like(c)
    us = ctypes.c_double(0)
    _check(lib().ghcm_cpu_fmm(c.ctypes.data, w, h, s.ctypes.data, s.shape[0],
                              out.ctypes.data, ctypes.byref(us)))
    return out, us.value


def fim_sequential(minimum_change: float, cost: np.ndarray, seeds):
    c = np.ascontiguousarray(cost, dtype=np.float32)
    h, w = c.shape
    s = _seeds_array(seeds)
    out = np.empty_like(c)
    us = ctypes.c_double(0)
    _check(lib().ghcm_cpu_fim(minimum_change, c.ctypes.data, w, h, s.ctypes.data, s.shape[0],
                              out.ctypes.data, ctypes.byref(us)))
    return out, us.value


def hcm_solve_cxx(cost: np.ndarray, seeds, block_size: int = 8, compute_unit_overload: float = 1.0,
                  fim_minimum_change: float = 1e-5, kernel: int = KERNEL_SYSTOLIC,
                  ordinal: int = 0):
    """eikonal_hcm_hip_setup + eikonal_hcm_hip_ex through the C++ layer."""
    c = np.ascontiguousarray(cost, dtype=np.float32)
    h, w = c.shape
    s = _seeds_array(seeds)
    out = np.empty_like(c)
    st = Stats()
    _check(lib().ghcm_hcm_solve_cxx(ordinal, block_size, compute_unit_overload,
                                    fim_minimum_change, kernel, c.ctypes.data, w, h,
                                    s.ctypes.data, s.shape[0], out.ctypes.data,
                                    ctypes.byref(st)))
    return out, st.to_dict()


# ---- synthetic workloads ----

def gen_ring_speed(w: int, h: int, seed: int, wall_speed: float = 1e-3,
                   scale: float = 1.0) -> np.ndarray:
    out = np.empty((h, w), dtype=np.float64)
    _check(lib().ghcm_gen_ring_speed(w, h, seed, wall_speed, scale, out.ctypes.data))
    return out


def gen_smooth_speed(w: int, h: int, seed: int, sigma: float = 32.0, lo: float = 0.1,
                     hi: float = 10.0) -> np.ndarray:
    out = np.empty((h, w), dtype=np.float64)
    _check(lib().ghcm_gen_smooth_speed(w, h, seed, sigma, lo, hi, out.ctypes.data))
    return out


def gen_points(w: int, h: int, seed: int, n: int) -> np.ndarray:
    out = np.empty((n, 2), dtype=np.uint32)
    _check(lib().ghcm_gen_points(w, h, seed, n, out.ctypes.data))
    return out


def cost_from_speed(speed: np.ndarray) -> np.ndarray:
    """cost = 1/speed in FP32 (src/test_runner.cpp:32-36); speed 0 -> +INF."""
    with np.errstate(divide="ignore"):
        return (np.float32(1.0) / np.asarray(speed, dtype=np.float32)).astype(np.float32)


# ---- field loaders, writers and the test runner (include/ghcm_io.h) -------------------

def _take_image(img: Image) -> np.ndarray:
    n = img.w * img.h * img.comp
    arr = np.ctypeslib.as_array(img.data, shape=(n,)).copy() if n else np.zeros(0, np.float32)
    lib().ghcm_image_free(ctypes.byref(img))
    return arr


def _path(p) -> bytes:
    return os.fsencode(p)


def png_decode8(data: bytes, req_comp: int = 0) -> np.ndarray:
    """PNG bytes -> (H, W, C) uint8, converted as stb_image does for `req_comp` channels."""
    L = lib()
    w, h, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    _check(L.ghcm_png_decode8(buf, len(data), req_comp, ctypes.byref(w), ctypes.byref(h),
                              ctypes.byref(c), None, 0))
    out = np.empty((h.value, w.value, c.value), dtype=np.uint8)
    _check(L.ghcm_png_decode8(buf, len(data), req_comp, ctypes.byref(w), ctypes.byref(h),
                              ctypes.byref(c), out.ctypes.data, out.nbytes))
    return out


def png_write8(path, pixels: np.ndarray) -> None:
    p = np.ascontiguousarray(pixels, dtype=np.uint8)
    if p.ndim == 2:
        p = p[:, :, None]
    _check(lib().ghcm_png_write8(_path(path), p.ctypes.data, p.shape[1], p.shape[0], p.shape[2]))


def image_load_f32(path, req_comp: int) -> np.ndarray:
    """stbi_loadf(path, req_comp) equivalent: (H, W, C) float32."""
    img = Image()
    _check(lib().ghcm_image_load_f32(_path(path), req_comp, ctypes.byref(img)))
    w, h, c = img.w, img.h, img.comp
    return _take_image(img).reshape(h, w, c)


def field_load_mono(path) -> np.ndarray:
    """field2d_img_monochannel (src/field_loaders.cpp:53-59): (H, W) float32."""
    img = Image()
    _check(lib().ghcm_field_load_mono(_path(path), ctypes.byref(img)))
    w, h = img.w, img.h
    return _take_image(img).reshape(h, w)


def field_load_rgb_luminosity(ctx: "Context", path) -> np.ndarray:
    """field2d_img_rgb_luminosity (src/field_loaders.cpp:62-102) on ctx's GPU."""
    img = Image()
    _check(lib().ghcm_field_load_rgb_luminosity(ctx._h, _path(path), ctypes.byref(img)))
    w, h = img.w, img.h
    return _take_image(img).reshape(h, w)


def write_time_field_png(path, field: np.ndarray, f_min=None, f_max=None) -> None:
    f = np.ascontiguousarray(field, dtype=np.float32)
    auto = f_min is None or f_max is None
    _check(lib().ghcm_write_time_field_png(_path(path), f.ctypes.data, f.shape[1], f.shape[0],
                                           0.0 if auto else f_min, 0.0 if auto else f_max,
                                           1 if auto else 0))


def write_field_hdr(path, field: np.ndarray) -> None:
    f = np.ascontiguousarray(field, dtype=np.float32)
    _check(lib().ghcm_write_field_hdr(_path(path), f.ctypes.data, f.shape[1], f.shape[0]))


def blob_write(path, field: np.ndarray) -> None:
    f = np.ascontiguousarray(field, dtype=np.float32)
    _check(lib().ghcm_blob_write(_path(path), f.ctypes.data, f.shape[1], f.shape[0]))


def blob_read(path) -> np.ndarray:
    img = Image()
    _check(lib().ghcm_blob_read(_path(path), ctypes.byref(img)))
    w, h = img.w, img.h
    return _take_image(img).reshape(h, w)


def test_run(solver: str, cost: np.ndarray, seeds, out_dir, solver_name=None, job_name="job",
             ctx: "Context" = None, trials: int = 5, fim_minimum_change: float = 1e-3) -> None:
    """test_run (src/test_runner.cpp:163-217) for one job with solver "fmm", "fim" or "hcm"."""
    c = np.ascontiguousarray(cost, dtype=np.float32)
    s = _seeds_array(seeds)
    _check(lib().ghcm_test_run(solver.encode(), ctx._h if ctx else None,
                               (solver_name or solver).encode(), job_name.encode(), c.ctypes.data,
                               c.shape[1], c.shape[0], s.ctypes.data, s.shape[0], _path(out_dir),
                               trials, fim_minimum_change))


test_run.__test__ = False   # not a pytest test


def field_cost_sin_checkerboard(w: int, h: int, speed_amplitude: float, checkerboard_res: float):
    out = np.empty((h, w), dtype=np.float32)
    _check(lib().ghcm_field_cost_sin_checkerboard(w, h, speed_amplitude, checkerboard_res,
                                                  out.ctypes.data))
    return out


def field_cost_random(w: int, h: int, seed: int):
    out = np.empty((h, w), dtype=np.float32)
    _check(lib().ghcm_field_cost_random(w, h, seed, out.ctypes.data))
    return out
