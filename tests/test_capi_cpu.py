"""C-ABI and C++-layer checks that need no GPU: symbols, defaults, errors, CPU solvers."""
import ctypes
import os
import re

import numpy as np
import pytest

import ghcm_amd as g
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in ("ghcm.h", "eikonal_c.h", "ghcm_io.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ghcm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = g.lib()
    declared = _declared_functions()
    assert declared == set(g.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(L, name), name


def test_abi_and_status_strings():
    assert g.lib().ghcm_abi_version() == 1
    assert g.status_string(g.OK) == "ok"
    assert "converge" in g.status_string(g.ERR_NOT_CONVERGED)
    assert g.status_string(999) == "unknown status"


def test_config_defaults_mirror_hcmconfig():
    # HcmConfig defaults, src/eikonal.h:16-26
    c = g.default_config()
    assert (c.tile_w, c.tile_h) == (8, 8)
    assert c.workers_overload == 1.0 and c.compute_unit_overload == 1.0
    assert c.min_change == pytest.approx(1e-5)
    assert c.cache_mode == 3 and c.kernel == g.KERNEL_SYSTOLIC
    assert c.scheduler == g.SCHED_DEVICE_TOPN


def test_struct_sizes_match_header():
    assert ctypes.sizeof(g.Config) == 56
    assert ctypes.sizeof(g.Stats) == 104
    assert ctypes.sizeof(g.Instance) == 40
    assert ctypes.sizeof(g.ServeConfig) == 32
    assert ctypes.sizeof(g.ServeResult) == 40
    assert ctypes.sizeof(g.ServeStats) == 112


def test_struct_layouts_match_c_compiler(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives ghcm.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"ghcm_config": g.Config, "ghcm_stats": g.Stats, "ghcm_instance": g.Instance,
               "ghcm_image": g.Image,
               "ghcm_tune_candidate": g.TuneCandidate, "ghcm_device_info": g.DeviceInfo,
               "ghcm_serve_config": g.ServeConfig, "ghcm_serve_result": g.ServeResult,
               "ghcm_serve_stats": g.ServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ghcm.h"', '#include "ghcm_io.h"',
             "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, got):
        cname, what, val = line.split()
        cls = mirrors[cname]
        want = ctypes.sizeof(cls) if what == "size" else getattr(cls, what).offset
        assert int(val) == want, (cname, what)


@pytest.mark.skipif(g.device_count() > 0, reason="checks the no-device error path")
def test_no_device_errors_are_reported_not_fatal():
    with pytest.raises(g.GhcmError) as e:
        g.Context(0)
    assert e.value.status == g.ERR_NO_DEVICE
    assert g.device_count() == 0


def test_invalid_config_rejected_before_device():
    L = g.lib()
    h = ctypes.c_void_p()
    bad = g.default_config(tile_w=7)
    rc = L.ghcm_ctx_create(0, ctypes.byref(bad), ctypes.byref(h))
    assert rc in (g.ERR_UNSUPPORTED, g.ERR_NO_DEVICE)
    bad = g.default_config(kernel=5)
    assert L.ghcm_ctx_create(0, ctypes.byref(bad), ctypes.byref(h)) == g.ERR_INVALID_ARG
    bad = g.default_config(min_change=-1.0)
    assert L.ghcm_ctx_create(0, ctypes.byref(bad), ctypes.byref(h)) == g.ERR_INVALID_ARG
    assert L.ghcm_ctx_create(0, None, None) == g.ERR_INVALID_ARG


def test_cpu_solvers_equal_oracle(rng):
    for _ in range(3):
        h, w = rng.integers(2, 90, 2)
        c = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
        c[rng.random((h, w)) < 0.1] = np.inf
        seeds = [(int(rng.integers(0, w)), int(rng.integers(0, h)))]
        c[seeds[0][1], seeds[0][0]] = 1.0
        a, _ = g.fmm_sequential(c, seeds)
        assert np.array_equal(a, oracle.fmm_f32(c, seeds))
        b, _ = g.fim_sequential(1e-5, c, seeds)
        ob, _ = oracle.fim_f32(c, seeds, 1e-5)
        assert np.array_equal(b, ob)


def test_cpu_solvers_reject_bad_seed():
    c = np.ones((3, 3), dtype=np.float32)
    with pytest.raises(g.GhcmError):
        g.fmm_sequential(c, [(3, 0)])
    with pytest.raises(g.GhcmError):
        g.fim_sequential(1e-5, c, [(0, 3)])


def test_bf16_round_trip():
    x = np.array([1.0, 0.1, 10.0, 1e-4, 65504.0, 3.0e38], dtype=np.float32)
    b = g.to_bf16(x)
    y = g.from_bf16(b)
    assert np.all(np.abs(y - x) <= np.abs(x) * 2.0 ** -8)
    assert g.to_bf16(np.array([1.0], np.float32))[0] == 0x3F80
