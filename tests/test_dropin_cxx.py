"""The drop-in C++ boundary compiled from outside the library: tests/dropin/dropin_caller.cpp
is the caller INTEGRATION.md section 1 describes (include/eikonal.h in place of src/eikonal.h,
a solver bound into SolverFn with std::bind as src/test_runner.cpp:117, 243-246 does), built
by g++ against the installed headers and linked to libghcm.so.  The CPU solvers run in the
CPU suite; eikonal_hcm_opencl_ex (the GPU solver) under -m gpu.  Results are checked against
the oracle."""
import os
import subprocess

import numpy as np
import pytest

import ghcm_amd as g
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "dropin", "dropin_caller.cpp")
LIBDIR = os.path.join(ROOT, "ghcm_amd")


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    g.lib()                                 # the library is built
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_caller")
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", "/opt/rocm/include", SRC, "-o", exe, "-L", LIBDIR, "-lghcm",
           f"-Wl,-rpath,{LIBDIR}", "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, tmp_path, solver, cost, seeds):
    h, w = cost.shape
    cin, tout = tmp_path / "cost.bin", tmp_path / "time.bin"
    np.ascontiguousarray(cost, np.float32).tofile(cin)
    args = [exe, solver, str(cin), str(w), str(h), str(tout)]
    for x, y in seeds:
        args += [str(x), str(y)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return np.fromfile(tout, np.float32).reshape(h, w)


def _field():
    speed = g.gen_ring_speed(200, 150, 21, 1e-3, 0.6)
    return g.cost_from_speed(speed), [(100, 75), (3, 140)]


def test_dropin_cpu_solvers_bound_into_solverfn(caller, tmp_path):
    cost, seeds = _field()
    t_fmm = _run(caller, tmp_path, "fmm", cost, seeds)
    # eikonal_fmm_sequential restates src/eikonal_fmm.cpp step for step: bit-equal to the oracle
    assert np.array_equal(t_fmm, oracle.fmm_f32(cost, seeds))
    t_fim = _run(caller, tmp_path, "fim", cost, seeds)
    ref_fim, _ = oracle.fim_f32(cost, seeds, 1e-3)
    assert np.array_equal(t_fim, ref_fim)


def test_dropin_usage_errors(caller):
    r = subprocess.run([caller, "fmm"], capture_output=True, text=True)
    assert r.returncode == 2


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["hcm", "hcm_sweep"])
def test_dropin_gpu_solver_bound_into_solverfn(caller, tmp_path, solver):
    cost, seeds = _field()
    print(f"case dropin c++ {solver}", flush=True)
    t = _run(caller, tmp_path, solver, cost, seeds)
    ref = oracle.fmm_f32(cost, seeds)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(t), fin)
    assert np.max(np.abs(t[fin].astype(np.float64) - ref[fin])) <= 1e-4 * ref[fin].max()
    for x, y in seeds:
        assert t[y, x] == 0.0
