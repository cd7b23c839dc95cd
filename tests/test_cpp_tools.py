"""The C++ side of the boundary: the CLI parser and the command-line tools
(CPU: parsing and error exits, no device work), the facade and run drivers
(GPU: tests/cpp/drivers_selftest.cpp restates the reference's
tests/test_executables.cpp), and the tools' outputs on small problems (GPU)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "qew_amd", "bin")
INC = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "qew_amd")


def tool(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not built")
    return p


def compile_cpp(src, out, link=False):
    cmd = ["g++", "-O1", "-std=c++17", "-I", INC, os.path.join(ROOT, "tests", "cpp", src), "-o", out]
    if link:
        cmd += ["-L", PKG, "-lqew_amd", f"-Wl,-rpath,{PKG}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_cli_parser_selftest(tmp_path):
    exe = compile_cpp("cli_selftest.cpp", str(tmp_path / "cli_selftest"))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout


def test_division_free_weights_selftest(tmp_path):
    """tests/cpp/div_selftest.cpp: the device weights' x / 6 without a
    division equals the correctly rounded division (host build of the same
    header)."""
    exe = compile_cpp("div_selftest.cpp", str(tmp_path / "div_selftest"))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout


@pytest.mark.parametrize("name", ["qew_static", "qew_dynamic", "qew_depinning", "qew_benchmark"])
def test_tools_help_and_option_errors(name):
    exe = tool(name)
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith(name)
    for bad in (["--bogus", "1"], ["--nx", "abc"], ["--nx"], ["positional"],
                ["--model", "elastic"]):
        r = subprocess.run([exe, *bad], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (bad, r.stdout, r.stderr)
        assert r.stderr.startswith(name + ":")


@pytest.mark.gpu
def test_facade_and_drivers_selftest(tmp_path):
    exe = compile_cpp("drivers_selftest.cpp", str(tmp_path / "drivers_selftest"), link=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_facade_per_point_api_selftest(tmp_path):
    """tests/cpp/facade_selftest.cpp: DeviceNoise::sample / sample_value /
    sample_dy, rk_site_velocity, rk_forward_zero and GradientWorkspace through
    the C++ facade (the reference's sampler and site-velocity tests)."""
    exe = compile_cpp("facade_selftest.cpp", str(tmp_path / "facade_selftest"), link=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_tools_write_reference_keys(tmp_path):
    from qew_amd.io import read_kv, read_output
    # no --output: the reference's default name, static_<model>.h5, in HDF5
    r = subprocess.run([tool("qew_static"), "--nx", "256", "--ny", "64", "--Lx", "2.56",
                        "--pinning_lengths", "1,10", "--model", "linear",
                        "--kokkos-num-threads=4"], capture_output=True, text=True, timeout=300,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    kv = read_output(tmp_path / "static_linear.h5")
    assert kv["physical_size"] == 2.56
    for lp in ("1", "10"):
        assert kv[f"driving_force=0/pinning_length={lp}/pinning_length"] == float(lp)
        assert kv[f"driving_force=0/pinning_length={lp}/h"].shape == (256,)
    assert r.stdout.count("converged") == 2
    # --batched 0: the reference's one minimisation per length, same minima
    r0 = subprocess.run([tool("qew_static"), "--nx", "256", "--ny", "64", "--Lx", "2.56",
                         "--pinning_lengths", "1,10", "--model", "linear", "--batched", "0",
                         "--output", str(tmp_path / "s0.kv")], capture_output=True, text=True,
                        timeout=300)
    assert r0.returncode == 0, r0.stderr
    kv0 = read_kv(tmp_path / "s0.kv")
    for lp in ("1", "10"):
        key = f"driving_force=0/pinning_length={lp}/h"
        assert np.max(np.abs(kv0[key] - kv[key])) < 1e-5

    out = tmp_path / "d.kv"
    r = subprocess.run([tool("qew_dynamic"), "--nx", "128", "--ny", "64", "--total_time", "20",
                        "--output", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    kv = read_kv(out)
    snaps = [k for k in kv if k.endswith("/h")]
    assert snaps and all(k.startswith("timestep=") for k in snaps)
    assert all(np.all(np.isfinite(kv[k])) for k in snaps)

    out = tmp_path / "p.h5"
    r = subprocess.run([tool("qew_depinning"), "--nx", "64", "--ny", "64", "--f_step", "2",
                        "--fc_rel_tol", "1e-3", "--output", str(out)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    kv = read_output(out)
    assert "f_c_upper_bound" in kv and kv["driving_force=0/h"].shape == (64,)
    assert kv["f_c_lower"] <= kv["f_c"] <= kv["f_c_upper"] <= kv["f_c_upper_bound"]
    assert "Line depinned near" in r.stdout


@pytest.mark.gpu
def test_benchmark_tool_prints_reference_summary():
    """qew_benchmark restates executables/benchmark/main.cpp: kernel timings on
    a 262144-site line, then a preconditioned L-BFGS relaxation of a
    qew_static-shaped line (n = 4096) that converges."""
    import re
    r = subprocess.run([tool("qew_benchmark"), "--reps", "3"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert re.match(r"device= 0  setup= *[0-9.]+s  grad= *[0-9.]+ ms/ea  obj= *[0-9.]+ ms/ea  "
                    r"dyn= *[0-9.]+ ms/ea$", lines[0]), lines[0]
    m = re.match(r"lbfgs\(n=4096, precond\): *[0-9.]+s  iters=(\d+)  feval=(\d+)  converged$", lines[1])
    assert m, lines[1]
    assert 0 < int(m.group(1)) <= int(m.group(2))
