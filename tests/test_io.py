"""The tools' output containers (include/qew/qew_io.hpp): the reference's
writer tests (tests/test_executables.cpp:199-289) restated. The C++ writers run
in tests/cpp/io_selftest.cpp (host only); the files are read back here with
qew_amd.io (HDF5 through ctypes, and the text archive) and, independently of
that reader, with the HDF5 distribution's h5dump."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from qew_amd import build as qbuild
from qew_amd.io import read_h5, read_kv, read_output

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "qew_amd")


def _h5dump():
    for r in ("/usr/local", "/opt/conda"):
        p = os.path.join(r, "bin", "h5dump")
        if os.path.exists(p):
            return p
    return shutil.which("h5dump")


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    flags = qbuild.hdf5_flags()
    if not flags:
        pytest.skip("no HDF5 headers and library on this host")
    d = tmp_path_factory.mktemp("io")
    exe = str(d / "io_selftest")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", INC, "-I", "/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", *flags, os.path.join(ROOT, "tests", "cpp", "io_selftest.cpp"),
                    "-o", exe, "-L", PKG, "-lqew_amd", "-ldl", f"-Wl,-rpath,{PKG}"],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe, str(d)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
    return d


def _expect_static(f):
    # WriteStaticKeysAndValues (tests/test_executables.cpp:201-229)
    assert f["physical_size"] == 4.0
    prefix = "driving_force=0/pinning_length=0.1"
    assert np.array_equal(f[prefix + "/h"], [1.0, 2.0, 3.0, 4.0])
    assert f[prefix + "/pinning_length"] == 0.1


def _expect_dynamic(f):
    # WriteDynamicKeysAndValues (tests/test_executables.cpp:231-255)
    prefix = "timestep=1/driving_force=0/line_tension=2"
    assert np.array_equal(f[prefix + "/h"], [5.0, 6.0])
    assert f[prefix + "/pinning_length"] == pytest.approx(0.1, rel=0, abs=0)
    assert f["physical_size"] == 2.0


def _expect_depinning(f):
    # WriteDepinningSkipsRunawayAndWritesFc (tests/test_executables.cpp:257-288)
    assert np.array_equal(f["driving_force=0/h"], [0.5, 0.5, 0.5])
    assert f["driving_force=0/pinning_length"] == 5.0
    assert "driving_force=0.8/h" not in f  # the runaway force stores no profile
    assert f["f_c_upper_bound"] == 0.8
    assert set(f) == {"physical_size", "driving_force=0/h", "driving_force=0/pinning_length",
                      "f_c_upper_bound"}


@pytest.mark.parametrize("ext", [".h5", ".kv"])
def test_writer_keys_and_values(written, ext):
    _expect_static(read_output(written / f"test_write_static{ext}"))
    _expect_dynamic(read_output(written / f"test_write_dynamic{ext}"))
    _expect_depinning(read_output(written / f"test_write_depinning{ext}"))


def test_hdf5_and_text_archive_agree(written):
    for name in ("static", "dynamic", "depinning"):
        a = read_h5(written / f"test_write_{name}.h5")
        b = read_kv(written / f"test_write_{name}.kv")
        assert set(a) == set(b)
        for k in a:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_dispatch_by_suffix(written):
    assert read_kv(written / "dispatch.kv") == {"x": 3.0}
    assert read_h5(written / "dispatch.h5") == {"x": 4.0}
    with open(written / "dispatch.h5", "rb") as f:
        assert f.read(8) == b"\x89HDF\r\n\x1a\n"  # the HDF5 signature


def test_h5dump_reads_the_same_layout(written):
    """An independent reader: h5dump prints every dataset with its dataspace
    (SCALAR for single values, SIMPLE {n} for profiles) and FP64 type."""
    exe = _h5dump()
    if exe is None:
        pytest.skip("h5dump not found")
    r = subprocess.run([exe, str(written / "test_write_static.h5")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert re.search(r'GROUP "driving_force=0"\s*{\s*GROUP "pinning_length=0.1"', out)
    m = re.search(r'DATASET "h" {\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE { \( 4 \) / \( 4 \) }'
                  r'\s*DATA {\s*\(0\): 1, 2, 3, 4', out)
    assert m, out
    m = re.search(r'DATASET "physical_size" {\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SCALAR'
                  r'\s*DATA {\s*\(0\): 4', out)
    assert m, out
