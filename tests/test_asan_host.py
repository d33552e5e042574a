"""Host AddressSanitizer run of the C ABI library's host code (argument and
basis checks, pair and enumeration tables, the host 2-RDM / 1-RDM builders)
and of the C oracle, on the golden s/p/d and X-Y-X CAS(4,4) records
(tests/asan/host_selftest.cpp, built by tests/asan/Makefile with only host
code instrumented).  Without a GPU the device calls return THISTLE_EDEVICE,
which the self-test expects."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN = os.path.join(ROOT, "tests", "asan")
EXE = os.path.join(ASAN, "_build", "host_selftest")


@pytest.fixture(scope="module")
def selftest():
    if not shutil.which("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not found")
    lib = os.path.join(ROOT, "thistle_amd", "csrc", "_build", "capi.o")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "thistle_amd", "csrc")], check=True)
    subprocess.run(["make", "-C", ASAN], check=True, capture_output=True)
    return EXE


def _record(path, g):
    ngto = g["l"].size
    confs = np.asarray(g["confs"], dtype=np.int32)
    civs = np.asarray(g["civs"], dtype=np.float64)
    if civs.ndim == 1:
        civs = civs[:, None]
    hdr = np.array([ngto, int(g["group"].max()), g["mmod"].shape[0], confs.shape[0],
                    confs.shape[1], civs.shape[1], g["q"].size], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        for k in ("l", "m", "n", "group"):
            f.write(np.asarray(g[k], dtype=np.int32).tobytes())
        for k in ("ga", "xx", "yy", "zz"):
            f.write(np.asarray(g[k], dtype=np.float64).tobytes())
        f.write(np.asfortranarray(g["mmod"]).tobytes(order="F"))
        f.write(np.asfortranarray(confs).tobytes(order="F"))
        f.write(np.asfortranarray(civs).tobytes(order="F"))
        f.write(np.asarray(g["q"], dtype=np.float64).tobytes())


@pytest.mark.parametrize("name", ["spd2", "xyx_cas44", "far"])
def test_host_code_under_asan(selftest, golden, name, tmp_path):
    rec = tmp_path / "rec.bin"
    _record(rec, golden(name))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               OMP_NUM_THREADS="2", HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    r = subprocess.run([selftest, str(rec)], capture_output=True, text=True, env=env,
                       timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert "AddressSanitizer" not in r.stderr
    assert r.returncode == 0
    assert "fails=0" in r.stdout
