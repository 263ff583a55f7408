import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    return {k: d[k] for k in d.files}


@pytest.fixture
def golden():
    return load_golden

EMU_DIR = os.path.join(ROOT, "tests", "emu")
EMU_LIB = os.path.join(EMU_DIR, "libpscf_emu.so")
SRC_DIR = os.path.join(ROOT, "pscf_amd", "csrc")


def _emu_stale():
    if not os.path.exists(EMU_LIB):
        return True
    t = os.path.getmtime(EMU_LIB)
    srcs = [os.path.join(SRC_DIR, f) for f in os.listdir(SRC_DIR)]
    srcs.append(os.path.join(EMU_DIR, "hip", "hip_runtime.h"))
    return any(os.path.getmtime(f) > t for f in srcs)


_libs = {}


def get_lib(kind):
    """'emu': the library's sources compiled for the CPU emulator (tests only);
    'gpu': the product gfx950 library."""
    if kind in _libs:
        return _libs[kind]
    from pscf_amd._lib import Lib, DEFAULT_LIB
    if kind == "emu":
        import fcntl
        with open(os.path.join(EMU_DIR, ".build.lock"), "w") as lk:   # one builder at a time
            fcntl.flock(lk, fcntl.LOCK_EX)
            if _emu_stale():
                import subprocess
                subprocess.run([os.path.join(EMU_DIR, "build_emu.sh")], check=True)
        L = Lib(EMU_LIB)
        L.tensor_device = "cpu"      # emulated device memory is host memory
    else:
        L = Lib(DEFAULT_LIB)
        assert L.device_count() >= 1, "no HIP device visible"
    _libs[kind] = L
    return L


@pytest.fixture(params=["emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def backend(request):
    return request.param
