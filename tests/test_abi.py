"""The C ABI: libiac_hip.so loads, exports every function include/iac_hip.h
declares, and reports errors through return codes (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import incohautocorrelate_amd as iac
from incohautocorrelate_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    txt = (ROOT / "include" / "iac_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(iac_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    lib = C.CDLL(str(_abi.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) <= set(_abi.SIGNATURES)


def test_strerror_and_version():
    L = iac.lib()
    assert L.iac_abi_version() == 1
    for code in (0, -1, -2, -3, -4, -5, -6, -7, -8):
        assert L.iac_strerror(code)
    assert L.iac_strerror(-2).decode().startswith("mesh size must be odd")


def test_null_and_bad_arguments_return_codes():
    L = iac.lib()
    assert L.iac_mesh_shape_big(None, 63, 1.0, None) == -1
    s = iac.MeshShape()
    mq = np.array([0.1, 0.1, 0.01], np.float32)
    assert L.iac_mesh_shape_big(_abi.fptr(mq), 64, 1.0, C.byref(s)) == -2
    assert L.iac_mesh_shape_small_cofq(_abi.fptr(mq), 10, 1.0, C.byref(s)) == -2
    assert L.iac_calc_kmap(None, 0, None, None, None, None) == -1
    assert L.iac_ctx_sync(None) == -1
    assert L.iac_ac3d_sparse(None, None, 0, None, None, 0, 0, 1, None) == -1
    assert L.iac_ppp_lap(None, None, None, 0, 0, None, None, None, None, 0, 1.0, 0.5, 0.9, 0, None, None) == -1
    n = C.c_int(-1)
    assert L.iac_device_count(C.byref(n)) == 0 and n.value >= 0


def test_context_without_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(iac.IACError) as e:
        iac.Context(0)
    assert e.value.code == -3


def test_host_library_and_eval_driver():
    """The C++ host mirror links against the C ABI; the iac_eval program starts without a GPU."""
    import ctypes
    import subprocess
    from pathlib import Path

    pkg = Path(__file__).resolve().parents[1] / "incohautocorrelate_amd"
    ctypes.CDLL(str(pkg / "libiac_host.so"))
    r = subprocess.run([str(pkg / "iac_eval"), "--version"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "abi" in r.stdout
    r = subprocess.run([str(pkg / "iac_eval")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    # the LAP mode checks its arguments before touching a device: wrong count, no panels,
    # panels larger than the frame
    for args in (["--lap"], ["--lap", "x", "1", "10", "", "1", "0.5", "0.9", "y"],
                 ["--lap", "x", "1", "10", "0:4:4", "1", "0.5", "0.9", "y"]):
        r = subprocess.run([str(pkg / "iac_eval"), *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "--lap" in r.stderr, args
    # so does the per-event queue mode: wrong count, a zero size, no frames, no mesh edge
    for args in (["--queue"], ["--queue", "pm", "0", "4", "p", "c", "o", "1", "31", "m"],
                 ["--queue", "pm", "4", "4", "p", "c", "o", "0", "31", "m"],
                 ["--queue", "pm", "4", "4", "p", "c", "o", "1", "0", "m"]):
        r = subprocess.run([str(pkg / "iac_eval"), *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "--queue" in r.stderr, args
