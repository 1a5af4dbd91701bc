"""The host mirror frees its device reservation on every exit path (CPU only).

A small C++ program against libiac_host.so hands Settings a one-device pool of
placeholder contexts (Settings::UseContexts; no device call is made), makes each
reserving method throw after its reservation -- a mesh on another device, a small
C(q) mesh on another device -- and checks that the pool's device can be reserved
again afterwards (OCL_ReserveDevice / OCL_FreeDevice, Settings.cpp:432-470)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "incohautocorrelate_amd"

PROGRAM = r"""
#include <cstdio>
#include <stdexcept>
#include "iac_host.hpp"

static iac_ctx *fake(uintptr_t v) { return reinterpret_cast<iac_ctx *>(v); }

int main()
{
    Settings S;
    S.echo = false;
    S.UseContexts({fake(0x1000)});
    Detector D(S);
    Detector::AutoCorrFlags flags;
    int failures = 0;
    {   // AutoCorrelateSparseList: the mesh lives on another device
        ACMesh M(&S);
        M.Checklist.BigMesh = true;
        M.ctx = fake(0x2000);
        D.Checklist.SparseHitList = true;
        bool thrown = false;
        try { D.AutoCorrelateSparseList(M, flags, flags, true, S, 1); } catch (const std::runtime_error &) { thrown = true; }
        const int dev = S.GPU_ReserveDevice();
        if (!thrown || dev != 0) { std::printf("AutoCorrelateSparseList: thrown %d, next reservation %d\n", thrown, dev); failures++; }
        if (dev >= 0) S.GPU_FreeDevice(dev);
    }
    {   // Merge_smallCofQ: the small C(q) mesh lives on another device
        ACMesh Big(&S), Small(&S);
        Big.Checklist.CofQBigMesh = true;
        Small.Checklist.CofQSmallMesh = true;
        Small.d_CQMesh = reinterpret_cast<double *>(0x3000);
        Small.ctx = fake(0x2000);
        std::vector<Settings::HitEvent> ev(2);
        bool thrown = false;
        try { D.Merge_smallCofQ(Big, Small, ev, 0, 2, S, flags); } catch (const std::runtime_error &) { thrown = true; }
        Small.d_CQMesh = nullptr;
        const int dev = S.GPU_ReserveDevice();
        if (!thrown || dev != 0) { std::printf("Merge_smallCofQ: thrown %d, next reservation %d\n", thrown, dev); failures++; }
        if (dev >= 0) S.GPU_FreeDevice(dev);
    }
    {   // an exhausted pool: the guard throws and leaves the held device alone
        const int held = S.GPU_ReserveDevice();
        ACMesh M(&S);
        M.Checklist.BigMesh = true;
        bool thrown = false;
        try { D.AutoCorrelateSparseList(M, flags, flags, true, S, 1); } catch (const std::runtime_error &) { thrown = true; }
        S.GPU_FreeDevice(held);
        const int dev = S.GPU_ReserveDevice();
        if (!thrown || held != 0 || dev != 0) { std::printf("exhausted pool: thrown %d, held %d, next %d\n", thrown, held, dev); failures++; }
        if (dev >= 0) S.GPU_FreeDevice(dev);
    }
    std::printf("failures %d\n", failures);
    return failures ? 1 : 0;
}
"""


def test_reservation_released_after_throw(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    if not (PKG / "libiac_host.so").exists():
        pytest.skip("host library not built")
    src = tmp_path / "guard.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "guard"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                    "-L", str(PKG), "-liac_host", "-liac_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,/opt/rocm/lib"],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
