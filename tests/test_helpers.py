"""The drop-in's app-side helpers (include/kernel/helper/*.h: level-set shape generators, sphere
packings, OBJ reader) on the host, and the reference's own tutorial apps compiling unchanged
against include/ (the tutorials' CUDA damping helper replaced by its HIP counterpart
include/globalDamping.h, which a one-line globalDamping.cuh includes)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = "/opt/rocm/bin/hipcc"
REF_APPS = Path("/root/reference/FunDEM/apps")


def test_helper_checks(tmp_path):
    exe = tmp_path / "check_helpers"
    r = subprocess.run([HIPCC, "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests/cpp/check_helpers.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(tmp_path / "quad.obj")], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout


APPS = sorted(REF_APPS.glob("tutorial/tutorial*.cpp")) + sorted(REF_APPS.glob("test/*.cpp"))


@pytest.mark.skipif(not APPS, reason="the reference apps are not on this machine")
@pytest.mark.parametrize("app", APPS, ids=[a.name for a in APPS])
def test_reference_app_compiles_unchanged(tmp_path, app):
    src = tmp_path / app.name
    shutil.copyfile(app, src)
    (tmp_path / "globalDamping.cuh").write_text('#include "globalDamping.h"\n')
    r = subprocess.run([HIPCC, "-std=c++17", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
