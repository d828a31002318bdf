"""The LSSolver drop-in (include/kernel/LSSolver.h) driving the GPU library from an app written as
a FunDEM app is (lsdem2.0_amd/apps/ls_sphere_bouncing.cpp; run on an MI355X).

A sphere of restitution 1 falls 1 m onto a plane: it must come back up to its drop height, and the
frame files must follow the reference's layout (<dir>_phase1/<series>/<stem>_####.vtu with base64
inline arrays and a UInt64 header)."""
import base64
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
APP = ROOT / "lsdem2.0_amd" / "apps" / "ls_sphere_bouncing"


def read_vtu(path: Path) -> dict:
    txt = path.read_text()
    out = {"TIME": float(re.search(r'Name="TIME"[^>]*>([^<]+)<', txt).group(1))}
    for m in re.finditer(r'<DataArray type="(\w+)"( Name="(\w+)")?( NumberOfComponents="(\d+)")? format="binary">\n'
                         r'([A-Za-z0-9+/=]*)\n</DataArray>', txt):
        raw = base64.b64decode(m.group(6))
        n = int(np.frombuffer(raw[:8], np.uint64)[0])
        dt = {"Float32": np.float32, "Int32": np.int32, "UInt8": np.uint8}[m.group(1)]
        a = np.frombuffer(raw[8:8 + n], dt)
        k = int(m.group(5) or 1)
        out[m.group(3) or "Points"] = a.reshape(-1, k) if k > 1 else a
    return out


def test_sphere_bounces_back(tmp_path):
    assert APP.exists(), "build the apps first (__graft_entry__.build())"
    prefix = tmp_path / "bounce"
    r = subprocess.run([str(APP), str(prefix)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    pdir = Path(str(prefix) + "_phase1")
    frames = sorted((pdir / "LSParticle").glob("Particle_*.vtu"))
    assert len(frames) == 202, r.stdout[-2000:] + r.stderr[-2000:]  # frame 0, 200 frames, the final one
    for sub, stem in (("LSParticle", "LSObjectMesh"), ("Wall", "LSObjectMesh"),
                      ("LSParticleInteraction", "LSInteraction"), ("LSParticle-WallInteraction", "LSInteraction")):
        assert (pdir / sub / f"{stem}_0000.vtu").exists(), (sub, stem)
        assert (pdir / sub / f"{stem}.pvd").exists()
    t, z = [], []
    for f in frames:
        d = read_vtu(f)
        t.append(d["TIME"])
        z.append(float(d["Points"][0, 2]))
    t, z = np.array(t), np.array(z)
    assert z[0] == pytest.approx(1.0)
    # free fall to the plane takes sqrt(2 h / g) = 0.4515 s; with e = 1 it rises back to ~1 m
    after = t > 0.6
    assert z[after].max() == pytest.approx(1.0, abs=0.03)
    assert z.min() < 0.02
    mesh = read_vtu(pdir / "LSParticle" / "LSObjectMesh_0000.vtu")
    assert mesh["Points"].shape == (36, 3) and len(mesh["types"]) == 60


CANTILEVER = ROOT / "lsdem2.0_amd" / "apps" / "bonded_cantilever"


def test_bonded_cantilever_tip_deflection(tmp_path):
    """The drop-in's helper headers (LevelSetObject::Sphere), bonds, the tutorial damping helpers and
    the interaction getters, end to end: 11 bonded spheres clamped at one end under a small tip load
    settle to Euler-Bernoulli's tip deflection F L^3 / (3 E I) (the parallel bonds' shear and
    bending stiffness B2, B3 make a chain of exact beam elements; shear deformation of this slender
    beam adds about 0.5 %)."""
    assert CANTILEVER.exists(), "build the apps first (__graft_entry__.build())"
    r = subprocess.run([str(CANTILEVER), str(tmp_path / "beam")], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    m = re.search(r"tip deflection (\S+) m, Euler-Bernoulli (\S+) m, ratio (\S+)", r.stdout)
    assert m, r.stdout[-2000:]
    assert 0.97 <= float(m.group(3)) <= 1.03
    b = re.search(r"bonds (\d+) active (\d+), particle contacts (\d+)", r.stdout)
    assert b and int(b.group(1)) == 10 and int(b.group(2)) == 10
