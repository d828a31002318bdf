"""The C ABI library loads without a GPU and exports every entry point declared in include/lsdem.h."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared(header: Path):
    txt = header.read_text()
    return sorted(set(re.findall(r"\b(lsdem_[a-z0-9_]+)\s*\(", txt)))


def extern_c(header: Path):
    txt = header.read_text()
    return sorted(set(re.findall(r'extern "C"[^;{]*?\b(\w+)\s*\(', txt)))


def test_all_symbols_exported(L):
    lib = L.load_library()
    names = declared(ROOT / "include" / "lsdem.h")
    assert len(names) >= 15
    helpers = extern_c(ROOT / "include" / "globalDamping.h")
    assert helpers == ["launchAddGlobalConstantForceTorque", "launchAddGlobalDampingForceTorque"]
    missing = [n for n in names + helpers if not hasattr(lib, n)]
    assert not missing, missing


def test_config_default_and_errors(L):
    lib = L.load_library()
    cfg = L.Config()
    lib.lsdem_config_default(C.byref(cfg))
    assert cfg.quirks == L.QUIRK_ALL and cfg.dt > 0 and cfg.block_size == 256
    ctx = C.c_void_p()
    bad = L.Config()
    lib.lsdem_config_default(C.byref(bad))
    bad.domain_max[0] = bad.domain_min[0]  # zero-size domain is rejected before any device call
    assert lib.lsdem_create(C.byref(bad), C.byref(ctx)) == 1
    lib.lsdem_last_error.restype = C.c_char_p
    assert b"domain" in lib.lsdem_last_error()


def test_mass_properties_matches_oracle(L):
    """Product and oracle compute mass properties with the same loop order: identical results."""
    import numpy as np
    from oracle_binding import oracle_lib
    r = 4e-3
    h = 0.4e-3
    n = 30
    o = -0.5 * (n - 1) * h
    ax = o + h * np.arange(n)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    sdf = (np.sqrt((X / 1.2) ** 2 + Y ** 2 + (Z / 0.8) ** 2) - r).astype(np.float64)
    a = L.mass_properties(L.load_library(), sdf, (o, o, o), h, 2500.0)
    b = L.mass_properties(oracle_lib(), sdf, (o, o, o), h, 2500.0, prefix="lsdem_oracle_")
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def test_struct_layouts_match_header(L, tmp_path):
    """The ctypes mirrors of the public structs have the C header's size and field offsets
    (a small C program built with gcc against include/lsdem.h prints them)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    mirrors = {"lsdem_config": L.Config, "lsdem_counters": L.Counters, "lsdem_profile": L.Profile,
               "lsdem_surrogate_params": L.SurrogateParams, "lsdem_surrogate_report": L.SurrogateReport,
               "lsdem_device_view": L.DeviceView}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lsdem.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, f"{cname}.{fname}"
