"""The committed macroscopic fixture (tests/golden/c1_macro.json) is the oracle's own output: its
start rebuilds identically and the oracle reproduces the first recorded checkpoint exactly."""
import importlib.util
import json
from pathlib import Path

from oracle_binding import oracle_sim

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_c1_fixture_first_checkpoint():
    spec = importlib.util.spec_from_file_location("make_c1_macro", GOLDEN / "make_c1_macro.py")
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    ref = json.loads((GOLDEN / "c1_macro.json").read_text())
    sc = mk.c1_start()
    sim = sc.load(oracle_sim(sc.cfg, threads=4))
    assert mk.record(sim, sc, 0) == ref["rows"][0]
    sim.step(ref["every"])
    assert mk.record(sim, sc, ref["every"]) == ref["rows"][1]
