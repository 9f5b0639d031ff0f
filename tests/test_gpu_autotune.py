"""The autotuner (morfo70_amd/autotune.py) on a live handle and a live batch: it times
every candidate schedule, applies the fastest, caches a single handle's choice per
(device, nx bucket, ny bucket), and a tuned schedule changes no result beyond the
rounding its reduced-system tile implies (row block size and graph replay change no bit;
16-unknown column segments round differently from 8-unknown ones)."""
from __future__ import annotations

import json

import numpy as np
import pytest

from morfo70_amd import _lib
from morfo70_amd import scenarios as S
from tests.helpers import nonuniform_beach

pytestmark = pytest.mark.gpu


def _states_agree(a, b, bitwise):
    for k in a:
        if bitwise:
            assert np.array_equal(a[k], b[k]), k
        else:
            scale = max(float(np.max(np.abs(b[k]))), 1e-300)
            assert float(np.max(np.abs(a[k] - b[k]))) <= 1e-12 * scale, k


def test_tune_handle_applies_caches_and_keeps_results(tmp_path, monkeypatch):
    from morfo70_amd import autotune
    from morfo70_amd.hydro import from_scenario
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "cache.json"))
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=64, ny=96, physics=ph, seed=41)
    print("[autotune] handle 64x96", flush=True)
    ref = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    ref.Run(6)
    st_ref = ref.state()
    ref.close()

    g = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    g.push()
    g.set_forcing()   # the trials run the handle's own device state
    cfg = autotune.tune_handle(g, steps=2)
    g.close()
    assert cfg["rows_block"] in (256, 512) and cfg["tile_cols"] in (0, 8, 16), cfg
    assert cfg["ms_per_step"] > 0
    with open(tmp_path / "cache.json") as f:
        cache = json.load(f)
    key = autotune._key(64, 96)
    assert cache[key] == cfg

    def no_timing(*args, **kwargs):
        raise AssertionError("a cached size was timed again")

    monkeypatch.setattr(autotune, "_time", no_timing)
    h = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    assert autotune.tune_handle(h) == cfg   # same buckets: taken from the cache
    h.Run(6)
    _states_agree(h.state(), st_ref, bitwise=cfg["tile_cols"] != 16)
    h.close()


def test_tune_batch_picks_a_candidate_and_keeps_results_bitwise():
    from morfo70_amd import autotune
    from morfo70_amd.batch import Batch
    from morfo70_amd.hydro import from_scenario
    scs = [S.seeded_beach(nx, ny, 2.0, 2.0, 90 + q, 20, random_waves=True)
           for q, (nx, ny) in enumerate([(130, 260), (96, 512), (200, 129)])]
    print("[autotune] batch of 3", flush=True)
    hs = [from_scenario(sc, solver=_lib.SOLVER_FUSED) for sc in scs]
    for h in hs:
        h.push()
        h.set_forcing()
    b = Batch(hs)
    cfg = autotune.tune_batch(b, steps=2)
    b.close()
    for h in hs:
        h.close()
    tried = cfg["tried"]
    assert {(t["rows_block"], t["graph"]) for t in tried} == {(256, True), (256, False), (512, True), (512, False)}
    assert cfg["ms_per_step"] == min(t["ms_per_step"] for t in tried)

    ref = []
    for sc in scs:
        g = from_scenario(sc, solver=_lib.SOLVER_FUSED)
        g.Run(6)
        ref.append(g.state())
        g.close()
    hs = [from_scenario(sc, solver=_lib.SOLVER_FUSED) for sc in scs]
    for h in hs:
        h.push()
        h.set_forcing()
    b = Batch(hs, graph=cfg["graph"])
    b.set_rows_block(cfg["rows_block"])
    b.run(6)
    for q, h in enumerate(hs):
        _states_agree(h.state(), ref[q], bitwise=True)
    b.close()
    for h in hs:
        h.close()
