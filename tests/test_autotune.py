"""The context autotuner (no reference counterpart: the reference fixes its one launch from
threads_per_block, MCGPU-PET.cu:452-487). Pinned by equality:
- every LDS layout (palette, spectrum and 511 keV cross-section row each in LDS or read from
  global memory) gives the same tallies bit for bit, and so does every tuned run;
- a decays-per-stream value the tuner chooses reproduces, bit for bit, a context created with
  that value (and the oracle states the same balanced layout, test_balanced_streams)."""
import ctypes as C
import sys

import numpy as np
import pytest

import mcpet
from harness import head_scenario, water_cube_scenario

KEYS = ("trues", "scatter", "image_trues", "image_scatter", "spectrum", "material_edep")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b, what):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (what, k)
    assert a["total_histories"] == b["total_histories"] and a["final_seed"] == b["final_seed"], what


def test_autotune_rejects_bad_options_without_a_device():
    o = mcpet.TuneOptions()
    o.stream_fraction = 0.0
    n = C.c_int32()
    assert mcpet.lib().mcpet_autotune(None, C.byref(o), None, 0, C.byref(n)) == 1
    assert C.sizeof(mcpet.TuneResult) == 48 and C.sizeof(mcpet.TuneOptions) == 88


@pytest.mark.gpu
def test_every_lds_layout_gives_the_same_tallies_and_tuning_keeps_them():
    _say("head 24x32x32, 2e5 decays: 8 LDS layouts, then autotune")
    sc = head_scenario(n_decays=2e5, tpb=16, material_dose=True)
    with mcpet.Simulation(sc.cfg, sc.phantom, sc.tables) as sim:
        sim.run(seed=99)
        ref = sim.fetch(psf=False)
        assert ref["trues"].sum() > 0
        for lds in range(8):
            sim.run(seed=99, lds_layout=lds)
            _same(sim.fetch(psf=False), ref, f"layout {lds}")
        res = sim.autotune(fraction=0.1)
        assert len(res) >= 7
        choice = res[-1]
        assert choice["block_size"] in (128, 256) and choice["phase_policy"] in (1, 2)
        assert 0 <= choice["lds_layout"] < 8 and choice["stream_decays"] == 0
        assert all(r["decays_per_ms"] > 0 for r in res)
        assert choice["decays_per_ms"] == max(r["decays_per_ms"] for r in res[:-1])
        sim.run(seed=99)                        # the tuned launch, layout and policy
        _same(sim.fetch(psf=False), ref, "tuned")


@pytest.mark.gpu
def test_tuned_decays_per_stream_reproduce_a_context_created_with_them():
    _say("water cube, 4e5 decays: stream-layout candidates 0, 7, 50")
    sc = water_cube_scenario(n_decays=4e5, tpb=8, material_dose=True)
    with mcpet.Simulation(sc.cfg, sc.phantom, sc.tables) as sim:
        res = sim.autotune(fraction=0.25, stream_decays=[0, 7, 50])
        choice = res[-1]
        assert sorted({r["stream_decays"] for r in res[:-1]}) == [0, 7, 50]
        sim.run(seed=4242)
        tuned = sim.fetch(psf=False)
    cfg = sc.cfg
    cfg.stream_decays = int(choice["stream_decays"])
    with mcpet.Simulation(cfg, sc.phantom, sc.tables) as sim:
        sim.run(seed=4242)
        _same(tuned, sim.fetch(psf=False), f"stream_decays {choice['stream_decays']}")
    # and a forced balanced layout is a different, statistically equal simulation
    cfg.stream_decays = 7 if choice["stream_decays"] != 7 else 50
    with mcpet.Simulation(cfg, sc.phantom, sc.tables) as sim:
        sim.run(seed=4242)
        other = sim.fetch(psf=False)
    a, b = int(tuned["trues"].sum()), int(other["trues"].sum())
    assert abs(a - b) < 6 * np.sqrt(a + b) + 10
