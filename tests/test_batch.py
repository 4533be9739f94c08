"""Batched persistent kernel (no reference counterpart; the reference runs one phantom per
process): n phantoms in one launch with a stream queue per phantom. Pinned by equality: each
phantom's tallies and event counts equal those of its own plain-kernel launch (which equals
the oracle, see test_gpu_parity), bit for bit, whatever the launch shape, and a captured
graph replays to the eager result."""
import ctypes as C

import numpy as np
import pytest

import mcpet
from conftest import MATERIALS
from mcpet import phantoms

KEYS = ("trues", "scatter", "image_trues", "image_scatter", "spectrum", "material_edep")


def _batch_inputs(count=3, seed=5):
    specs = phantoms.random_phantoms(count, seed=seed, xy=(16, 40), z=(16, 32), histories=(4e4, 4e5))
    names = phantoms.head_materials(MATERIALS)
    cfg = phantoms.make_config(seed=31337, threads_per_block=16, acquisition_time_s=10.0, materials=names,
                               material_dose=True)
    phs = [mcpet.phantom_from_arrays(*sp) for sp in specs]
    dmax = np.max([p.density_max() for p in phs], axis=0)      # tables shared by the batch
    tab = mcpet.load_materials(names, dmax)
    return cfg, phs, tab


def _same(a, b, counts=True):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    assert a["total_histories"] == b["total_histories"] and a["final_seed"] == b["final_seed"]


def test_batch_rejects_bad_arguments_without_a_device():
    cfg = mcpet.Config()
    b = C.c_void_p()
    assert mcpet.lib().mcpet_batch_create(C.byref(cfg), None, 0, None, C.byref(b)) == 1 and not b.value


@pytest.mark.gpu
def test_batch_equals_plain_launch_per_phantom():
    cfg, phs, tab = _batch_inputs()
    with mcpet.Batch(cfg, phs, tab) as b:
        b.run(count_events=True)
        got = [(b.fetch(k, psf=False), b.stats(k)) for k in range(len(phs))]
    for k, ph in enumerate(phs):
        with mcpet.Simulation(cfg, ph, tab) as sim:
            sim.run(kernel=mcpet.KERNEL_PLAIN)
            ref = sim.fetch(psf=False)
            rs = sim.stats()
        g, gs = got[k]
        _same(g, ref)
        for key in ("decays", "streams", "woodcock_steps", "material_changes", "compton", "compton_shell_evals",
                    "rayleigh", "photoelectric", "pairs_tallied", "coincidences", "rn_draws"):
            assert gs[key] == rs[key], (k, key, gs[key], rs[key])
        assert gs["coincidences"] > 0


@pytest.mark.gpu
def test_batch_graph_replays_match_eager_runs():
    cfg, phs, tab = _batch_inputs(count=3, seed=6)
    with mcpet.Batch(cfg, phs, tab) as b:
        b.run(seed=777)
        eager = [b.fetch(k, psf=False) for k in range(len(phs))]
        for rep in range(3):
            b.run(seed=777, use_graph=True)
            b.sync()
            for k in range(len(phs)):
                _same(b.fetch(k, psf=False), eager[k])
            b.run(seed=778 + rep)        # eager work between the replays
            b.sync()


@pytest.mark.gpu
def test_batch_autotune_keeps_results():
    cfg, phs, tab = _batch_inputs(count=4, seed=7)
    with mcpet.Batch(cfg, phs, tab) as b:
        b.run(seed=99)
        before = [b.fetch(k, psf=False) for k in range(len(phs))]
        res = b.autotune(0.1)
        assert len(res) >= 4 and all(r["decays_per_ms"] > 0 for r in res)
        assert {r["phase_policy"] for r in res} == {1, 2}      # both phase-selection policies tried
        b.run(seed=99)
        for k in range(len(phs)):
            _same(b.fetch(k, psf=False), before[k])


@pytest.mark.gpu
def test_batch_with_balanced_streams_equals_each_phantoms_own_run():
    """The balanced stream layout (stream_decays > 0) in a batch: per-phantom stream ranges
    and random-number ids as in each phantom's own persistent run, bit for bit."""
    cfg, phs, tab = _batch_inputs(seed=9)
    cfg.stream_decays = 7
    with mcpet.Batch(cfg, phs, tab) as b:
        b.run(seed=4242, use_graph=True)
        got = [b.fetch(k, psf=False) for k in range(len(phs))]
        st = [b.stats(k) for k in range(len(phs))]
    for k, ph in enumerate(phs):
        with mcpet.Simulation(cfg, ph, tab) as sim:
            sim.run(seed=4242, kernel=mcpet.KERNEL_PERSISTENT)
            ref = sim.fetch(psf=False)
            rs = sim.stats()
        _same(got[k], ref)
        assert st[k]["streams"] == rs["streams"] and st[k]["decays"] == rs["decays"] > 0


@pytest.mark.gpu
def test_48_phantom_workload_graph_replay_equals_each_persistent_run():
    """The many-phantom workload at its stated size: 48 random phantoms (32-160 x 32-160 x
    32-96 voxels at 2 mm, 1e6-1e8 photon histories each in proportion to their activity,
    10 s), a 168 x 168 x 1293 sinogram per phantom, all in one batched persistent launch
    captured and replayed as a HIP graph. After the replay, every phantom's tallies, final
    seed and counters equal those of its own persistent launch, bit for bit."""
    import sys
    specs = phantoms.random_phantoms(48)
    names = phantoms.head_materials(MATERIALS)
    cfg = phantoms.make_config(seed=20261202, threads_per_block=32, acquisition_time_s=10.0, nrad=168,
                               materials=names)
    phs = [mcpet.phantom_from_arrays(*sp) for sp in specs]
    tab = mcpet.load_materials(names, np.max([p.density_max() for p in phs], axis=0))
    print(f"48-phantom batch: {sum(p.nvox for p in phs)} voxels, graph capture and replay", flush=True)
    with mcpet.Batch(cfg, phs, tab) as b:
        b.run(seed=11)                           # eager, then the graph (captured) and its replay
        b.sync()
        b.run(seed=4321, use_graph=True)
        b.sync()
        b.run(seed=4321, use_graph=True)
        b.sync()
        total = 0
        for k, ph in enumerate(phs):
            got, gs = b.fetch(k, psf=False), b.stats(k)
            with mcpet.Simulation(cfg, ph, tab) as sim:
                sim.run(seed=4321, kernel=mcpet.KERNEL_PERSISTENT)
                ref, rs = sim.fetch(psf=False), sim.stats()
            _same(got, ref)
            for key in ("decays", "streams", "pairs_tallied", "coincidences"):
                assert gs[key] == rs[key], (k, key, gs[key], rs[key])
            total += gs["decays"]
            del got, ref
        print(f"48-phantom batch: {total} decays, every phantom equals its own run", flush=True)
        sys.stdout.flush()
    assert 1e8 < total < 2e9
    for p in phs:
        p.close()
    tab.close()
