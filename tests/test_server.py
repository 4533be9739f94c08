"""Serving loop (no reference counterpart): continuous batching of simulation requests under
an HBM budget. Pinned by equality and invariants: every request's result equals its own
standalone run bit for bit (alone vs in a mix, chunked into steps); the admitted device
footprint never exceeds the budget; a request larger than the budget is rejected."""
import numpy as np
import pytest

import mcpet
from conftest import MATERIALS
from mcpet import phantoms

pytestmark = pytest.mark.gpu
KEYS = ("trues", "scatter", "image_trues", "image_scatter", "spectrum", "material_edep")


def _inputs(count, seed):
    specs = phantoms.random_phantoms(count, seed=seed, xy=(16, 48), z=(16, 32), histories=(4e4, 4e5))
    names = phantoms.head_materials(MATERIALS)
    cfg = phantoms.make_config(seed=4711, threads_per_block=16, acquisition_time_s=10.0, materials=names,
                               material_dose=True, nrows=20, nzs=39, mrd=19, span=3)
    phs = [mcpet.phantom_from_arrays(*sp) for sp in specs]
    tab = mcpet.load_materials(names, np.max([p.density_max() for p in phs], axis=0))
    return cfg, phs, tab


def test_requests_in_a_mix_equal_standalone_runs_within_budget():
    cfg, phs, tab = _inputs(6, seed=11)
    per_request = 2 * 4 * cfg.nbins + 8 * 1024 * 1024      # about one request's footprint
    budget = int(2.6 * per_request) + 8 * 1024 * 1024
    with mcpet.Server(tab, budget, max_running=8, streams_per_step=3000) as srv:
        ids = [srv.submit(cfg, ph, seed=100 + k) for k, ph in enumerate(phs)]
        srv.drain()
        st = srv.stats()
        assert [srv.poll(i) for i in ids] == ["done"] * len(ids)
        got = [srv.fetch(i, psf=False) for i in ids]
    assert st["completed"] == len(ids) and st["rejected"] == 0
    assert st["peak_admitted_bytes"] <= budget and st["peak_running"] >= 2
    assert st["steps"] > len(ids)                    # requests were split over several steps
    assert 0 < st["latency_p50_ms"] <= st["latency_p99_ms"] <= st["latency_max_ms"]
    for k, ph in enumerate(phs):
        with mcpet.Simulation(cfg, ph, tab) as sim:
            sim.run(seed=100 + k)
            ref = sim.fetch(psf=False)
        for key in KEYS:
            assert np.array_equal(got[k][key], ref[key]), (k, key)
        assert got[k]["total_histories"] == ref["total_histories"] and got[k]["final_seed"] == ref["final_seed"]


def test_oversized_request_is_rejected():
    cfg, phs, tab = _inputs(1, seed=12)
    with mcpet.Server(tab, 1024 * 1024, max_running=2, streams_per_step=1000) as srv:
        with pytest.raises(mcpet.McpetError):
            srv.submit(cfg, phs[0])
        assert srv.stats()["rejected"] == 1


def test_bad_requests_are_refused_at_submit_and_the_queue_proceeds():
    """Requests that could not be created (a dose region of interest outside the phantom) or
    run (an unusable seed) are rejected when submitted, with the library's reason; fetching one
    reports it and forgets it; the valid requests behind them complete and equal their
    standalone runs."""
    cfg, phs, tab = _inputs(2, seed=13)
    nz, ny, nx = phs[0].shape
    bad = phantoms.make_config(seed=4711, threads_per_block=16, acquisition_time_s=10.0,
                               materials=phantoms.head_materials(MATERIALS), material_dose=True, nrows=20, nzs=39,
                               mrd=19, span=3, voxel_dose=True, roi=(0, nx, 0, ny - 1, 0, nz - 1))   # x one past the end
    with mcpet.Server(tab, 4 << 30, max_running=4, streams_per_step=2000) as srv:
        with pytest.raises(mcpet.McpetError, match="region of interest"):
            srv.submit(bad, phs[0])
        rid_bad = max(srv._meta)
        assert srv.poll(rid_bad) == "rejected"
        with pytest.raises(mcpet.McpetError, match="seed"):
            srv.submit(cfg, phs[1], seed=-2147483400)            # below -m2: abMODm would divide by zero
        good = [srv.submit(cfg, ph, seed=200 + k) for k, ph in enumerate(phs)]
        srv.drain()
        assert [srv.poll(i) for i in good] == ["done", "done"]
        with pytest.raises(mcpet.McpetError, match="rejected"):
            srv.fetch(rid_bad, psf=False)
        assert srv.poll(rid_bad) == "unknown"
        st = srv.stats()
        assert st["rejected"] == 2 and st["completed"] == 2 and st["failed"] == 0
        got = [srv.fetch(i, psf=False) for i in good]
    for k, ph in enumerate(phs):
        with mcpet.Simulation(cfg, ph, tab) as sim:
            sim.run(seed=200 + k)
            ref = sim.fetch(psf=False)
        for key in KEYS:
            assert np.array_equal(got[k][key], ref[key]), (k, key)
