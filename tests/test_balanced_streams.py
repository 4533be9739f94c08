"""Balanced decay-stream layout (the stream_decays extension; the reference's layout is
threads_per_block streams per emitting voxel, MCGPU-PET_kernel.cu:223-260): hot voxels get
proportionally more streams so every stream carries about stream_decays decays.

CPU tier: the oracle's stream count against its definition computed here from the phantom,
and the balanced layout against the reference layout statistically (decays, trues, scatter,
per-voxel emissions). GPU tier: the device against the oracle bit for bit in this layout, for
the fused tallies and for the singles pipeline."""
import numpy as np
import pytest

import mcpet
from conftest import AIR, WATER
from harness import Scenario, assert_tallies_equal
from mcpet import phantoms


def _scenario(stream_decays, n_decays=2e5, seed=777, tpb=8, acq_s=1.0, **kw):
    mat, dens, act, vs = phantoms.water_cube(n_decays=n_decays, acquisition_time_s=acq_s)
    cfg = phantoms.make_config(seed=seed, threads_per_block=tpb, materials=[AIR, WATER], stream_decays=stream_decays,
                               acquisition_time_s=acq_s, **kw)
    return Scenario(cfg, mat, dens, act, vs, [AIR, WATER]), act


def _expected_streams(act, cfg):
    tau = float(np.float32(cfg.mean_life))
    T = cfg.acquisition_time_ps * 1e-12
    e = act.astype(np.float64).ravel() * (tau * (1 - np.exp(-T / tau)))
    e = e[act.ravel() > np.float32(1e-7)]
    return int(np.maximum(1, np.ceil(e / cfg.stream_decays)).sum())


def test_oracle_stream_count_follows_the_definition():
    for sd in (1, 7, 300):
        sc, act = _scenario(sd, n_decays=2e4)
        o = sc.run_oracle("fmath", nthreads=8)
        assert o["counters"]["streams"] == _expected_streams(act, sc.cfg), sd


def test_balanced_layout_is_statistically_the_reference_layout():
    a, _ = _scenario(0, n_decays=3e5, seed=11)
    b, _ = _scenario(20, n_decays=3e5, seed=12)
    oa = a.run_oracle("fmath", nthreads=8)
    ob = b.run_oracle("fmath", nthreads=8)
    da, db = oa["counters"]["decays"], ob["counters"]["decays"]
    assert abs(da - db) < 5 * np.sqrt(da + db)
    for k in ("trues", "scatter"):
        x, y = int(oa[k].sum()), int(ob[k].sum())
        assert abs(x - y) < 5 * np.sqrt(x + y) + 5, (k, x, y)
    # per emitting voxel: the trues imaged at the emission voxel (hot voxels included)
    ia, ib = oa["image_trues"].astype(float), ob["image_trues"].astype(float)
    hot = np.argsort(ia)[-3:]
    for v in hot:
        assert abs(ia[v] - ib[v]) < 5 * np.sqrt(ia[v] + ib[v]) + 5


@pytest.mark.gpu
@pytest.mark.parametrize("sd", [5, 64])
def test_device_balanced_layout_bit_exact(sd):
    sc, _ = _scenario(sd, n_decays=1e5, material_dose=True)
    o = sc.run_oracle("fmath")
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)
    assert_tallies_equal(g, o)
    gp = sc.run_gpu(kernel=mcpet.KERNEL_PLAIN)
    assert_tallies_equal(gp, o)


@pytest.mark.gpu
def test_device_balanced_layout_singles_bit_exact():
    import oracle_py
    sc, _ = _scenario(9, n_decays=5e4, tally_psf_sinogram=3, singles_capacity=400_000, coincidence_window_ps=4500,
                      acq_s=1e-3)
    o = sc.run_oracle("fmath", nthreads=8)
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)
    key = ["time_ps", "event", "kind", "scatter", "energy", "z", "phi"]
    assert np.array_equal(np.sort(g["singles"], order=key).view(np.uint8), np.sort(o["singles"], order=key).view(np.uint8))
    assert g["final_seed"] == o["final_seed"]
    geom = {"ncrystals": sc.cfg.ncrystals, "nrows": sc.cfg.nrows, "fovz": sc.cfg.fovz, "nbins": sc.cfg.nbins}
    sinos, counts, raw = oracle_py.pair_singles(o["singles"], 4500, geom, sc.oracle("fmath"))
    for i, k in enumerate(("trues", "scatter", "randoms", "spurious")):
        assert np.array_equal(g[k], sinos[i]), k
