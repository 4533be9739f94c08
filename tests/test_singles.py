"""Singles pipeline (an extension without reference counterpart: the reference tallies pairs
only, MCGPU-PET_kernel.cu:1623-1836): decay branching and prompt gammas, the singles list,
the coincidence sorter and its randoms.

CPU tier: the pairing statement against closed forms for Poisson singles (the randoms of a
cluster-of-two sorter, and 2*tau*S1*S2 between two detectors); the oracle's singles path
against the oracle's reference (fused) path at low count rate; branching ratios (binomial).
GPU tier: the device radix sort against a stable CPU sort; the singles list, spectrum, doses,
counters and coincidence sinograms bit-exact against the oracle and the CPU pairing."""
import numpy as np
import pytest

import mcpet
import oracle_py
from conftest import AIR, WATER
from harness import Scenario, assert_tallies_equal, water_cube_scenario
from mcpet import phantoms

WINDOW = 4500   # ps


def _poisson_singles(rate_hz, duration_s, rng, kinds=None):
    n = rng.poisson(rate_hz * duration_s)
    t = np.sort(rng.uniform(0, duration_s * 1e12, n)).round().astype(np.uint64)
    s = np.zeros(n, oracle_py.SINGLE_DTYPE)
    s["time_ps"] = t
    s["event"] = np.arange(n, dtype=np.uint64) + 1     # every single its own decay
    if kinds is not None:
        s["phi"] = kinds
    return s


def test_pairing_randoms_closed_form():
    """Uncorrelated Poisson singles at rate S: a single opens a two-single cluster with
    probability (1-p) p (1-p), p = 1 - exp(-S W), every such pair is random."""
    rng = np.random.default_rng(1)
    S, T = 2.0e7, 0.05
    s = _poisson_singles(S, T, rng)
    _, _, raw = oracle_py.pair_singles(s, WINDOW, {"nbins": 1})
    p = 1.0 - np.exp(-S * (WINDOW + 0.5) * 1e-12)
    expect = len(s) * p * (1 - p) ** 2
    assert raw[0] == raw[1] == raw[3] == 0
    assert abs(raw[2] - expect) < 5 * np.sqrt(expect), (raw[2], expect)
    # multiples: clusters of >= 3 start with probability (1-p) p^2
    em = len(s) * (1 - p) * p * p
    assert abs(raw[4] - em) < 5 * np.sqrt(em) + 5, (raw[4], em)


def test_pairing_two_detector_randoms_rate():
    """Two detectors with singles rates S1, S2 (S = S1 + S2): randoms between them
    R12 = S (1-p)^2 p 2 S1 S2 / S^2, which is the textbook 2 tau S1 S2 as S tau -> 0
    (p = 1 - exp(-S tau) ~ S tau; (1-p)^2 is the chance that the pair is not in a multiple)."""
    rng = np.random.default_rng(2)
    S1, S2, T = 6.0e6, 1.5e7, 0.1
    s1 = _poisson_singles(S1, T, rng)
    s2 = _poisson_singles(S2, T, rng)
    s2["event"] += np.uint64(10 ** 9)
    s1["phi"], s2["phi"] = 1.0, -1.0
    s = np.concatenate([s1, s2])
    s = s[np.argsort(s["time_ps"], kind="stable")]
    st = np.sort(s, order=["time_ps", "event", "kind"], kind="stable")
    t = st["time_ps"].astype(np.int64)
    gap = np.diff(t)
    start = (gap <= WINDOW) & np.concatenate([[True], gap[:-1] > WINDOW]) & ~np.concatenate([gap[1:] <= WINDOW, [False]])
    idx = np.flatnonzero(start)
    cross = int(np.sum(st["phi"][idx] != st["phi"][idx + 1]))
    _, _, raw = oracle_py.pair_singles(s, WINDOW, {"nbins": 1})
    assert raw[2] == len(idx)
    S, tau = S1 + S2, (WINDOW + 0.5) * 1e-12
    p = 1.0 - np.exp(-S * tau)
    expect = S * T * (1 - p) ** 2 * p * 2 * S1 * S2 / S ** 2
    assert abs(cross - expect) < 5 * np.sqrt(expect), (cross, expect)
    textbook = 2 * tau * S1 * S2 * T
    assert abs(expect / textbook - (1 - p) ** 2 * p / (S * tau)) < 1e-12


def _singles_scenario(n_decays=1e5, tpb=8, seed=4242, acq_s=1.0, **kw):
    mat, dens, act, vs = phantoms.water_cube(n_decays=n_decays, acquisition_time_s=acq_s)
    cfg = phantoms.make_config(seed=seed, threads_per_block=tpb, materials=[AIR, WATER], acquisition_time_s=acq_s,
                               tally_psf_sinogram=3, coincidence_window_ps=WINDOW, **kw)
    return Scenario(cfg, mat, dens, act, vs, [AIR, WATER])


def _geom(cfg):
    return {"ncrystals": cfg.ncrystals, "nrows": cfg.nrows, "fovz": cfg.fovz, "nbins": cfg.nbins}


def test_singles_path_matches_fused_path_at_low_rate():
    """At a count rate where decays never overlap within the window, pairing the oracle's
    singles recovers the annihilation pairs that the reference's fused tally bins: the same
    number of trues and scattered coincidences within Poisson errors, no randoms."""
    acq = 200.0      # 1e5 decays over 200 s: mean gap between decays 2 ms >> 4.5 ns
    sc = _singles_scenario(n_decays=1e5, acq_s=acq, singles_capacity=400_000)
    o = sc.run_oracle("libm", nthreads=8)
    sinos, counts, raw = oracle_py.pair_singles(o["singles"], WINDOW, _geom(sc.cfg), sc.oracle("libm"))
    assert raw[2] + raw[3] + raw[4] <= 2
    mat, dens, act, vs = phantoms.water_cube(n_decays=1e5, acquisition_time_s=acq)
    cfg2 = phantoms.make_config(seed=999, threads_per_block=8, materials=[AIR, WATER], acquisition_time_s=acq)
    f = Scenario(cfg2, mat, dens, act, vs, [AIR, WATER]).run_oracle("libm", nthreads=8)
    for k, key in ((0, "trues"), (1, "scatter")):
        a, b = counts[k], int(f[key].sum())
        assert abs(a - b) < 5 * np.sqrt(a + b) + 5, (key, a, b)
    # every detected single inside the window is in the list; the spectrum counts all singles
    assert o["singles_requested"] == len(o["singles"])
    assert o["spectrum"].sum() >= len(o["singles"])


def test_branching_ratios_binomial():
    """Positron branching 0.4 and prompt-gamma branching 0.7 (400 keV): the emitted pairs and
    prompt gammas are binomial in the number of decays."""
    sc = _singles_scenario(n_decays=3e4, singles_capacity=200_000, positron_fraction=0.4, prompt_gamma_fraction=0.7,
                           prompt_gamma_energy=400e3)
    o = sc.run_oracle("fmath", nthreads=8)
    c = o["counters"]
    n = c["decays"]
    for key, p in (("pairs_emitted", 0.4), ("prompts_emitted", 0.7)):
        assert abs(c[key] - n * p) < 5 * np.sqrt(n * p * (1 - p)), (key, c[key], n * p)
    kinds = np.bincount(o["singles"]["kind"], minlength=3)
    assert kinds[2] > 0 and kinds[0] > 0 and kinds[1] > 0


# ------------------------------------------------------------------------------------ GPU tier

@pytest.mark.gpu
@pytest.mark.parametrize("n,bits", [(0, 64), (1, 64), (255, 64), (4096, 64), (4097, 40), (100_003, 44), (1_000_000, 64)])
def test_device_radix_sort_is_a_stable_sort(n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    keys &= np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    if n > 10:
        keys[rng.integers(0, n, n // 3)] = keys[rng.integers(0, n, n // 3)]     # many equal keys
    passes = (bits + 7) // 8
    perm = mcpet.selftest_sort(keys, passes)
    assert np.array_equal(perm, np.argsort(keys, kind="stable").astype(np.uint32))


def _check_against_oracle(sc, nthreads=8, keep=None):
    """GPU singles pipeline vs the oracle on sc (keep: a dict that receives the oracle's result)."""
    o = sc.run_oracle("fmath", nthreads=nthreads)
    if keep is not None:
        keep["oracle"] = o
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)
    # singles list: the same records (the device list comes sorted by time)
    gs, os_ = g["singles"], o["singles"]
    assert g["singles_requested"] == o["singles_requested"] and len(gs) == len(os_)
    assert np.all(np.diff(gs["time_ps"].astype(np.int64)) >= 0)
    key = ["time_ps", "event", "kind", "scatter", "energy", "z", "phi"]
    assert np.array_equal(np.sort(gs, order=key).view(np.uint8), np.sort(os_, order=key).view(np.uint8))
    for k in ("spectrum", "material_edep"):
        assert np.array_equal(g[k], o[k]), k
    assert g["total_histories"] == o["total_histories"] and g["final_seed"] == o["final_seed"]
    for k, v in o["counters"].items():
        if k != "coincidences":     # in this mode: the pairing's binned trues + scatter (below)
            assert g["stats"][k] == v, (k, g["stats"][k], v)
    # coincidences: the device sorter equals the CPU statement on the oracle's singles
    sinos, counts, raw = oracle_py.pair_singles(os_, WINDOW, _geom(sc.cfg), sc.oracle("fmath"))
    assert g["stats"]["coincidences"] == counts[0] + counts[1]
    for i, k in enumerate(("trues", "scatter", "randoms", "spurious")):
        assert np.array_equal(g[k], sinos[i]), k
        assert g["coincidences"][k] == counts[i], k
    assert g["coincidences"]["multiples"] == counts[4]
    return g, counts, raw


@pytest.mark.gpu
def test_singles_pipeline_bit_exact():
    # 1e5 decays in 1 ms: singles overlap, so there are randoms and multiples to sort out
    sc = _singles_scenario(n_decays=1e5, acq_s=1e-3, singles_capacity=400_000, material_dose=True)
    g, counts, raw = _check_against_oracle(sc)
    assert counts[0] > 100 and counts[2] > 10 and raw[4] > 0


@pytest.mark.gpu
def test_singles_with_prompt_gammas_and_branching_bit_exact():
    sc = _singles_scenario(n_decays=5e4, acq_s=1e-3, singles_capacity=400_000, positron_fraction=0.6,
                           prompt_gamma_fraction=0.5, prompt_gamma_energy=450e3)
    g, counts, raw = _check_against_oracle(sc)
    assert counts[3] > 0 and g["stats"]["prompts_emitted"] > 0


@pytest.mark.gpu
def test_singles_capacity_overflow_is_reported():
    sc = _singles_scenario(n_decays=2e4, acq_s=1e-3, singles_capacity=500)
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT, count_events=False)
    assert g["singles_count"] == 500 and g["singles_requested"] > 500
    assert len(g["singles"]) == 500 and np.all(np.diff(g["singles"]["time_ps"].astype(np.int64)) >= 0)


@pytest.mark.gpu
def test_fused_mode_with_prompt_gammas_bit_exact():
    """Prompt gammas in the reference's pair tallies: transported (dose), never tallied."""
    sc = water_cube_scenario(n_decays=5e4, tpb=16, material_dose=True, prompt_gamma_fraction=0.5,
                             prompt_gamma_energy=300e3, positron_fraction=0.8)
    o = sc.run_oracle("fmath")
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)
    assert_tallies_equal(g, o)
