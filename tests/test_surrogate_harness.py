"""Scatter surrogate harness (no reference counterpart): the NRMSE / time-to-5 % formulas, the
consistency check that triggers the Monte Carlo fallback, the input normalisation and the
training phantoms, each checked against Poisson simulation or an invariant on the CPU."""
import numpy as np

from mcpet import phantoms
from mcpet import surrogate as S


def _smooth_field(rng, shape=S.GRID):
    z, y, x = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        c = rng.uniform(-0.7, 0.7, 3)
        f += rng.uniform(0.5, 2) * np.exp(-((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2) / rng.uniform(0.1, 0.5))
    return f + 0.05


def test_decays_for_gives_a_5_percent_poisson_estimate():
    """Monte Carlo estimates at the decays N5 the formula predicts have NRMSE 5 % (Poisson
    counts drawn around the reference)."""
    rng = np.random.default_rng(1)
    r = _smooth_field(rng) * 1e-6                     # expected counts per decay
    n5 = S.decays_for(r)
    errs = [S.nrmse(rng.poisson(r * n5) / n5, r) for _ in range(20)]
    assert abs(np.mean(errs) - 0.05) < 0.002, np.mean(errs)
    # and 4x the decays halves it
    errs4 = [S.nrmse(rng.poisson(r * 4 * n5) / (4 * n5), r) for _ in range(10)]
    assert abs(np.mean(errs4) - 0.025) < 0.001


def test_poisson_floor_is_the_expected_nrmse_of_the_counts():
    rng = np.random.default_rng(2)
    lam = _smooth_field(rng) * 30
    floor = S.poisson_floor(lam)
    errs = [S.nrmse(rng.poisson(lam).astype(float), lam) for _ in range(20)]
    assert abs(np.mean(errs) / floor - 1) < 0.02


def test_consistency_z_is_standard_normal_for_consistent_data_and_large_for_a_bias():
    rng = np.random.default_rng(3)
    pred = _smooth_field(rng) * 2e-7                 # per decay
    n_low = 5e6
    z = [S.consistency_z(pred, rng.poisson(pred * n_low), n_low) for _ in range(200)]
    assert abs(np.mean(z)) < 0.25 and 0.8 < np.std(z) < 1.25, (np.mean(z), np.std(z))
    assert np.mean(np.abs(z) > 5) == 0.0
    # a prediction 10 % too low everywhere, or with a 25 % dent over a corner, is rejected
    obs = rng.poisson(pred * n_low)
    assert S.consistency_z(pred * 0.9, obs, n_low) > 5
    dented = pred.copy()
    dented[:16, :28, :28] *= 0.75
    assert S.consistency_z(dented, obs, n_low) > 5


def test_normalize_is_invariant_to_the_number_of_decays():
    rng = np.random.default_rng(4)
    s = rng.poisson(3.0, S.GRID).astype(float)
    t = rng.poisson(9.0, S.GRID).astype(float)
    x1, f1 = S.normalize(s, t, 5e6)
    x2, f2 = S.normalize(2 * s, 2 * t, 1e7)
    assert np.array_equal(x1, x2) and f1 == f2           # counts x 2^k scale exactly
    assert abs(x1[1].mean() - 1) < 1e-6 and x1.dtype == np.float32 and x1.shape == (2,) + S.GRID
    # the output factor turns normalised units back into counts per decay
    assert abs(float(f1) * float(x1[1].sum()) * 5e6 / t.sum() - 1) < 1e-6


def test_normalize_with_maps_scales_activity_by_its_maximum():
    """4-channel input: channels 0-1 as without maps; channel 2 the activity integrals times
    float32(1 / max) (maximum exactly 1), channel 3 the attenuation factors unchanged."""
    rng = np.random.default_rng(5)
    s = rng.poisson(3.0, S.GRID).astype(np.float32)
    t = rng.poisson(9.0, S.GRID).astype(np.float32)
    act = (rng.random(S.GRID) * 3e4).astype(np.float32)
    att = rng.random(S.GRID).astype(np.float32)
    x2, f2 = S.normalize(s, t, 5e6)
    x4, f4 = S.normalize(s, t, 5e6, (act, att))
    assert x4.shape == (4,) + S.GRID and x4.dtype == np.float32 and f4 == f2
    assert np.array_equal(x4[:2], x2) and np.array_equal(x4[3], att)
    assert x4[2].max() == 1.0 and np.array_equal(x4[2], act * np.float32(1.0 / float(act.max())))
    z, _ = S.normalize(s, t, 5e6, (np.zeros_like(act), att))
    assert np.all(z[2] == 0)


def test_packaged_weights_fit_the_four_channel_network():
    """The packaged parameters are a width-32, 4-input-channel network's (the layer list of
    mcpet_surrogate_create), and their training log says so."""
    import json
    import os
    w = np.load(S.WEIGHTS, allow_pickle=False)
    assert w.dtype == np.float32 and w.size == S.param_count(S.IN_CHANNELS, S.WIDTH) == 1275841
    assert S.param_count(2, 32) == 1274113 and S.param_count(4, 56) == 3901969
    with open(os.path.splitext(S.WEIGHTS)[0] + ".json") as f:
        info = json.load(f)
    assert info["in_channels"] == 4 and info["width"] == 32 and list(info["grid"]) == list(S.GRID)
    assert np.all(np.isfinite(w))


def test_workload_config():
    c = S.config([], 5e6)
    assert c.acquisition_time_ps == 5_000_000_000 and c.stream_decays == 5 and c.nrad == 168 and c.nsinos == 1293
    assert S.config([], 7e8).stream_decays == 668 and S.stream_decays(1e3) == 4


def test_scatter_phantoms_are_seeded_and_scaled():
    a = phantoms.scatter_phantoms(3, 5, n_decays=1e9, acquisition_time_s=1.0)
    b = phantoms.scatter_phantoms(3, 5, n_decays=1e9, acquisition_time_s=1.0)
    c = phantoms.scatter_phantoms(1, 6, n_decays=1e9, acquisition_time_s=1.0)
    frac = phantoms.F18_MEAN_LIFE * (1 - np.exp(-1.0 / phantoms.F18_MEAN_LIFE))
    for (m, d, act, vs), (m2, d2, act2, _) in zip(a, b):
        assert m.shape == (64, 128, 128) and vs == (0.2, 0.2, 0.2)
        assert np.array_equal(m, m2) and np.array_equal(act, act2)
        assert abs(act.sum(dtype=np.float64) * frac / 1e9 - 1) < 1e-5
        assert set(np.unique(m)) <= set(range(1, 9)) and (m == 2).sum() > 0
        assert np.all(act[m == 1] == 0) and np.all(d > 0)
    assert not np.array_equal(a[0][0], c[0][0])
