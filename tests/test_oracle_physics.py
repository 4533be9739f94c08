"""Physics of the CPU oracle pinned by closed forms and invariants that do not restate its code:
exponential attenuation with coefficients read from the raw material files, Klein-Nishina,
Thomson, the sampled form-factor distribution, rotation geometry, source isotropy and
non-collinearity, the radioactive decay law, Michelogram geometry, the detector energy
resolution, and tally identities. Both oracle builds (libm and shared elementary functions)
are exercised."""
import gzip

import numpy as np
import pytest
from scipy import stats

import oracle_py
from conftest import AIR, WATER, names15
from mcpet import phantoms

E511 = np.float32(510998.95)
EPS = np.float32(0.00015)


def raw_table(path):
    rows = []
    with gzip.open(path, "rt") as f:
        for line in f:
            if line.startswith("#"):
                if rows:
                    break
                continue
            rows.append([float(x) for x in line.split()])
    return np.array(rows)


def mu_at(path, e, rho_nom, rho):
    """(total, Compton, Rayleigh, photo) attenuation coefficients (1/cm) at energy e and density rho,
    from the raw mean free paths (linear interpolation of inverse mfp, as tabulated)."""
    t = raw_table(path)
    out = []
    for col in (4, 2, 1, 3):
        inv = 1.0 / (t[:, col] * rho_nom)
        out.append(np.interp(e, t[:, 0], inv) * rho)
    return np.array(out)


def box_oracle(mat, dens, vs, flavor="libm", names=(AIR, WATER), **cfgkw):
    act = np.zeros(mat.shape, np.float32)
    act.flat[0] = 1.0
    cfg = phantoms.make_config(materials=list(names), **cfgkw)
    dm = np.full(15, -999.0, np.float32)
    for m in np.unique(mat):
        dm[m - 1] = dens[mat == m].max()
    tab = oracle_py.OracleTables(names15(names), dm, flavor)
    vmd = np.stack([mat.ravel().astype(np.float32) + np.float32(1e-4), dens.ravel(), act.ravel()], 1)
    return oracle_py.Oracle(phantoms.config_dict(cfg), vmd, mat.shape, vs, tab, flavor)


@pytest.fixture(scope="module")
def water_box():
    n = 40
    return box_oracle(np.full((n, n, n), 2, np.int32), np.ones((n, n, n), np.float32), (1.0, 1.0, 1.0))


def test_locate_voxel(water_box):
    o = water_box
    assert o.locate(0.5, 0.5, 0.5) == 0
    assert o.locate(1.5, 0.5, 0.5) == 1 and o.locate(0.5, 1.5, 0.5) == 40 and o.locate(0.5, 0.5, 1.5) == 1600
    assert o.locate(39.9, 39.9, 39.9) == 40 ** 3 - 1
    assert o.locate(float(EPS) * 0.99, 1, 1) == -100 and o.locate(40 - float(EPS) * 0.5, 1, 1) == -100
    assert o.locate(float(EPS) * 1.01, 1.0, 1.0) == 0 + 40 + 1600
    assert o.locate(np.nan, 1, 1) == -100 and o.locate(-1, 1, 1) == -100


@pytest.mark.parametrize("flavor", ["libm", "fmath"])
def test_first_interaction_exponential_attenuation(flavor):
    n = 40
    o = box_oracle(np.full((n, n, n), 2, np.int32), np.ones((n, n, n), np.float32), (1.0, 1.0, 1.0), flavor)
    N = 200_000
    pos = np.tile([20.0, 20.0, 20.0], (N, 1))
    d = np.tile([1.0, 0.0, 0.0], (N, 1))
    dist, typ = o.probe_first_interaction(7, 3, pos, d, E511)
    mu = mu_at(WATER, float(E511), 1.0, 1.0)
    L = 20.0 - float(EPS)
    p_esc = np.exp(-mu[0] * L)
    esc = (typ == 0).mean()
    assert abs(esc - p_esc) < 4 * np.sqrt(p_esc * (1 - p_esc) / N), (esc, p_esc)
    x = dist[typ > 0]
    cdf = lambda t: (1 - np.exp(-mu[0] * t)) / (1 - np.exp(-mu[0] * L))
    assert stats.kstest(x, cdf).pvalue > 1e-3
    frac = np.array([(typ == k).sum() for k in (1, 2, 3)]) / (typ > 0).sum()
    want = mu[1:] / mu[1:].sum()
    nint = (typ > 0).sum()
    for f, w in zip(frac, want):
        assert abs(f - w) < 5 * np.sqrt(w * (1 - w) / nint) + 1e-5, (frac, want)


def test_woodcock_through_heterogeneous_slabs():
    """Air, then 10 cm water, air, 3 cm bone-density water, air: the uncollided fraction is
    exp(-sum mu_i L_i) with the Woodcock majorant set by the 1.85 g/cm3 voxels."""
    nx = 40
    mat = np.ones((4, 4, nx), np.int32)
    dens = np.full((4, 4, nx), 0.0012, np.float32)
    mat[:, :, 10:20] = 2
    dens[:, :, 10:20] = 1.0
    mat[:, :, 25:28] = 2
    dens[:, :, 25:28] = 1.85
    o = box_oracle(mat, dens, (1.0, 1.0, 1.0))
    N = 300_000
    pos = np.tile([0.5, 2.0, 2.0], (N, 1))
    d = np.tile([1.0, 0.0, 0.0], (N, 1))
    dist, typ = o.probe_first_interaction(11, 5, pos, d, E511)
    mw = mu_at(WATER, float(E511), 1.0, 1.0)[0]
    ma = mu_at(AIR, float(E511), 0.00120479, 0.0012)[0]
    tau = ma * (40 - 0.5 - 13 - float(EPS)) + mw * 10 + mw * 1.85 * 3
    p = np.exp(-tau)
    esc = (typ == 0).mean()
    assert abs(esc - p) < 4 * np.sqrt(p * (1 - p) / N), (esc, p)
    # where the first interactions happen: fraction inside the first water slab
    in_w = ((dist > 9.5) & (dist < 19.5)).mean()
    want = np.exp(-ma * 9.5) * (1 - np.exp(-mw * 10))
    assert abs(in_w - want) < 4 * np.sqrt(want * (1 - want) / N), (in_w, want)


def klein_nishina_cdf(c, k=float(E511) / 510998.918):
    cs = np.linspace(-1, 1, 200001)
    r = 1.0 / (1 + k * (1 - cs))
    pdf = r * r * (r + 1 / r - (1 - cs * cs))
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(cs))])
    return np.interp(c, cs, cdf / cdf[-1])


@pytest.mark.parametrize("flavor", ["libm", "fmath"])
def test_compton_klein_nishina_at_511(water_box, flavor):
    o = water_box if flavor == "libm" else box_oracle(np.full((4, 4, 4), 2, np.int32), np.ones((4, 4, 4), np.float32),
                                                      (1.0, 1.0, 1.0), "fmath")
    e, c = o.probe_gcoa(13, 2, 300_000, E511, 1)
    assert np.all(e <= E511) and np.all(e > 0) and np.all(np.abs(c) <= 1)
    # binding effects only matter for very forward angles at 511 keV: compare cos < 0.99
    sel = c < 0.99
    cc = c[sel]
    F = lambda t: klein_nishina_cdf(t) / klein_nishina_cdf(0.99)
    ks = stats.kstest(cc, F)
    assert ks.statistic < 0.006, ks
    # energy: Compton line E/(1+k(1-cos)) broadened by the electron momentum (Doppler)
    ratio = e / (float(E511) / (1 + float(E511) / 510998.918 * (1 - c)))
    assert abs(np.median(ratio) - 1) < 2e-3 and np.std(ratio[c < 0.5]) < 0.03


def test_rayleigh_thomson_limit(water_box):
    c = water_box.probe_graa(17, 1, 200_000, np.float32(100.0), 1)     # xmax < 0.01: Thomson
    assert abs(c.mean()) < 0.006
    assert abs((c * c).mean() - 0.4) < 0.004          # E[cos^2] for pdf (3/8)(1+cos^2)


def test_rayleigh_form_factor_distribution(water_box):
    o = water_box
    c = o.probe_graa(19, 4, 200_000, E511, 1)
    xmax = float(E511) * 8.065535669099010e-5
    tab = o.tables.ray
    xco = np.array(tab.xco[128:256], np.float64)
    pco = np.array(tab.pco[128:256], np.float64)
    x2max = min(xmax * xmax, xco[-1])
    x2 = x2max * (1 - c) / 2
    # expected accepted CDF at the grid points: integral of the Thomson weight over dP (trapezoid)
    k = np.searchsorted(xco, x2max)
    xs = np.concatenate([xco[:k], [x2max]])
    ps = np.concatenate([pco[:k], [np.interp(x2max, xco, pco)]])
    w = lambda x: (1 + (1 - 2 * x / x2max) ** 2) / 2
    dP = np.diff(ps) * 0.5 * (w(xs[1:]) + w(xs[:-1]))
    Fexp = np.concatenate([[0], np.cumsum(dP)]) / dP.sum()
    Femp = np.array([(x2 < xv).mean() for xv in xs])
    assert np.max(np.abs(Femp - Fexp)) < 0.01


def test_rotate_geometry():
    rng = np.random.default_rng(3)
    N = 50_000
    v = rng.normal(size=(N, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    v[:10] = [0, 0, 1]
    v[10:20] = [0, 0, -1]
    ct = rng.uniform(-1, 1, N)
    ph = rng.uniform(0, 2 * np.pi, N)
    out = oracle_py.probe_rotate(v.astype(np.float32), ct, ph)
    assert np.max(np.abs(np.linalg.norm(out, axis=1) - 1)) < 2e-6
    assert np.max(np.abs((out * v).sum(1) - ct)) < 3e-6
    st = np.sqrt(1 - ct[:10] ** 2)
    assert np.allclose(out[:10], np.stack([st * np.cos(ph[:10]), st * np.sin(ph[:10]), ct[:10]], 1), atol=2e-6)
    st = np.sqrt(1 - ct[10:20] ** 2)
    assert np.allclose(out[10:20], np.stack([-st * np.cos(ph[10:20]), st * np.sin(ph[10:20]), -ct[10:20]], 1), atol=2e-6)
    # azimuth about a fixed axis is uniform
    a = np.tile(np.array([0.6, 0.0, 0.8], np.float32), (N, 1))
    out = oracle_py.probe_rotate(a, np.zeros(N), rng.uniform(0, 2 * np.pi, N))
    perp = out - (out * a).sum(1)[:, None] * a
    assert np.all(np.abs(perp.mean(0)) < 0.012)


@pytest.mark.parametrize("flavor", ["libm", "fmath"])
def test_source_isotropy_position_and_noncollinearity(flavor):
    o = box_oracle(np.full((5, 5, 5), 2, np.int32), np.ones((5, 5, 5), np.float32), (0.7, 0.7, 0.7), flavor)
    N = 200_000
    p1, d1, d2 = o.probe_source(21, 8, N, 2, 3, 1)
    u = p1 / np.float32(0.7) - np.array([2, 3, 1])
    assert u.min() >= float(EPS) - 2e-6 and u.max() <= 1 - float(EPS) + 2e-6     # float rounding of (i+r)/inv
    for k in range(3):
        assert stats.kstest(u[:, k], "uniform", args=(float(EPS), 1 - 2 * float(EPS))).pvalue > 1e-3
    assert stats.kstest(d1[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-3
    phi = np.arctan2(d1[:, 1], d1[:, 0])
    assert stats.kstest(phi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > 1e-3
    alpha = np.arccos(np.clip(-(d1 * d2).sum(1), -1, 1))
    sigma = 2 * 0.21276 * np.pi / 180
    assert abs(alpha.mean() / (sigma * np.sqrt(2 / np.pi)) - 1) < 0.01
    assert abs(np.sqrt((alpha ** 2).mean()) / sigma - 1) < 0.01
    o.s.bg_mode = 1
    p1, d1, d2 = o.probe_source(23, 9, N, 2, 3, 1)
    assert abs((d1 * d2).sum(1).mean()) < 0.01 and stats.kstest(d2[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-3


def test_decay_law_and_photon_count():
    """Expected decays A*tau*(1-exp(-T/tau)) with the activity decaying during the scan
    (not A*T), two photons per decay."""
    mat = np.full((1, 1, 1), 1, np.int32)
    dens = np.full((1, 1, 1), 0.0012, np.float32)
    o = box_oracle(mat, dens, (1.0, 1.0, 1.0), names=(AIR,), threads_per_block=64, acquisition_time_s=4.0,
                   mean_life=2.0)
    o.vmd[0, 2] = 2.0e4
    r = o.run(nthreads=4)
    dec = r["counters"]["decays"]
    want = 2.0e4 * 2.0 * (1 - np.exp(-2.0))
    assert abs(dec - want) < 4.5 * np.sqrt(want), (dec, want)
    assert r["total_histories"] == 2 * dec
    assert r["counters"]["streams"] == 64


def test_michelogram_geometry(water_box):
    o = water_box
    P = o.p
    ncr, nang, nrad = P["ncrystals"], P["nangles"], P["nrad"]
    rng = np.random.default_rng(8)
    for _ in range(3000):
        ix1, ix2 = rng.integers(0, ncr, 2)
        iz1, iz2 = rng.integers(0, P["nrows"], 2)
        b = o.sinogram_bin(int(ix1), int(iz1), int(ix2), int(iz2))
        assert -1 <= b < P["nbins"]
        if abs(int(iz1) - int(iz2)) > P["mrd"]:
            assert b == -1
        b2 = o.sinogram_bin(int(ix2), int(iz2), int(ix1), int(iz1))
        if b >= 0 and b2 >= 0:
            assert b % (nang * nrad) == b2 % (nang * nrad)        # same angle and radius after swapping
            if iz1 == iz2:
                assert b == b2
    # opposite crystals -> central radial bin; parallel chords share the angle bin
    for ix1 in range(0, nang, 7):
        b = o.sinogram_bin(ix1, 10, ix1 + nang, 10)
        assert b % nrad == nrad // 2
        ref = (o.sinogram_bin(ix1, 10, ix1 + nang, 10) // nrad) % nang
        for k in (1, 5, 30):
            bb = o.sinogram_bin((ix1 + k) % ncr, 10, (ix1 + nang - k) % ncr, 10)
            if bb >= 0:
                assert (bb // nrad) % nang == ref
    # radial index grows with the chord's distance from the centre, R*|cos(pi*delta/ncr)|
    dist, rad = [], []
    for dlt in range(nang - nrad // 2 + 1, nang + nrad // 2):
        b = o.sinogram_bin(40, 10, 40 + dlt, 10)
        if b >= 0:
            dist.append(abs(np.cos(np.pi * dlt / ncr)))
            rad.append(abs(b % nrad - nrad // 2))
    assert stats.spearmanr(dist, rad).correlation > 0.999


def test_energy_resolution_and_tally_identities():
    """Trues-only coincidences: every spectrum entry is a blurred 511 keV photon, so the
    spectrum is N(511 keV, sigma = 0.12/2.35 * 511 keV). Image sums equal sinogram sums and the
    per-voxel dose over the whole phantom equals the per-material dose, exactly."""
    mat, dens, act, vs = phantoms.water_cube(n_decays=2e5)
    cfg = phantoms.make_config(threads_per_block=32, materials=[AIR, WATER], tally_type=1, material_dose=True,
                               voxel_dose=True, roi=(0, 8, 0, 8, 0, 8))
    dm = np.full(15, -999.0, np.float32)
    dm[1] = 1.0
    tab = oracle_py.OracleTables(names15([AIR, WATER]), dm, "libm")
    vmd = np.stack([mat.ravel() + np.float32(1e-4), dens.ravel(), act.ravel()], 1).astype(np.float32)
    r = oracle_py.Oracle(phantoms.config_dict(cfg), vmd, mat.shape, vs, tab, "libm").run(nthreads=8)
    assert r["scatter"].sum() == 0
    assert r["image_trues"].sum() == r["trues"].sum() and r["image_scatter"].sum() == 0
    sp = r["spectrum"].astype(np.float64)
    assert sp.sum() == 2 * r["counters"]["pairs_tallied"]
    e = np.arange(len(sp)) + 0.5
    mean = (sp * e).sum() / sp.sum()
    sd = np.sqrt((sp * (e - mean) ** 2).sum() / sp.sum() - 1 / 12)
    sigma = 0.12 / 2.35 * 510.99895
    assert abs(mean - 510.999) < 0.5 and abs(sd / sigma - 1) < 0.02, (mean, sd, sigma)
    assert int(r["voxel_edep"][:, 0].sum()) == int(r["material_edep"][1, 0])
    assert int(r["voxel_edep"][:, 1].sum()) == int(r["material_edep"][1, 1])


def test_libm_and_shared_math_builds_agree():
    """The two oracle builds on independent seeds: totals within Poisson errors. On the same
    seed they follow the same histories almost always (both sets of elementary functions are
    correctly rounded but for rare cases)."""
    mat, dens, act, vs = phantoms.water_cube(n_decays=2e5)
    cfg = phantoms.make_config(threads_per_block=32, materials=[AIR, WATER], seed=5)
    dm = np.full(15, -999.0, np.float32)
    dm[1] = 1.0
    vmd = np.stack([mat.ravel() + np.float32(1e-4), dens.ravel(), act.ravel()], 1).astype(np.float32)
    res = {}
    for fl, seed in (("libm", 5), ("fmath", 6), ("fmath", 5)):
        p = phantoms.config_dict(cfg)
        p["seed"] = seed
        tab = oracle_py.OracleTables(names15([AIR, WATER]), dm, fl)
        res[(fl, seed)] = oracle_py.Oracle(p, vmd, mat.shape, vs, tab, fl).run(nthreads=8)
    a, b = res[("libm", 5)], res[("fmath", 6)]
    for k in ("trues", "scatter"):
        x, y = int(a[k].sum()), int(b[k].sum())
        assert abs(x - y) / np.sqrt(x + y) < 4
    same = res[("fmath", 5)]
    assert (a["image_trues"] == same["image_trues"]).mean() > 0.9
