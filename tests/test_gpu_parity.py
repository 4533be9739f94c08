"""GPU parity: the MI355X library against the CPU oracle on the same inputs.

Bit-exact tier: the oracle built with the library's elementary functions ("fmath") follows
the reference algorithm stream by stream, so every integer tally, the PSF multiset, the
dose sums, the final seed and the event counters must be identical.
Statistical tier: the oracle built on the C library ("libm") is an independent arithmetic;
totals must agree within Poisson errors and the energy spectrum by a chi-square test.
"""
import sys

import numpy as np
import pytest
from scipy import stats

import mcpet
import oracle_py
from harness import assert_tallies_equal, head_scenario, sample_scenario, water_cube_scenario

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_ranecu_bit_exact():
    _say("ranecu: 12 streams x 4096 draws, float and double")
    ids = np.array([0, 1, 2, 31, 1000, 123456, 999999, 2**31 - 2, 2**31 - 1, 2**31 + 7, 5 * 2**33 + 3, 10**11], np.int64)
    for seed in (1, 12345, 2147483562, -17):
        g = mcpet.selftest_ranecu(seed, ids, 4096)
        o = oracle_py.ranecu_block(seed, ids, 4096)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
        gd = mcpet.selftest_ranecu(seed, ids, 512, as_double=True)
        od = oracle_py.ranecu_block(seed, ids, 512, as_double=True)
        assert np.array_equal(gd.view(np.uint64), od.view(np.uint64))


def test_fmath_host_device_bit_exact():
    _say("fmath: 200k inputs, 15 functions, device vs gcc host")
    rng = np.random.default_rng(5)
    n = 200_000
    x = np.concatenate([rng.uniform(-20, 20, n // 4), rng.uniform(-1, 1, n // 4),
                        np.exp(rng.uniform(-30, 6, n // 4)) * rng.choice([-1, 1], n // 4),
                        rng.uniform(-700, 700, n // 8), rng.uniform(-3e4, 3e4, n - 3 * (n // 4) - n // 8)])
    y = rng.uniform(-1.5, 1.5, n)
    gd, gf = mcpet.selftest_fmath(x, y)
    hd, hf = oracle_py.fmath_eval(x, y)
    for i in range(7):
        bad = np.flatnonzero(gd[i].view(np.uint64) != hd[i].view(np.uint64))
        assert bad.size == 0, f"double fn {i}: {bad.size} mismatches, first x={x[bad[0]]!r} y={y[bad[0]]!r}"
    for i in range(8):
        badf = np.flatnonzero(gf[i].view(np.uint32) != hf[i].view(np.uint32))
        assert badf.size == 0, f"float fn {i}: {badf.size} mismatches, first x={x[badf[0]]!r} y={y[badf[0]]!r}"


@pytest.mark.parametrize("kernel", [mcpet.KERNEL_PLAIN, mcpet.KERNEL_PERSISTENT])
def test_water_cube_bit_exact(kernel):
    _say(f"water cube 2e5 decays, tpb 32, sinogram+PSF+doses, kernel {kernel}")
    sc = water_cube_scenario(n_decays=2e5, tpb=32, tally_psf_sinogram=0, psf_size=400_000, material_dose=True,
                             voxel_dose=True, roi=(1, 7, 0, 8, 2, 6))
    g = sc.run_gpu(kernel=kernel)
    o = sc.run_oracle("fmath")
    assert o["trues"].sum() > 1000 and o["scatter"].sum() > 100
    assert_tallies_equal(g, o)


def test_water_cube_full_workload_bit_exact():
    """The small workload at its full size: the 9 x 9 x 9 water cube (1 cm voxels, three hot
    voxels at 100x), 1e7 decays, sinogram, emission images, spectrum and material dose,
    bit-exact against the oracle."""
    _say("water cube 9x9x9, 1e7 decays (the small workload's full size), tpb 32, persistent kernel")
    sc = water_cube_scenario(n_decays=1e7, tpb=32, material_dose=True)
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)
    o = sc.run_oracle("fmath", nthreads=16)
    assert o["trues"].sum() > 100000 and o["counters"]["decays"] > 9.9e6
    assert_tallies_equal(g, o)


def test_sample_simulation_bit_exact():
    _say("reference sample .in (R 9.05 cm), 2 s acquisition, PSF + sinogram + material dose")
    sc = sample_scenario(acq_s=2.0)
    g = sc.run_gpu()
    o = sc.run_oracle("fmath")
    assert o["total_histories"] > 40000
    assert_tallies_equal(g, o)


@pytest.mark.parametrize("extra", [dict(bg_mode=1), dict(tally_type=1, tally_psf_sinogram=0, psf_size=100_000),
                                   dict(tally_type=2, tally_psf_sinogram=1, psf_size=100_000)])
def test_modes_bit_exact(extra):
    _say(f"water cube modes {extra}")
    bg = extra.pop("bg_mode", 0)
    sc = water_cube_scenario(n_decays=5e4, tpb=16, **extra)
    sc.cfg.bg_mode = bg
    g = sc.run_gpu()
    o = sc.run_oracle("fmath")
    assert_tallies_equal(g, o)


def test_head_small_bit_exact():
    _say("head phantom 32x32x24 at 8 mm, 8 tissues, 1e5 decays")
    sc = head_scenario(n_decays=1e5, material_dose=True, tally_psf_sinogram=0, psf_size=200_000)
    g = sc.run_gpu()
    o = sc.run_oracle("fmath")
    assert o["counters"]["compton"] > 10000
    assert_tallies_equal(g, o)


def statistical_criteria(g, o, cfg, shape, image_block=1):
    """The transport's statistical criteria between two independent simulations of the same
    workload: totals of trues and scatter within 3 sigma, energy spectrum chi-square p > 0.01,
    per-bin z-scores of the emission images (in blocks of image_block voxels per axis) and of the
    sinograms (summed over planes, blocks of 8 angles x 3 radial bins) with mean within 0.2 of 0
    and spread 0.8-1.25."""
    for k in ("trues", "scatter"):
        a, b = int(g[k].sum()), int(o[k].sum())
        z = (a - b) / np.sqrt(a + b)
        _say(f"{k}: gpu {a} oracle {b} z={z:.2f}")
        assert abs(z) < 3.0, f"{k}: gpu {a} oracle {b} z={z:.2f}"
    ga, oa = g["spectrum"].astype(float), o["spectrum"].astype(float)
    m = (ga + oa) > 20
    chi2 = (((ga[m] - oa[m]) ** 2) / (ga[m] + oa[m])).sum()
    p = stats.chi2.sf(chi2, m.sum() - 1)
    _say(f"spectrum chi2={chi2:.1f} dof={m.sum() - 1} p={p:.3g}")
    assert p > 0.01, f"spectrum chi2={chi2:.1f} dof={m.sum() - 1} p={p:.2g}"

    def blocks(img):
        a = img.astype(float).reshape(shape)
        nb = [n // image_block * image_block for n in shape]
        a = a[:nb[0], :nb[1], :nb[2]]
        return a.reshape(nb[0] // image_block, image_block, nb[1] // image_block, image_block,
                         nb[2] // image_block, image_block).sum(axis=(1, 3, 5)).ravel()

    for which in ("image_trues", "image_scatter"):
        gi, oi = blocks(g[which]), blocks(o[which])
        mm = (gi + oi) > 50
        zz = (gi[mm] - oi[mm]) / np.sqrt(gi[mm] + oi[mm])
        _say(f"{which} z: {mm.sum()} bins, mean {zz.mean():.3f} std {zz.std():.3f}")
        assert mm.sum() > 20 and abs(zz.mean()) < 0.2 and 0.8 < zz.std() < 1.25, (which, zz.mean(), zz.std())

    def coarse(t):
        a = t.astype(float).reshape(cfg.nsinos, cfg.nangles, cfg.nrad).sum(0)
        a = a[:cfg.nangles // 8 * 8, :cfg.nrad // 3 * 3]
        return a.reshape(cfg.nangles // 8, 8, cfg.nrad // 3, 3).sum(axis=(1, 3)).ravel()

    for which in ("trues", "scatter"):
        gs, os_ = coarse(g[which]), coarse(o[which])
        ms = (gs + os_) > 50
        zs = (gs[ms] - os_[ms]) / np.sqrt(gs[ms] + os_[ms])
        _say(f"{which} sinogram z: {ms.sum()} bins, mean {zs.mean():.3f} std {zs.std():.3f}")
        assert ms.sum() > 100 and abs(zs.mean()) < 0.2 and 0.8 < zs.std() < 1.25, (which, zs.mean(), zs.std())


def test_statistical_vs_libm_oracle():
    _say("water cube 1e6 decays: GPU vs libm oracle on an independent seed, Poisson totals, spectrum chi2, image z")
    sc = water_cube_scenario(n_decays=1e6, tpb=64, seed=99, material_dose=True)
    g = sc.run_gpu()
    sc_o = water_cube_scenario(n_decays=1e6, tpb=64, seed=1234567, material_dose=True)
    o = sc_o.run_oracle("libm")
    statistical_criteria(g, o, sc.cfg, sc.phantom.shape)


def test_statistical_vs_libm_oracle_on_the_head():
    """The same criteria on heterogeneous geometry: the head phantom at 64 x 64 x 48 voxels of
    4 mm (scalp, fat, skull bone at 1.85 g/cm3, grey and white matter, CSF, sinus air), 3e6
    decays, where the tissue boundaries exercise material changes, Compton events in bone and
    the air gaps; GPU against the oracle built on the system libm with an independent seed."""
    _say("head 64x64x48 at 4 mm, 3e6 decays: GPU vs libm oracle on an independent seed")
    sc = head_scenario(shape=(48, 64, 64), voxel_cm=0.4, n_decays=3e6, tpb=16, seed=2027)
    g = sc.run_gpu(count_events=False)
    sc_o = head_scenario(shape=(48, 64, 64), voxel_cm=0.4, n_decays=3e6, tpb=16, seed=8675309)
    o = sc_o.run_oracle("libm", nthreads=16)
    assert o["trues"].sum() > 50000 and o["scatter"].sum() > 20000
    statistical_criteria(g, o, sc.cfg, sc.phantom.shape, image_block=8)


def test_head_full_size_invariants():
    """The bench workload (128x128x96 head, 2 mm, 1e9 decays in 10 s) on the GPU: size-independent
    checks (trues/scatter sums equal image sums, photon count = 2 x decays with the expected
    Poisson mean, counters consistent) plus a statistical match of the per-decay rates against
    the oracle on a sample of the same phantom."""
    from mcpet import phantoms
    from conftest import MATERIALS
    _say("head 128x128x96, 1e9 decays: invariants")
    mat, dens, act, vs = phantoms.head()
    names = phantoms.head_materials(MATERIALS)
    cfg = phantoms.make_config(seed=20261203, threads_per_block=32, acquisition_time_s=10.0, nrad=168, materials=names)
    ph = mcpet.phantom_from_arrays(mat, dens, act, vs)
    tab = mcpet.load_materials(names, ph.density_max())
    with mcpet.Simulation(cfg, ph, tab) as sim:
        sim.run(count_events=True)
        out = sim.fetch(psf=False)
        st = sim.stats()
    decays = st["decays"]
    assert out["total_histories"] == 2 * decays
    expected = float(act.astype(np.float64).sum()) * phantoms.F18_MEAN_LIFE * (1 - np.exp(-10.0 / phantoms.F18_MEAN_LIFE))
    assert abs(decays - expected) < 6 * np.sqrt(expected), (decays, expected)
    assert out["trues"].sum() == out["image_trues"].sum()
    assert out["scatter"].sum() == out["image_scatter"].sum()
    assert st["coincidences"] == out["trues"].sum() + out["scatter"].sum()
    assert out["spectrum"].sum() <= 2 * st["pairs_tallied"]
    # oracle on every 499th voxel of the same phantom and seed
    sc_o = oracle_py.Oracle(phantoms.config_dict(cfg), ph.voxel_mat_dens(), ph.shape, vs,
                            oracle_py.OracleTables(names + [""] * (15 - len(names)), ph.density_max(), "libm"), "libm")
    o = sc_o.run(nthreads=16, voxel_stride=499)
    od = o["counters"]["decays"]
    for key in ("coincidences", "compton", "woodcock_steps"):
        rg, ro = st[key] / decays, o["counters"][key] / od
        assert abs(rg - ro) / ro < 0.05, f"{key}/decay gpu {rg:.4f} oracle-sample {ro:.4f}"


def test_head_full_shape_slice_bit_exact():
    """The head workload's phantom and scanner at their full size (128 x 128 x 96 at 2 mm, 8
    tissues, 168 x 168 x 1293 sinogram, 32 streams per emitting voxel) over a 10 ms slice of
    its acquisition (1e6 decays at the workload's rate), bit-exact against the oracle."""
    from mcpet import phantoms
    from conftest import MATERIALS
    _say("head 128x128x96 full shape, 10 ms slice (1e6 decays), tpb 32")
    mat, dens, act, vs = phantoms.head()                 # 1e9 decays in 10 s
    names = phantoms.head_materials(MATERIALS)
    cfg = phantoms.make_config(seed=4711, threads_per_block=32, acquisition_time_s=0.01, nrad=168, materials=names,
                               material_dose=True)
    from harness import Scenario
    sc = Scenario(cfg, mat, dens, act, vs, names)
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)
    o = sc.run_oracle("fmath", nthreads=16)
    assert abs(o["counters"]["decays"] - 1e6) < 6e3 and o["trues"].sum() > 20000
    assert_tallies_equal(g, o)


def test_production_kernel_equals_instrumented_and_oracle():
    _say("water cube 2e5 decays: uninstrumented persistent kernel vs oracle (tallies, PSF, doses, lite counters)")
    sc = water_cube_scenario(n_decays=2e5, tpb=32, tally_psf_sinogram=0, psf_size=500_000, material_dose=True)
    o = sc.run_oracle("fmath")
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT, count_events=False)
    assert_tallies_equal(g, o, counters="lite")


@pytest.mark.parametrize("k_distinct,fmt", [(6, "u8 palette"), (3000, "u16 palette"), (0, "8-byte words")])
def test_voxel_layouts_bit_exact(k_distinct, fmt):
    """Phantoms with few, many and all-distinct (density, material) pairs select the u8-palette,
    u16-palette and 8-byte device layouts; each equals the oracle bit for bit, and forcing the
    8-byte layout gives the same tallies."""
    from conftest import AIR, WATER
    from harness import Scenario
    from mcpet import phantoms
    _say(f"48^3 voxels, {k_distinct or 'all'} distinct densities ({fmt}), 1e5 decays")
    n = 48
    rng = np.random.default_rng(100 + k_distinct)
    mat = rng.choice([1, 2], size=(n, n, n), p=[0.1, 0.9]).astype(np.int32)
    if k_distinct:
        vals = rng.uniform(0.3, 1.9, k_distinct).astype(np.float32)
        dens = vals[rng.integers(0, k_distinct, mat.shape)]
    else:
        dens = rng.uniform(0.3, 1.9, mat.shape).astype(np.float32)
    frac = phantoms.F18_MEAN_LIFE * (1.0 - np.exp(-1.0 / phantoms.F18_MEAN_LIFE))
    act = np.full(mat.shape, 1e5 / (mat.size * frac), np.float32)
    cfg = phantoms.make_config(seed=777, threads_per_block=4, materials=[AIR, WATER], material_dose=True)
    sc = Scenario(cfg, mat, dens, act, (0.3, 0.3, 0.3), [AIR, WATER])
    o = sc.run_oracle("fmath")
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)
    assert_tallies_equal(g, o)
    sc.cfg.voxel_format = 1
    g8 = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT, count_events=False)
    assert_tallies_equal(g8, o, counters="lite")
