"""The coincidence phase-space file (PSF) pinned by properties that do not restate its code
(reference tally_PSF_coincidences, MCGPU-PET_kernel.cu:530-644, record MCGPU-PET.h:235-247):

- the two records of a pair are consecutive and share the emission time and voxel, and the
  emission time T - remaining lies in [0, T) with the density of a decaying source
  (activity proportional to exp(-t/tau));
- the record with the shorter travel comes first;
- every record hits the cylinder inside its height, with a unit direction, index2 = 0 and a
  scatter flag in 0..3;
- an unscattered photon (index1 = 0) travelled in a straight line: going back from its hit
  point (R cos phi, R sin phi, z) against its direction by travel_time_ps x c lands inside the
  emission voxel, and both photons of a true pair land on the same point;
- the energies of unscattered photons follow the detector's Gaussian, sigma = R/2.35 x 511 keV,
  truncated to the energy window, and every energy lies inside the window;
- with both the sinogram and the PSF tallied (mode 0), each pair draws its own energy blur for
  either tally, so the true pairs in the PSF and the trues of the sinogram are two samples of the
  same acceptance: their difference is within a few standard deviations of zero;
- the pair counter is twice the number of stored pairs while the file has room.

The CPU test checks the oracle (both builds), the GPU test the library's device PSF."""
import sys

import numpy as np
import pytest
from scipy import stats

from harness import water_cube_scenario

C_CM_PER_PS = 2.99792458e-2        # speed of light, cm/ps (SI definition of the metre)
E_ANNIHILATION = 510998.95         # eV, the pair photon energy the source emits
TAU_S = 1.0                        # a short mean life so that the time density is far from flat
ACQ_S = 1.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def psf_scenario(n_decays, seed):
    # three threads per voxel: a high per-stream rate, so the activity's decay between two
    # decays of one stream is small (the time-density check assumes a smooth decay)
    return water_cube_scenario(n_decays=n_decays, tpb=3, seed=seed, tally_psf_sinogram=0,
                               psf_size=int(n_decays), mean_life=TAU_S, acquisition_time_s=ACQ_S)


def detector_frame(cfg, shape, vs):
    """Effective cylinder (centre, radius): a negative radius centres it on the voxel box."""
    nz, ny, nx = shape
    R = float(cfg.detector.radius)
    if R < 0:
        return np.array([nx * vs[0] / 2, ny * vs[1] / 2, nz * vs[2] / 2]), -R
    return np.array(list(cfg.detector.center), np.float64), R


def check_psf(psf, requested, trues_sum, cfg, shape, vs):
    """Assert every invariant listed in the module docstring; returns a dict of figures."""
    n = len(psf)
    assert n % 2 == 0 and n > 2000, n
    assert requested == n, (requested, n)            # every requested slot stored (room left)
    a, b = psf[0::2], psf[1::2]
    # pairs: same emission time and voxel, time inside the acquisition
    T_ps = cfg.acquisition_time_ps
    assert np.array_equal(a["emission_time_ps"], b["emission_time_ps"])
    assert np.array_equal(a["emission_absvox"], b["emission_absvox"])
    t = a["emission_time_ps"].astype(np.int64)
    assert t.min() >= 0 and t.max() < T_ps
    # shorter travel first
    assert np.all(a["travel_time_ps"] <= b["travel_time_ps"])
    # on the cylinder inside its height, unit direction, flags
    H = float(cfg.detector.height)
    assert np.all(np.abs(psf["z"]) < 0.5 * H)
    v = np.stack([psf["vx"], psf["vy"], psf["vz"]], 1).astype(np.float64)
    assert np.all(np.abs(np.linalg.norm(v, axis=1) - 1.0) < 1e-5)
    assert np.all(psf["index2"] == 0)
    assert set(np.unique(psf["index1"]).tolist()) <= {0, 1, 2, 3}
    # energies inside the window (open interval)
    E = psf["energy"].astype(np.float64)
    assert np.all(E > cfg.e_low) and np.all(E < cfg.e_high)
    # straight-line back-projection of unscattered photons into the emission voxel
    centre, R = detector_frame(cfg, shape, vs)
    nz, ny, nx = shape
    phi = psf["phi"].astype(np.float64)
    hit = np.stack([R * np.cos(phi), R * np.sin(phi), psf["z"].astype(np.float64)], 1) + centre
    d = psf["travel_time_ps"].astype(np.float64) * C_CM_PER_PS
    emit = hit - v * d[:, None]
    vox = psf["emission_absvox"].astype(np.int64)
    ijk = np.stack([vox % nx, (vox // nx) % ny, vox // (nx * ny)], 1)
    lo, hi = ijk * np.asarray(vs), (ijk + 1) * np.asarray(vs)
    un = psf["index1"] == 0
    tol = 2e-3
    inside = np.all((emit >= lo - tol) & (emit <= hi + tol), axis=1)
    assert un.sum() > 1000
    bad = np.flatnonzero(un & ~inside)
    assert bad.size == 0, f"{bad.size} unscattered records do not trace back into their voxel, e.g. {emit[bad[0]]} " \
                          f"voxel {ijk[bad[0]]}"
    true_pair = (a["index1"] == 0) & (b["index1"] == 0)
    sep = np.linalg.norm(emit[0::2] - emit[1::2], axis=1)[true_pair]
    assert sep.max() < 2 * tol, sep.max()
    # the detector's energy response on unscattered photons, truncated to the window
    sig = cfg.e_resol / 2.35 * E_ANNIHILATION
    lo_s, hi_s = (cfg.e_low - E_ANNIHILATION) / sig, (cfg.e_high - E_ANNIHILATION) / sig
    p_energy = stats.kstest(E[un], stats.truncnorm(lo_s, hi_s, loc=E_ANNIHILATION, scale=sig).cdf).pvalue
    assert p_energy > 1e-3, p_energy
    # emission times of a decaying source: density exp(-t/tau) on [0, T)
    tau_ps = cfg.mean_life * 1e12
    norm = -np.expm1(-T_ps / tau_ps)
    p_time = stats.kstest(t.astype(np.float64), lambda x: -np.expm1(-x / tau_ps) / norm).pvalue
    mean_want = tau_ps - T_ps * np.exp(-T_ps / tau_ps) / norm
    sd_want = np.sqrt(tau_ps ** 2 - T_ps ** 2 * np.exp(-T_ps / tau_ps) / norm ** 2)
    z_time = (t.mean() - mean_want) / (sd_want / np.sqrt(len(t)))
    assert p_time > 1e-3 and abs(z_time) < 4.5, (p_time, z_time)
    # mode 0: PSF true pairs and sinogram trues are two acceptances of the same pairs
    n_true = int(true_pair.sum())
    pw = stats.norm.cdf(hi_s) - stats.norm.cdf(lo_s)
    p_pair = pw * pw                                  # both energies of a true pair in the window
    sd = np.sqrt(2.0 * n_true / p_pair * p_pair * (1 - p_pair))
    diff = n_true - int(trues_sum)
    assert abs(diff) <= 5 * sd + 3, (n_true, int(trues_sum), sd)
    return dict(pairs=n // 2, true_pairs=n_true, sino_trues=int(trues_sum), p_energy=p_energy, p_time=p_time,
                z_time=z_time, max_sep_cm=float(sep.max()))


@pytest.mark.parametrize("flavor", ["fmath", "libm"])
def test_oracle_psf_invariants(flavor):
    sc = psf_scenario(4e5, seed=2468)
    o = sc.run_oracle(flavor, nthreads=8)
    r = check_psf(o["psf"], o["psf_requested"], o["trues"].sum(), sc.cfg, sc.phantom.shape, sc.vs)
    _say(f"oracle ({flavor}) PSF: {r}")


@pytest.mark.gpu
def test_gpu_psf_invariants():
    _say("water cube 3e6 decays, tau = T = 1 s, PSF + sinogram (mode 0): PSF invariants on the device output")
    sc = psf_scenario(3e6, seed=97531)
    g = sc.run_gpu(count_events=False)
    r = check_psf(g["psf"], g["psf_requested"], g["trues"].sum(), sc.cfg, sc.phantom.shape, sc.vs)
    _say(f"GPU PSF: {r}")
