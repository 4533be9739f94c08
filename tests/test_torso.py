"""Torso workload: 256 x 256 x 200 voxels at 2.5 mm (lungs at 0.3 g/cm3, spine and ribs, liver,
two lesions), a positron emitter with 50 % beta+ branching and a prompt gamma with every decay,
5e9 expected decays in 10 s, a 4.5 ns coincidence window, 64-bit fixed-point dose tallies over
every voxel, a 46 cm detector radius (the 64 x 64 cm box has a 45.3 cm half-diagonal).

The prompt gamma is emitted at 603 keV. The only material files (the reference's air and water)
end at 515 keV, so the interaction tables are continued to 610 keV (mcpet_load_materials_extended:
Compton by Klein-Nishina scaling, Rayleigh and photoelectric by power laws, pinned in
test_inputs.py); the oracle builds the same continued tables.

GPU tier: a 2 ms slice of the acquisition at the full count rate (1e6 decays, so singles
overlap and the sorter finds randoms and spurious coincidences), the whole torso, bit-exact
against the oracle (singles list, coincidence sinograms, spectrum, voxel and material doses,
counters); the full 10 s acquisition (5e9 decays) in the reference's fused pair tallies,
checked through size-independent properties (decay and branching counts, images against
sinograms, voxel doses against material doses)."""
import os

import numpy as np
import pytest

import mcpet
from conftest import WATER
from harness import Scenario
from mcpet import phantoms
from test_singles import _check_against_oracle

PROMPT_E = 603.0e3
E_MAX = 610.0e3   # interaction tables continued past the files' 515 keV
WINDOW = 4500     # ps
SHAPE = (200, 256, 256)


def torso_scenario(acq_s, seed=606, **kw):
    mat, dens, act, vs = phantoms.torso()          # activity: 5e9 expected decays in 10 s
    names = phantoms.torso_materials(os.path.dirname(WATER))
    frac = phantoms.F18_MEAN_LIFE * (1.0 - np.exp(-acq_s / phantoms.F18_MEAN_LIFE))
    n = float(act.sum(dtype=np.float64)) * frac
    cfg = phantoms.make_config(seed=seed, threads_per_block=1, acquisition_time_s=acq_s, radius=-46.0, height=25.0,
                               fovz=25.0, materials=names, positron_fraction=0.5, prompt_gamma_fraction=1.0,
                               prompt_gamma_energy=PROMPT_E, material_dose=True, voxel_dose=True,
                               roi=(0, SHAPE[2] - 1, 0, SHAPE[1] - 1, 0, SHAPE[0] - 1),
                               stream_decays=int(max(4, np.ceil(n / 2.0 ** 20))), **kw)
    return Scenario(cfg, mat, dens, act, vs, names, e_max=E_MAX), n


def test_torso_phantom_shape_and_tissues():
    mat, dens, act, vs = phantoms.torso()
    assert mat.shape == SHAPE and vs == (0.25, 0.25, 0.25)
    counts = np.bincount(mat.ravel(), minlength=8)
    assert np.all(counts[1:] > 0)                                # every tissue present
    assert dens[mat == 5].max() == np.float32(0.30) and dens[mat == 4].max() == np.float32(1.85)
    frac = phantoms.F18_MEAN_LIFE * (1.0 - np.exp(-10.0 / phantoms.F18_MEAN_LIFE))
    assert abs(act.sum(dtype=np.float64) * frac / 5e9 - 1) < 1e-5
    assert act[mat == 1].max() == 0 and act[mat == 7].min() > act[mat == 6].max() > act[mat == 2].max()
    # the body fits the 46 cm detector radius around the box centre
    zz, yy, xx = np.nonzero(mat > 1)
    r = np.hypot((xx + 0.5) * 0.25 - 32.0, (yy + 0.5) * 0.25 - 32.0)
    assert r.max() < 46.0


def test_torso_phantom_is_seeded():
    a = phantoms.torso(seed=3, shape=(40, 64, 64), voxel_cm=1.0)
    b = phantoms.torso(seed=3, shape=(40, 64, 64), voxel_cm=1.0)
    c = phantoms.torso(seed=4, shape=(40, 64, 64), voxel_cm=1.0)
    assert all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    assert not np.array_equal(a[0], c[0])


@pytest.mark.gpu
def test_torso_slice_at_full_rate_bit_exact():
    sc, n = torso_scenario(0.002, tally_psf_sinogram=3, coincidence_window_ps=WINDOW, singles_capacity=4_000_000)
    assert abs(n / 1e6 - 1) < 1e-3
    keep = {}
    g, counts, raw = _check_against_oracle(sc, nthreads=16, keep=keep)
    assert np.array_equal(g["voxel_edep"], keep["oracle"]["voxel_edep"])
    assert int(g["voxel_edep"][:, 0].sum()) == int(g["material_edep"][:, 0].sum())
    assert int(g["voxel_edep"][:, 1].sum()) == int(g["material_edep"][:, 1].sum())
    st = g["stats"]
    d = st["decays"]
    assert abs(d - n) < 5 * np.sqrt(n)
    assert st["prompts_emitted"] == d                            # a prompt gamma with every decay
    assert abs(st["pairs_emitted"] - 0.5 * d) < 5 * np.sqrt(0.25 * d)
    assert counts[0] > 100 and counts[1] > 100 and counts[2] > 0 and counts[3] > 0, counts
    kinds = np.bincount(g["singles"]["kind"], minlength=3)
    assert kinds[2] > 0 and kinds[0] > 0 and kinds[1] > 0


@pytest.mark.gpu
def test_torso_full_acquisition_properties():
    sc, n = torso_scenario(10.0)
    assert abs(n / 5e9 - 1) < 1e-5
    g = sc.run_gpu(kernel=mcpet.KERNEL_PERSISTENT)      # with the event-counting replay (same seeds)
    st = g["stats"]
    d = st["decays"]
    assert abs(d - n) < 5 * np.sqrt(n)
    assert st["prompts_emitted"] == d
    assert abs(st["pairs_emitted"] - 0.5 * d) < 5 * np.sqrt(0.25 * d)
    assert int(g["total_histories"]) == 2 * d
    assert g["trues"].sum() == g["image_trues"].sum() > 1e7
    assert g["scatter"].sum() == g["image_scatter"].sum() > 1e6
    assert st["coincidences"] == g["trues"].sum() + g["scatter"].sum()
    assert g["spectrum"].sum() <= 2 * st["pairs_tallied"]          # both photons of every accepted pair
    assert int(g["voxel_edep"][:, 0].sum()) == int(g["material_edep"][:, 0].sum()) > 0
    assert int(g["voxel_edep"][:, 1].sum()) == int(g["material_edep"][:, 1].sum())
    # no energy in a material slot the phantom does not use; no emission from air
    assert np.all(g["material_edep"][7:] == 0)
    mat = phantoms.torso()[0]
    assert g["image_trues"].reshape(SHAPE)[mat == 1].sum() == 0 and g["image_scatter"].reshape(SHAPE)[mat == 1].sum() == 0
