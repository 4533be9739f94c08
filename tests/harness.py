"""Shared set-up for the parity tests: one scenario (config + phantom + tables) run through
the library (GPU) and through the CPU oracle on the same inputs."""
import os

import numpy as np

import mcpet
import oracle_py
from mcpet import phantoms
from conftest import AIR, WATER, MATERIALS, names15


class Scenario:
    """e_max > 0: interaction tables continued up to e_max eV (library and oracle alike)."""

    def __init__(self, cfg, mat, dens, act, vs, names, e_max=0.0):
        self.cfg = cfg
        self.names = list(names)
        self.phantom = mcpet.phantom_from_arrays(mat, dens, act, vs)
        self.vs = vs
        self.e_max = e_max
        self.tables = mcpet.load_materials(self.names, self.phantom.density_max(), e_max=e_max)
        self._otab = {}

    def oracle(self, flavor="fmath"):
        if flavor not in self._otab:
            self._otab[flavor] = oracle_py.OracleTables(names15(self.names), self.phantom.density_max(), flavor,
                                                        e_max=self.e_max)
        return oracle_py.Oracle(phantoms.config_dict(self.cfg), self.phantom.voxel_mat_dens(), self.phantom.shape,
                                self.vs, self._otab[flavor], flavor)

    def run_oracle(self, flavor="fmath", nthreads=8, **kw):
        return self.oracle(flavor).run(nthreads=nthreads, **kw)

    def run_gpu(self, kernel=mcpet.KERNEL_AUTO, count_events=True, **kw):
        with mcpet.Simulation(self.cfg, self.phantom, self.tables) as sim:
            sim.run(kernel=kernel, count_events=count_events, **kw)
            out = sim.fetch()
            out["stats"] = sim.stats()
        return out


def water_cube_scenario(n_decays=2e5, tpb=32, seed=12345, **cfgkw):
    mat, dens, act, vs = phantoms.water_cube(n_decays=n_decays)
    cfg = phantoms.make_config(seed=seed, threads_per_block=tpb, materials=[AIR, WATER], **cfgkw)
    return Scenario(cfg, mat, dens, act, vs, [AIR, WATER])


def sample_scenario(acq_s=0.05, seed=4321, tpb=32, **cfgkw):
    """The reference sample simulation (MCGPU-PET.in + phantom_9x9x9cm.vox) with a fixed seed
    and a shorter acquisition."""
    from conftest import SAMPLE
    cfg = mcpet.read_input(os.path.join(SAMPLE, "MCGPU-PET.in"))
    cfg.seed = seed
    cfg.threads_per_block = tpb
    cfg.acquisition_time_ps = int(acq_s * 1e12 + 0.5)
    cfg.detector.psf_size = 2_000_000
    for k, v in cfgkw.items():
        if hasattr(cfg.detector, k):
            setattr(cfg.detector, k, v)
        else:
            setattr(cfg, k, v)
    ph = mcpet.load_voxels(os.path.join(SAMPLE, "phantom_9x9x9cm.vox"), cfg)
    vmd = ph.voxel_mat_dens()
    mat = (vmd[:, 0].astype(np.int32)).reshape(ph.shape)
    names = [AIR, WATER]
    for i, n in enumerate(names):
        cfg.material_files[i].value = os.fsencode(n)
    return Scenario(cfg, mat, vmd[:, 1].reshape(ph.shape), vmd[:, 2].reshape(ph.shape), (1.0, 1.0, 1.0), names)


def head_scenario(shape=(24, 32, 32), voxel_cm=0.8, n_decays=1e5, tpb=32, seed=777, **cfgkw):
    mat, dens, act, vs = phantoms.head(shape=shape, voxel_cm=voxel_cm, n_decays=n_decays)
    names = phantoms.head_materials(MATERIALS)
    cfg = phantoms.make_config(seed=seed, threads_per_block=tpb, acquisition_time_s=10.0, nrad=168, materials=names,
                               **cfgkw)
    return Scenario(cfg, mat, dens, act, vs, names)


def psf_sorted(p):
    """PSF records as an order-independent multiset (rows of 11 uint32 words, sorted)."""
    raw = np.ascontiguousarray(p).view(np.uint32).reshape(len(p), 11)
    order = np.lexsort(raw.T[::-1])
    return raw[order]


def assert_tallies_equal(g, o, counters="all"):
    for k in ("trues", "scatter", "image_trues", "image_scatter", "spectrum"):
        assert np.array_equal(g[k], o[k]), f"{k} differs: gpu sum {g[k].sum()} oracle sum {o[k].sum()}"
    assert int(g["total_histories"]) == int(o["total_histories"])
    assert np.array_equal(g["material_edep"], o["material_edep"])
    if "voxel_edep" in o:
        assert np.array_equal(g["voxel_edep"], o["voxel_edep"])
    if "psf" in o:
        assert g["psf_requested"] == o["psf_requested"]
        assert np.array_equal(psf_sorted(g["psf"]), psf_sorted(o["psf"]))
    assert g["final_seed"] == o["final_seed"]
    lite = ("decays", "streams", "pairs_tallied", "coincidences")
    for k, v in o["counters"].items():
        if counters == "all" or k in lite:
            assert g["stats"][k] == v, f"counter {k}: gpu {g['stats'][k]} oracle {v}"
