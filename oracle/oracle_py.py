"""ctypes binding of the CPU oracle (pet_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product. Two builds of the same C source:
  "fmath" -- the library's deterministic elementary functions (bit-exact comparisons);
  "libm"  -- the C library's logf/expf/... (an independent arithmetic, statistical checks).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXMAT = 15
NP_RAY = 128
MAXBINS_RAY = 25005
MAXSH = 30

COUNTER_NAMES = ["decays", "streams", "woodcock_steps", "material_changes", "compton", "compton_shell_evals",
                 "rayleigh", "photoelectric", "pairs_tallied", "coincidences", "rn_draws", "pairs_emitted",
                 "prompts_emitted"]


class OrcRayleigh(C.Structure):
    _fields_ = [("xco", C.c_float * (NP_RAY * MAXMAT)), ("pco", C.c_float * (NP_RAY * MAXMAT)),
                ("aco", C.c_float * (NP_RAY * MAXMAT)), ("bco", C.c_float * (NP_RAY * MAXMAT)),
                ("pmax", C.c_float * (MAXBINS_RAY * MAXMAT)), ("itlco", C.c_uint8 * (NP_RAY * MAXMAT)),
                ("ituco", C.c_uint8 * (NP_RAY * MAXMAT))]


class OrcCompton(C.Structure):
    _fields_ = [("fco", C.c_float * (MAXMAT * MAXSH)), ("uico", C.c_float * (MAXMAT * MAXSH)),
                ("fj0", C.c_float * (MAXMAT * MAXSH)), ("noscco", C.c_int * MAXMAT)]


class OrcSim(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("inv_voxel_size", C.c_float * 3), ("size_bbox", C.c_float * 3),
        ("voxel_mat_dens", C.POINTER(C.c_float)),
        ("num_values", C.c_int), ("e0", C.c_float), ("ide", C.c_float), ("woodcock", C.POINTER(C.c_float)),
        ("mfp_a", C.POINTER(C.c_float)), ("mfp_b", C.POINTER(C.c_float)), ("rayleigh", C.POINTER(OrcRayleigh)),
        ("compton", C.POINTER(OrcCompton)),
        ("acquisition_time_ps", C.c_uint64), ("mean_life", C.c_float), ("threads_per_block", C.c_int), ("seed", C.c_int),
        ("bg_mode", C.c_int),
        ("det_center", C.c_float * 3), ("det_radius", C.c_float), ("det_height", C.c_float), ("psf_size", C.c_int),
        ("tally_psf_sinogram", C.c_int), ("tally_type", C.c_int),
        ("e_resol", C.c_float), ("e_low", C.c_float), ("e_high", C.c_float), ("fovz", C.c_float),
        ("nrows", C.c_int), ("ncrystals", C.c_int), ("nangles", C.c_int), ("nrad", C.c_int), ("nzs", C.c_int),
        ("nbins", C.c_int), ("ne", C.c_int), ("mrd", C.c_int), ("span", C.c_int), ("nsinos", C.c_int),
        ("material_dose", C.c_int), ("roi", C.c_short * 6),
        ("trues", C.POINTER(C.c_int)), ("scatter", C.POINTER(C.c_int)), ("image_trues", C.POINTER(C.c_int)),
        ("image_scatter", C.POINTER(C.c_int)), ("spectrum", C.POINTER(C.c_int)), ("psf", C.c_void_p),
        ("voxel_edep", C.POINTER(C.c_uint64)), ("material_edep", C.POINTER(C.c_uint64)),
        ("psf_index", C.c_int), ("total_histories", C.c_uint64), ("final_seed", C.c_int), ("final_seed_written", C.c_int),
        ("counters", C.c_uint64 * 16), ("nthreads", C.c_int), ("voxel_stride", C.c_int), ("voxel_offset", C.c_int),
        ("wall_seconds", C.c_double), ("positron_fraction", C.c_float), ("prompt_fraction", C.c_float),
        ("prompt_energy", C.c_float), ("singles", C.c_void_p), ("singles_cap", C.c_longlong), ("singles_n", C.c_longlong),
        ("stream_decays", C.c_int), ("stream_first", C.c_void_p)]


SINGLE_DTYPE = np.dtype([("time_ps", "<u8"), ("event", "<u8"), ("energy", "<f4"), ("z", "<f4"), ("phi", "<f4"),
                         ("scatter", "<i2"), ("kind", "<i2")])

_libs = {}


def load(flavor="fmath"):
    if flavor in _libs:
        return _libs[flavor]
    path = os.path.join(HERE, f"liboracle_{flavor}.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (python mcgpu-pet_amd/build.py)")
    L = C.CDLL(path)
    P = C.POINTER
    L.orc_abMODm.argtypes = [C.c_int, C.c_int, C.c_int]
    L.orc_abMODm.restype = C.c_int
    L.orc_init_PRNG.argtypes = [C.c_longlong, C.c_int, C.c_int, P(C.c_int), P(C.c_int)]
    L.orc_init_PRNG.restype = None
    L.orc_ranecu.argtypes = [P(C.c_int), P(C.c_int)]
    L.orc_ranecu.restype = C.c_float
    L.orc_ranecu_double.argtypes = [P(C.c_int), P(C.c_int)]
    L.orc_ranecu_double.restype = C.c_double
    L.orc_update_seed_PRNG.argtypes = [C.c_int, C.c_ulonglong, P(C.c_int)]
    L.orc_update_seed_PRNG.restype = None
    L.orc_ranecu_block.argtypes = [C.c_int, P(C.c_longlong), C.c_int, C.c_int, P(C.c_float), P(C.c_double), C.c_int]
    L.orc_ranecu_block.restype = None
    L.orc_sinogram_sizes.argtypes = [C.c_int] * 5 + [P(C.c_int)] * 3
    L.orc_sinogram_sizes.restype = None
    L.orc_load_material.argtypes = [C.c_char_p, P(C.c_float), P(C.c_float), P(C.c_int), P(C.c_float), P(C.c_float),
                                    P(P(C.c_float)), P(P(C.c_float)), P(P(C.c_float)), P(OrcRayleigh), P(OrcCompton)]
    L.orc_load_material.restype = C.c_int
    L.orc_load_material_ext.argtypes = [C.c_char_p, P(C.c_float), C.c_double, P(C.c_float), P(C.c_int), P(C.c_float),
                                        P(C.c_float), P(P(C.c_float)), P(P(C.c_float)), P(P(C.c_float)), P(OrcRayleigh),
                                        P(OrcCompton)]
    L.orc_load_material_ext.restype = C.c_int
    L.orc_free.argtypes = [C.c_void_p]
    L.orc_free.restype = None
    L.orc_run.argtypes = [P(OrcSim)]
    L.orc_run.restype = C.c_int
    L.orc_locate_voxel.argtypes = [P(OrcSim), C.c_float, C.c_float, C.c_float]
    L.orc_locate_voxel.restype = C.c_int
    L.orc_probe_rotate.argtypes = [C.c_int, P(C.c_float), P(C.c_double), P(C.c_double), P(C.c_float)]
    L.orc_probe_rotate.restype = None
    L.orc_probe_gcoa.argtypes = [P(OrcSim), C.c_int, C.c_longlong, C.c_int, C.c_float, C.c_int, P(C.c_float), P(C.c_double)]
    L.orc_probe_gcoa.restype = None
    L.orc_probe_graa.argtypes = [P(OrcSim), C.c_int, C.c_longlong, C.c_int, C.c_float, C.c_int, P(C.c_double)]
    L.orc_probe_graa.restype = None
    L.orc_probe_source.argtypes = [P(OrcSim), C.c_int, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_float),
                                   P(C.c_float), P(C.c_float)]
    L.orc_probe_source.restype = None
    L.orc_probe_first_interaction.argtypes = [P(OrcSim), C.c_int, C.c_longlong, C.c_int, P(C.c_float), P(C.c_float),
                                              C.c_float, P(C.c_float), P(C.c_int)]
    L.orc_probe_first_interaction.restype = None
    L.orc_sinogram_bin.argtypes = [P(OrcSim), C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_sinogram_bin.restype = C.c_int
    L.orc_detector_bin.argtypes = [P(OrcSim), P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float)]
    L.orc_detector_bin.restype = C.c_int
    if flavor == "fmath":
        L.orc_fmath_eval.argtypes = [P(C.c_double), P(C.c_double), C.c_int, P(C.c_double), P(C.c_float)]
        L.orc_fmath_eval.restype = None
    _libs[flavor] = L
    return L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class OracleTables:
    """Interaction tables built by the oracle's own restatement of load_material (C:1565-1831)."""

    def __init__(self, names, density_max, flavor="fmath", e_max=0.0):
        L = load(flavor)
        buf = C.create_string_buffer(MAXMAT * 250)
        for i, n in enumerate(list(names)[:MAXMAT]):
            b = os.fsencode(n) if n else b""
            C.memmove(C.addressof(buf) + 250 * i, b, min(len(b), 249))
        dm = np.ascontiguousarray(density_max, dtype=np.float32)
        self.density_nominal = np.zeros(MAXMAT, np.float32)
        n = C.c_int()
        e0 = C.c_float()
        ide = C.c_float()
        w = C.POINTER(C.c_float)()
        a = C.POINTER(C.c_float)()
        b = C.POINTER(C.c_float)()
        self.ray = OrcRayleigh()
        self.cmp = OrcCompton()
        rc = L.orc_load_material_ext(buf, _f(dm), float(e_max), _f(self.density_nominal), C.byref(n), C.byref(e0),
                                     C.byref(ide), C.byref(w), C.byref(a), C.byref(b), C.byref(self.ray), C.byref(self.cmp))
        if rc != 0:
            raise RuntimeError(f"orc_load_material failed ({rc})")
        self.num_values = n.value
        self.e0 = e0.value
        self.ide = ide.value
        N = self.num_values
        self.woodcock = np.ctypeslib.as_array(w, shape=(N, 2)).copy()
        self.mfp_a = np.ctypeslib.as_array(a, shape=(N, MAXMAT, 3)).copy()
        self.mfp_b = np.ctypeslib.as_array(b, shape=(N, MAXMAT, 3)).copy()
        L.orc_free(C.cast(w, C.c_void_p))
        L.orc_free(C.cast(a, C.c_void_p))
        L.orc_free(C.cast(b, C.c_void_p))


class Oracle:
    """One simulation set-up (phantom + tables + parameters) for the CPU restatement.

    params: dict with the mcpet_config-equivalent fields (see config_dict in tests)."""

    def __init__(self, params, voxel_mat_dens, shape, voxel_size, tables, flavor="fmath"):
        self.L = load(flavor)
        self.flavor = flavor
        self.p = dict(params)
        self.vmd = np.ascontiguousarray(voxel_mat_dens, dtype=np.float32).reshape(-1, 3)
        self.nz, self.ny, self.nx = shape
        self.tables = tables
        s = OrcSim()
        s.nx, s.ny, s.nz = self.nx, self.ny, self.nz
        vs = np.asarray(voxel_size, np.float32)
        ivs = (np.float32(1.0) / vs).astype(np.float32)
        bbox = (np.array([self.nx, self.ny, self.nz], np.float32) * vs).astype(np.float32)
        s.inv_voxel_size[:] = [float(v) for v in ivs]
        s.size_bbox[:] = [float(v) for v in bbox]
        s.voxel_mat_dens = _f(self.vmd)
        s.num_values = tables.num_values
        s.e0 = tables.e0
        s.ide = tables.ide
        self._w = np.ascontiguousarray(tables.woodcock, np.float32)
        self._a = np.ascontiguousarray(tables.mfp_a, np.float32)
        self._b = np.ascontiguousarray(tables.mfp_b, np.float32)
        s.woodcock, s.mfp_a, s.mfp_b = _f(self._w), _f(self._a), _f(self._b)
        s.rayleigh = C.pointer(tables.ray)
        s.compton = C.pointer(tables.cmp)
        p = self.p
        s.acquisition_time_ps = int(p["acquisition_time_ps"])
        s.mean_life = p["mean_life"]
        s.threads_per_block = p["threads_per_block"]
        s.seed = p["seed"]
        s.bg_mode = p.get("bg_mode", 0)
        c = list(p["det_center"])
        radius = p["det_radius"]
        if radius < 0:   # C:316-322
            c = [float(np.float32(bbox[i]) * np.float32(0.5)) for i in range(3)]
            radius = -radius
        s.det_center[:] = c
        s.det_radius = radius
        s.det_height = p["det_height"]
        s.psf_size = p.get("psf_size", 0)
        s.tally_psf_sinogram = p.get("tally_psf_sinogram", 2)
        s.tally_type = p.get("tally_type", 0)
        for k in ("e_resol", "e_low", "e_high", "fovz", "nrows", "ncrystals", "nangles", "nrad", "nzs", "nbins", "ne",
                  "mrd", "span", "nsinos"):
            setattr(s, k, p[k])
        s.material_dose = p.get("material_dose", 0)
        roi = p.get("roi", (32500, -32500, 32500, -32500, 32500, -32500))
        s.roi[:] = list(roi)
        pf = p.get("positron_fraction", 0.0)
        s.positron_fraction = pf if 0.0 < pf < 1.0 else 1.0
        s.prompt_fraction = p.get("prompt_gamma_fraction", 0.0)
        s.prompt_energy = p.get("prompt_gamma_energy", 0.0)
        s.singles_cap = int(p.get("singles_capacity", 0))
        s.stream_decays = int(p.get("stream_decays", 0))
        self.s = s

    def locate(self, x, y, z):
        return self.L.orc_locate_voxel(C.byref(self.s), x, y, z)

    def run(self, nthreads=1, voxel_stride=1, voxel_offset=0, psf=True):
        s = self.s
        p = self.p
        nvox = self.nx * self.ny * self.nz
        out = {"trues": np.zeros(p["nbins"], np.int32), "scatter": np.zeros(p["nbins"], np.int32),
               "image_trues": np.zeros(nvox, np.int32), "image_scatter": np.zeros(nvox, np.int32),
               "spectrum": np.zeros(max(1, p["ne"]), np.int32), "material_edep": np.zeros((MAXMAT, 2), np.uint64)}
        s.trues = out["trues"].ctypes.data_as(C.POINTER(C.c_int))
        s.scatter = out["scatter"].ctypes.data_as(C.POINTER(C.c_int))
        s.image_trues = out["image_trues"].ctypes.data_as(C.POINTER(C.c_int))
        s.image_scatter = out["image_scatter"].ctypes.data_as(C.POINTER(C.c_int))
        s.spectrum = out["spectrum"].ctypes.data_as(C.POINTER(C.c_int))
        s.material_edep = out["material_edep"].ctypes.data_as(C.POINTER(C.c_uint64))
        psf_buf = None
        if psf and s.tally_psf_sinogram in (0, 1) and s.psf_size > 0:
            psf_buf = np.zeros(s.psf_size * 44, np.uint8)
            s.psf = psf_buf.ctypes.data
        else:
            s.psf = None
        roi = list(s.roi)
        if roi[1] > -1:
            n_roi = (roi[1] - roi[0] + 1) * (roi[3] - roi[2] + 1) * (roi[5] - roi[4] + 1)
            out["voxel_edep"] = np.zeros((n_roi, 2), np.uint64)
            s.voxel_edep = out["voxel_edep"].ctypes.data_as(C.POINTER(C.c_uint64))
        else:
            s.voxel_edep = None
        singles_buf = None
        if s.tally_psf_sinogram == 3 and s.singles_cap > 0:
            singles_buf = np.zeros(s.singles_cap * 32, np.uint8)
            s.singles = singles_buf.ctypes.data
        else:
            s.singles = None
        s.nthreads = nthreads
        s.voxel_stride = voxel_stride
        s.voxel_offset = voxel_offset
        rc = self.L.orc_run(C.byref(s))
        if rc != 0:
            raise RuntimeError("orc_run failed")
        out["total_histories"] = s.total_histories
        out["psf_requested"] = s.psf_index
        if psf_buf is not None:
            n = min(s.psf_index, s.psf_size)
            out["psf"] = psf_buf[:n * 44].view(_psf_dtype()).copy()
        if singles_buf is not None:
            n = min(s.singles_n, s.singles_cap)
            out["singles"] = singles_buf[:n * 32].view(SINGLE_DTYPE).copy()
            out["singles_requested"] = s.singles_n
        out["final_seed"] = s.final_seed if s.final_seed_written else 0
        out["counters"] = {k: int(s.counters[i]) for i, k in enumerate(COUNTER_NAMES)}
        out["wall_seconds"] = s.wall_seconds
        return out

    def sinogram_bin(self, ix1, iz1, ix2, iz2):
        return self.L.orc_sinogram_bin(C.byref(self.s), ix1, iz1, ix2, iz2)

    def probe_gcoa(self, seed, stream, n, energy, mat):
        e = np.zeros(n, np.float32)
        c = np.zeros(n, np.float64)
        self.L.orc_probe_gcoa(C.byref(self.s), seed, stream, n, energy, mat, _f(e), c.ctypes.data_as(C.POINTER(C.c_double)))
        return e, c

    def probe_graa(self, seed, stream, n, energy, mat):
        c = np.zeros(n, np.float64)
        self.L.orc_probe_graa(C.byref(self.s), seed, stream, n, energy, mat, c.ctypes.data_as(C.POINTER(C.c_double)))
        return c

    def probe_source(self, seed, stream, n, ix, iy, iz):
        p1 = np.zeros((n, 3), np.float32)
        d1 = np.zeros((n, 3), np.float32)
        d2 = np.zeros((n, 3), np.float32)
        self.L.orc_probe_source(C.byref(self.s), seed, stream, n, ix, iy, iz, _f(p1), _f(d1), _f(d2))
        return p1, d1, d2

    def probe_first_interaction(self, seed, stream, pos, dirs, energy):
        pos = np.ascontiguousarray(pos, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(pos)
        dist = np.zeros(n, np.float32)
        typ = np.zeros(n, np.int32)
        self.L.orc_probe_first_interaction(C.byref(self.s), seed, stream, n, _f(pos), _f(dirs), energy, _f(dist),
                                           typ.ctypes.data_as(C.POINTER(C.c_int)))
        return dist, typ

    def detector_bin(self, pos, d, posf, df):
        a = [np.ascontiguousarray(v, np.float32) for v in (pos, d, posf, df)]
        return self.L.orc_detector_bin(C.byref(self.s), *[_f(v) for v in a])


def probe_rotate(dirs, costh, phi, flavor="fmath"):
    L = load(flavor)
    dirs = np.ascontiguousarray(dirs, np.float32)
    costh = np.ascontiguousarray(costh, np.float64)
    phi = np.ascontiguousarray(phi, np.float64)
    out = np.zeros_like(dirs)
    L.orc_probe_rotate(len(dirs), _f(dirs), costh.ctypes.data_as(C.POINTER(C.c_double)),
                       phi.ctypes.data_as(C.POINTER(C.c_double)), _f(out))
    return out


def ranecu_block(seed, ids, n, as_double=False, flavor="fmath"):
    L = load(flavor)
    ids = np.ascontiguousarray(ids, np.int64)
    if as_double:
        out = np.zeros((len(ids), n), np.float64)
        L.orc_ranecu_block(seed, ids.ctypes.data_as(C.POINTER(C.c_longlong)), len(ids), n, None,
                           out.ctypes.data_as(C.POINTER(C.c_double)), 1)
    else:
        out = np.zeros((len(ids), n), np.float32)
        L.orc_ranecu_block(seed, ids.ctypes.data_as(C.POINTER(C.c_longlong)), len(ids), n, _f(out), None, 0)
    return out


def _psf_dtype():
    return np.dtype([("emission_time_ps", "<u8"), ("travel_time_ps", "<f4"), ("emission_absvox", "<i4"),
                     ("energy", "<f4"), ("z", "<f4"), ("phi", "<f4"), ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4"),
                     ("index1", "<i2"), ("index2", "<i2")])


def fmath_eval(x, y):
    """Host (gcc) evaluation of the library's elementary functions (fmath build only)."""
    L = load("fmath")
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = len(x)
    od = np.zeros((7, n), np.float64)
    of = np.zeros((8, n), np.float32)
    L.orc_fmath_eval(x.ctypes.data_as(C.POINTER(C.c_double)), y.ctypes.data_as(C.POINTER(C.c_double)), n,
                     od.ctypes.data_as(C.POINTER(C.c_double)), _f(of))
    return od, of


def pair_singles(singles, window_ps, geom, oracle=None):
    """Coincidence pairing of a singles list (the CPU statement of the library's pairing):
    singles ordered by (time, event, kind); a coincidence is a cluster of exactly two singles
    with consecutive gap <= window and gaps > window on both sides; clusters of three or more
    are multiples. Classes: 0 true, 1 scattered, 2 random, 3 spurious (prompt gamma of the
    same decay). Pairs are binned with the reference's Michelogram arithmetic (earlier single
    first) through the oracle's binning, crystal indices as in tally_Sinogram (K:1749-1752).
    geom: dict with ncrystals, nrows, fovz, nbins. Returns (sinograms[4, nbins], binned counts
    per class + multiples [5], unbinned counts per class + multiples [5]); with oracle=None
    nothing is binned."""
    s = np.sort(np.asarray(singles), order=["time_ps", "event", "kind"], kind="stable")
    t = s["time_ps"].astype(np.uint64)
    n = len(s)
    sinos = np.zeros((4, geom["nbins"]), np.int64)
    counts = np.zeros(5, np.int64)
    raw = np.zeros(5, np.int64)
    if n < 2:
        return sinos, counts, raw
    W = np.uint64(window_ps)
    gap = t[1:] - t[:-1]                                    # gap[i] = t[i+1] - t[i]
    idx = np.arange(n - 1)
    start = gap <= W
    start &= np.concatenate([[True], gap[:-1] > W])         # left gap > W (or first)
    multiple = start & np.concatenate([gap[1:] <= W, [False]])
    counts[4] = int(multiple.sum())
    pairs = idx[start & ~multiple]
    a, b = s[pairs], s[pairs + 1]
    same = a["event"] == b["event"]
    prompt = (a["kind"] == 2) | (b["kind"] == 2)
    unsc = (a["scatter"] == 0) & (b["scatter"] == 0)
    cls = np.where(~same, 2, np.where(prompt, 3, np.where(unsc, 0, 1)))
    raw[:4] = np.bincount(cls, minlength=4)[:4]
    raw[4] = counts[4]
    if oracle is None:
        return sinos, counts, raw

    def crystal(phi):
        return np.floor(((phi.astype(np.float64) / (2.0 * np.pi) + 0.5) * geom["ncrystals"]).astype(np.float32)).astype(np.int64)

    def ring(z):
        q = (z.astype(np.float32) / np.float32(geom["fovz"])).astype(np.float64)
        return np.floor(((q + 0.5) * geom["nrows"]).astype(np.float32)).astype(np.int64)

    ix1, iz1, ix2, iz2 = crystal(a["phi"]), ring(a["z"]), crystal(b["phi"]), ring(b["z"])
    for k in range(len(pairs)):
        ibin = oracle.sinogram_bin(int(ix1[k]), int(iz1[k]), int(ix2[k]), int(iz2[k]))
        if ibin >= 0:
            sinos[cls[k], ibin] += 1
            counts[cls[k]] += 1
    return sinos, counts, raw
