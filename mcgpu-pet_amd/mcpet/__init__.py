"""Python mirror of the libmcpet C ABI (include/mcpet.h), via ctypes.

The product is the shared library libmcpet.so (gfx950 HIP kernels + host C++); this
module only marshals arguments. It never computes tallies itself and raises if the
library is missing: there is no CPU fallback.

Reference correspondence (MCGPU-PET): read_input -> read_input (MCGPU-PET.cu:866),
load_voxels -> load_voxels (C:1381), load_materials -> load_material (C:1565),
Simulation.run -> track_particles launch (C:452-487), Simulation.fetch -> result copies
(C:499-535), write_outputs -> output writers (C:544-616, 3193-3322, 2389-2650).
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("MCPET_LIB") or os.path.join(_PKG, "lib", "libmcpet.so")

MAX_MATERIALS = 15
MAX_SHELLS = 30
NPHASES = 10
NP_RAYLEIGH = 128
MAX_ENERGYBINS_RAYLEIGH = 25005
NAME_LEN = 250

KERNEL_AUTO, KERNEL_PLAIN, KERNEL_PERSISTENT = 0, 1, 2

STATUS = {0: "OK", 1: "E_ARG", 2: "E_IO", 3: "E_FORMAT", 4: "E_RANGE", 5: "E_NOMEM", 6: "E_HIP", 7: "E_STATE",
          8: "E_CAPACITY"}


class McpetError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Detector(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("height", C.c_float), ("psf_size", C.c_int32),
                ("tally_psf_sinogram", C.c_int32), ("tally_type", C.c_int32)]


class PsfRecord(C.Structure):
    _pack_ = 4
    _fields_ = [("emission_time_ps", C.c_uint64), ("travel_time_ps", C.c_float), ("emission_absvox", C.c_int32),
                ("energy", C.c_float), ("z", C.c_float), ("phi", C.c_float), ("vx", C.c_float), ("vy", C.c_float),
                ("vz", C.c_float), ("index1", C.c_int16), ("index2", C.c_int16)]


# numpy view of the 44-byte PSF record (MCGPU-PET.h:235-247)
PSF_DTYPE = np.dtype([("emission_time_ps", "<u8"), ("travel_time_ps", "<f4"), ("emission_absvox", "<i4"),
                      ("energy", "<f4"), ("z", "<f4"), ("phi", "<f4"), ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4"),
                      ("index1", "<i2"), ("index2", "<i2")])
assert PSF_DTYPE.itemsize == 44


class RayleighTable(C.Structure):
    _fields_ = [("xco", C.c_float * (NP_RAYLEIGH * MAX_MATERIALS)), ("pco", C.c_float * (NP_RAYLEIGH * MAX_MATERIALS)),
                ("aco", C.c_float * (NP_RAYLEIGH * MAX_MATERIALS)), ("bco", C.c_float * (NP_RAYLEIGH * MAX_MATERIALS)),
                ("pmax", C.c_float * (MAX_ENERGYBINS_RAYLEIGH * MAX_MATERIALS)),
                ("itlco", C.c_uint8 * (NP_RAYLEIGH * MAX_MATERIALS)), ("ituco", C.c_uint8 * (NP_RAYLEIGH * MAX_MATERIALS))]


class ComptonTable(C.Structure):
    _fields_ = [("fco", C.c_float * (MAX_MATERIALS * MAX_SHELLS)), ("uico", C.c_float * (MAX_MATERIALS * MAX_SHELLS)),
                ("fj0", C.c_float * (MAX_MATERIALS * MAX_SHELLS)), ("noscco", C.c_int32 * MAX_MATERIALS)]


class Config(C.Structure):
    _fields_ = [("seed", C.c_int32), ("gpu_id", C.c_int32), ("threads_per_block", C.c_int32), ("fact_object", C.c_float),
                ("acquisition_time_ps", C.c_uint64), ("mean_life", C.c_float), ("activity", C.c_float * MAX_MATERIALS),
                ("psf_file", C.c_char * NAME_LEN), ("detector", Detector), ("material_dose", C.c_int32),
                ("voxel_dose", C.c_int32), ("dose_file", C.c_char * NAME_LEN), ("roi_x_min", C.c_int16),
                ("roi_x_max", C.c_int16), ("roi_y_min", C.c_int16), ("roi_y_max", C.c_int16), ("roi_z_min", C.c_int16),
                ("roi_z_max", C.c_int16), ("e_resol", C.c_float), ("e_low", C.c_float), ("e_high", C.c_float),
                ("fovz", C.c_float), ("nrows", C.c_int32), ("ncrystals", C.c_int32), ("nangles", C.c_int32),
                ("nrad", C.c_int32), ("nzs", C.c_int32), ("res", C.c_int32), ("ne", C.c_int32), ("mrd", C.c_int32),
                ("span", C.c_int32), ("nseg", C.c_int32), ("nsinos", C.c_int32), ("nbins", C.c_int32),
                ("nvoxs", C.c_int32), ("voxel_file", C.c_char * NAME_LEN),
                ("material_files", (C.c_char * NAME_LEN) * MAX_MATERIALS), ("bg_mode", C.c_int32), ("voxel_format", C.c_int32),
                ("positron_fraction", C.c_float), ("prompt_gamma_fraction", C.c_float),
                ("prompt_gamma_energy", C.c_float), ("stream_decays", C.c_int32), ("singles_capacity", C.c_int64),
                ("coincidence_window_ps", C.c_uint64)]


class Phantom(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("voxel_size", C.c_float * 3),
                ("inv_voxel_size", C.c_float * 3), ("size_bbox", C.c_float * 3),
                ("density_max", C.c_float * MAX_MATERIALS), ("voxel_mat_dens", C.POINTER(C.c_float)),
                ("owned", C.c_int32)]


class Tables(C.Structure):
    _fields_ = [("num_values", C.c_int32), ("e0", C.c_float), ("ide", C.c_float),
                ("density_nominal", C.c_float * MAX_MATERIALS), ("woodcock", C.POINTER(C.c_float)),
                ("mfp_a", C.POINTER(C.c_float)), ("mfp_b", C.POINTER(C.c_float)),
                ("rayleigh", C.POINTER(RayleighTable)), ("compton", C.POINTER(ComptonTable)), ("owned", C.c_int32),
                ("num_file_values", C.c_int32)]


class Tallies(C.Structure):
    _fields_ = [("trues", C.POINTER(C.c_int32)), ("scatter", C.POINTER(C.c_int32)),
                ("image_trues", C.POINTER(C.c_int32)), ("image_scatter", C.POINTER(C.c_int32)),
                ("spectrum", C.POINTER(C.c_int32)), ("psf", C.c_void_p), ("psf_capacity", C.c_int64),
                ("voxel_edep", C.POINTER(C.c_uint64)), ("material_edep", C.POINTER(C.c_uint64)),
                ("total_histories", C.c_uint64), ("psf_count", C.c_int64), ("psf_requested", C.c_int64),
                ("sum_trues", C.c_int64), ("sum_scatter", C.c_int64), ("final_seed", C.c_int32),
                ("randoms", C.POINTER(C.c_int32)), ("spurious", C.POINTER(C.c_int32)), ("singles", C.c_void_p),
                ("singles_buffer", C.c_int64), ("singles_count", C.c_int64), ("singles_requested", C.c_int64),
                ("coincidences", C.c_int64 * 5)]


class RunParams(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("block_size", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("zero_tallies", C.c_int32), ("seed", C.c_int32), ("acquisition_time_ps", C.c_uint64),
                ("phase_policy", C.c_int32), ("phase_cost", C.c_int32 * NPHASES),
                ("count_events", C.c_int32), ("use_graph", C.c_int32), ("lds_layout", C.c_int32)]


LDS_PALETTE, LDS_SPECTRUM, LDS_XS511 = 1, 2, 4


class TuneResult(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("blocks_per_cu", C.c_int32), ("phase_policy", C.c_int32), ("ms", C.c_double),
                ("decays_per_ms", C.c_double), ("lds_layout", C.c_int32), ("stream_decays", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneOptions(C.Structure):
    _fields_ = [("stream_fraction", C.c_double), ("tune_layout", C.c_int32), ("tune_streams", C.c_int32),
                ("n_stream_decays", C.c_int32), ("stream_decays", C.c_int64 * 8)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("zero_ms", C.c_double), ("decays", C.c_uint64), ("streams", C.c_uint64),
                ("woodcock_steps", C.c_uint64), ("material_changes", C.c_uint64), ("compton", C.c_uint64),
                ("compton_shell_evals", C.c_uint64), ("rayleigh", C.c_uint64), ("photoelectric", C.c_uint64),
                ("pairs_tallied", C.c_uint64), ("coincidences", C.c_uint64), ("rn_draws", C.c_uint64),
                ("phase_exec", C.c_uint64 * NPHASES), ("phase_lanes", C.c_uint64 * NPHASES),
                ("phase_cycles", C.c_uint64 * NPHASES), ("coinc_ms", C.c_double), ("singles", C.c_uint64),
                ("pairs_emitted", C.c_uint64), ("prompts_emitted", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["phase_exec"] = list(self.phase_exec)
        d["phase_lanes"] = list(self.phase_lanes)
        d["phase_cycles"] = list(self.phase_cycles)
        return d


_lib = None


def lib():
    """Load libmcpet.so (built by mcgpu-pet_amd/build.py). Raises if absent: no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise McpetError(2, f"{LIB_PATH} not built (run python mcgpu-pet_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.mcpet_read_input.argtypes = [C.c_char_p, P(Config), C.c_char_p, C.c_size_t]
    L.mcpet_load_voxels.argtypes = [C.c_char_p, P(Config), P(Phantom), C.c_char_p, C.c_size_t]
    L.mcpet_phantom_from_arrays.argtypes = [C.c_int32, C.c_int32, C.c_int32, P(C.c_float), P(C.c_int32), P(C.c_float),
                                            P(C.c_float), C.c_float, P(Phantom), C.c_char_p, C.c_size_t]
    L.mcpet_free_phantom.argtypes = [P(Phantom)]
    L.mcpet_free_phantom.restype = None
    L.mcpet_load_materials.argtypes = [C.c_void_p, P(C.c_float), P(Tables), C.c_char_p, C.c_size_t]
    L.mcpet_load_materials_extended.argtypes = [C.c_void_p, P(C.c_float), C.c_double, P(Tables), C.c_char_p, C.c_size_t]
    L.mcpet_free_tables.argtypes = [P(Tables)]
    L.mcpet_free_tables.restype = None
    L.mcpet_sinogram_sizes.argtypes = [P(Config)]
    L.mcpet_sinogram_sizes.restype = None
    L.mcpet_create.argtypes = [P(Config), P(Phantom), P(Tables), P(C.c_void_p)]
    L.mcpet_run.argtypes = [C.c_void_p, P(RunParams)]
    L.mcpet_sync.argtypes = [C.c_void_p]
    L.mcpet_fetch.argtypes = [C.c_void_p, P(Tallies)]
    L.mcpet_get_stats.argtypes = [C.c_void_p, P(Stats)]
    L.mcpet_destroy.argtypes = [C.c_void_p]
    L.mcpet_destroy.restype = None
    L.mcpet_last_error.argtypes = [C.c_void_p]
    L.mcpet_last_error.restype = C.c_char_p
    L.mcpet_get_detector.argtypes = [C.c_void_p, P(Detector)]
    L.mcpet_voxel_dose_report.argtypes = [P(Config), P(Phantom), P(C.c_uint64), C.c_uint64, C.c_void_p]
    L.mcpet_write_outputs.argtypes = [C.c_char_p, P(Config), P(Phantom), P(Tables), P(Tallies), C.c_double, C.c_char_p,
                                      C.c_size_t]
    L.mcpet_selftest_ranecu.argtypes = [C.c_int32, P(C.c_int64), C.c_int32, C.c_int32, P(C.c_float), P(C.c_double),
                                        C.c_int32]
    L.mcpet_selftest_fmath.argtypes = [P(C.c_double), P(C.c_double), C.c_int32, P(C.c_double), P(C.c_float)]
    L.mcpet_selftest_copy.argtypes = [C.c_size_t, C.c_int32, P(C.c_double)]
    L.mcpet_stream_seeds.argtypes = [C.c_int32, P(C.c_int64), C.c_int32, P(C.c_int32), P(C.c_int32)]
    L.mcpet_version.restype = C.c_char_p
    L.mcpet_selftest_sort.argtypes = [P(C.c_uint64), C.c_int64, C.c_int32, P(C.c_uint32)]
    L.mcpet_autotune.argtypes = [C.c_void_p, P(TuneOptions), P(TuneResult), C.c_int32, P(C.c_int32)]
    L.mcpet_selftest_sort_noalloc.argtypes = [P(C.c_uint64), C.c_int64, C.c_int32, P(C.c_uint32)]
    L.mcpet_selftest_gemm_bf16.argtypes = [P(C.c_float), P(C.c_float), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                           P(C.c_float)]
    L.mcpet_struct_sizes.argtypes = [C.POINTER(C.c_size_t), C.c_int32]
    L.mcpet_batch_create.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_void_p)]
    L.mcpet_batch_run.argtypes = [C.c_void_p, C.c_void_p]
    L.mcpet_batch_sync.argtypes = [C.c_void_p]
    L.mcpet_batch_fetch.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    L.mcpet_batch_get_stats.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    L.mcpet_batch_autotune.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    L.mcpet_batch_destroy.argtypes = [C.c_void_p]
    L.mcpet_batch_last_error.argtypes = [C.c_void_p]
    L.mcpet_batch_last_error.restype = C.c_char_p
    L.mcpet_server_create.argtypes = [C.c_void_p, C.c_size_t, C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]
    L.mcpet_server_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
    L.mcpet_server_step.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.mcpet_server_drain.argtypes = [C.c_void_p]
    L.mcpet_server_poll.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]
    L.mcpet_server_fetch.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    L.mcpet_server_get_stats.argtypes = [C.c_void_p, C.c_void_p]
    L.mcpet_server_destroy.argtypes = [C.c_void_p]
    L.mcpet_server_submit_rebinned.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                               C.c_int32, C.c_int32, C.POINTER(C.c_int64)]
    L.mcpet_server_fetch_rebinned.argtypes = [C.c_void_p, C.c_int64, P(C.c_float), P(C.c_float), P(C.c_float),
                                              P(C.c_float), C.POINTER(C.c_int64)]
    L.mcpet_server_submit_estimate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                               C.POINTER(C.c_int64)]
    L.mcpet_server_fetch_estimate.argtypes = [C.c_void_p, C.c_int64, P(C.c_float), P(C.c_float), C.POINTER(C.c_int64)]
    L.mcpet_server_last_error.argtypes = [C.c_void_p]
    L.mcpet_server_last_error.restype = C.c_char_p
    L.mcpet_struct_sizes.restype = C.c_int32
    L.mcpet_sinogram_planes.argtypes = [P(Config), P(C.c_int32)]
    L.mcpet_rebin_sinogram.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_float)]
    L.mcpet_tally_rebinned.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
    L.mcpet_batch_rebin_sinogram.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                             P(C.c_float)]
    L.mcpet_project_maps.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(C.c_float), P(C.c_float)]
    L.mcpet_batch_project_maps.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_float),
                                           P(C.c_float)]
    L.mcpet_surrogate_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                         P(C.c_void_p)]
    L.mcpet_surrogate_param_count.argtypes = [C.c_void_p]
    L.mcpet_surrogate_forward_ms.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.mcpet_surrogate_param_count.restype = C.c_int64
    L.mcpet_surrogate_get_params.argtypes = [C.c_void_p, P(C.c_float)]
    L.mcpet_surrogate_set_params.argtypes = [C.c_void_p, P(C.c_float)]
    L.mcpet_surrogate_forward.argtypes = [C.c_void_p, P(C.c_float), C.c_int32, P(C.c_float), P(C.c_double)]
    L.mcpet_surrogate_gradients.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int32, P(C.c_float),
                                            P(C.c_float)]
    L.mcpet_surrogate_train_step.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int32, C.c_float,
                                             P(C.c_float)]
    L.mcpet_surrogate_estimate.argtypes = [C.c_void_p, C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_int64)]
    L.mcpet_surrogate_destroy.argtypes = [C.c_void_p]
    L.mcpet_surrogate_destroy.restype = None
    L.mcpet_surrogate_last_error.argtypes = [C.c_void_p]
    L.mcpet_surrogate_last_error.restype = C.c_char_p
    _lib = L
    return L


def _check(st, ctx=None, err=None):
    if st != 0:
        msg = err.value.decode(errors="replace") if err is not None and err.value else \
            lib().mcpet_last_error(ctx).decode(errors="replace")
        raise McpetError(st, msg)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def read_input(path):
    cfg = Config()
    err = C.create_string_buffer(1024)
    _check(lib().mcpet_read_input(os.fsencode(path), C.byref(cfg), err, 1024), err=err)
    return cfg


def sinogram_sizes(cfg):
    lib().mcpet_sinogram_sizes(C.byref(cfg))
    return cfg


def sinogram_planes(cfg):
    """Ring sum (single-slice rebinning plane) of every sinogram of cfg's michelogram."""
    out = np.zeros(cfg.nsinos, np.int32)
    _check(lib().mcpet_sinogram_planes(C.byref(cfg), out.ctypes.data_as(C.POINTER(C.c_int32))))
    return out


TALLY_INDEX = {"trues": 0, "scatter": 1, "randoms": 2, "spurious": 3}


def load_voxels(path, cfg=None):
    ph = Phantom()
    err = C.create_string_buffer(1024)
    _check(lib().mcpet_load_voxels(os.fsencode(path), C.byref(cfg) if cfg is not None else None, C.byref(ph), err, 1024),
           err=err)
    return PhantomHandle(ph)


def phantom_from_arrays(material, density, activity, voxel_size, fact_object=1.0):
    """material/density/activity: arrays shaped (nz, ny, nx) (x fastest in memory)."""
    material = np.ascontiguousarray(material, dtype=np.int32)
    density = np.ascontiguousarray(density, dtype=np.float32)
    activity = np.ascontiguousarray(activity, dtype=np.float32)
    nz, ny, nx = material.shape
    vs = (C.c_float * 3)(*[float(v) for v in voxel_size])
    ph = Phantom()
    err = C.create_string_buffer(1024)
    _check(lib().mcpet_phantom_from_arrays(nx, ny, nz, vs, material.ctypes.data_as(C.POINTER(C.c_int32)),
                                           _fptr(density), _fptr(activity), float(fact_object), C.byref(ph), err, 1024),
           err=err)
    return PhantomHandle(ph)


class PhantomHandle:
    """Owns a library-allocated phantom."""

    def __init__(self, ph):
        self.ph = ph

    @property
    def shape(self):
        return (self.ph.nz, self.ph.ny, self.ph.nx)

    @property
    def nvox(self):
        return self.ph.nx * self.ph.ny * self.ph.nz

    def voxel_mat_dens(self):
        return np.ctypeslib.as_array(self.ph.voxel_mat_dens, shape=(self.nvox, 3)).copy()

    def density_max(self):
        return np.array(self.ph.density_max[:], dtype=np.float32)

    def close(self):
        if self.ph is not None and self.ph.owned:
            lib().mcpet_free_phantom(C.byref(self.ph))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _names_buffer(names):
    buf = ((C.c_char * NAME_LEN) * MAX_MATERIALS)()
    for i, n in enumerate(names[:MAX_MATERIALS]):
        b = os.fsencode(n) if n else b""
        buf[i].value = b[:NAME_LEN - 1]
    return buf


def load_materials(names, density_max, e_max=0.0):
    """names: up to 15 material file paths ('' for unused slots). e_max (eV) above the files'
    last energy: the grid continued up to e_max (mcpet_load_materials_extended)."""
    if isinstance(names, Config):
        names = [names.material_files[i].value.decode() for i in range(MAX_MATERIALS)]
    names = list(names) + [""] * (MAX_MATERIALS - len(names))
    dm = np.ascontiguousarray(density_max, dtype=np.float32)
    t = Tables()
    err = C.create_string_buffer(1024)
    buf = _names_buffer(names)
    _check(lib().mcpet_load_materials_extended(C.cast(buf, C.c_void_p), _fptr(dm), float(e_max), C.byref(t), err, 1024),
           err=err)
    return TablesHandle(t)


class TablesHandle:
    def __init__(self, t):
        self.t = t

    def arrays(self):
        t = self.t
        n = t.num_values
        return {
            "num_values": n, "e0": t.e0, "ide": t.ide,
            "density_nominal": np.array(t.density_nominal[:], dtype=np.float32),
            "woodcock": np.ctypeslib.as_array(t.woodcock, shape=(n, 2)).copy(),
            "mfp_a": np.ctypeslib.as_array(t.mfp_a, shape=(n, MAX_MATERIALS, 3)).copy(),
            "mfp_b": np.ctypeslib.as_array(t.mfp_b, shape=(n, MAX_MATERIALS, 3)).copy(),
            "pmax": np.ctypeslib.as_array(t.rayleigh.contents.pmax)[:n * MAX_MATERIALS].reshape(n, MAX_MATERIALS).copy(),
            "num_file_values": t.num_file_values,
        }

    def close(self):
        if self.t is not None and self.t.owned:
            lib().mcpet_free_tables(C.byref(self.t))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """A device context: phantom and tables uploaded once, tallies device-resident."""

    def __init__(self, cfg, phantom, tables):
        self.cfg = cfg
        self.phantom = phantom
        self.tables = tables
        self.tally_grid = None   # the grid of mcpet_tally_rebinned, None while off
        self.ctx = C.c_void_p()
        st = lib().mcpet_create(C.byref(cfg), C.byref(phantom.ph), C.byref(tables.t), C.byref(self.ctx))
        if st != 0:
            raise McpetError(st, lib().mcpet_last_error(None).decode(errors="replace"))

    def run(self, kernel=KERNEL_AUTO, block_size=0, blocks_per_cu=0, zero_tallies=True, seed=0, acquisition_time_ps=0,
            phase_policy=0, phase_cost=None, count_events=False, use_graph=False, lds_layout=None):
        """Launch the transport (asynchronous). count_events selects the instrumented
        persistent kernel, which fills every event count of stats(). lds_layout: LDS_* bits of
        an explicit LDS layout (None: the autotuned or default one)."""
        rp = _run_params(kernel, block_size, blocks_per_cu, zero_tallies, seed, acquisition_time_ps, phase_policy,
                         phase_cost, count_events, use_graph, lds_layout)
        _check(lib().mcpet_run(self.ctx, C.byref(rp)), self.ctx)

    def autotune(self, fraction=0.02, layout=True, stream_decays=None):
        """Time launch shapes, phase policies, the LDS layout and (if stream_decays lists
        candidates, 0 = threads_per_block streams per voxel) the decays per stream on every
        round(1/fraction)-th stream; later runs use the fastest. Returns the timed candidates,
        the last entry being the choice."""
        o = TuneOptions()
        o.stream_fraction = float(fraction)
        o.tune_layout = 1 if layout else 0
        cands = list(stream_decays or [])
        if len(cands) > 8:
            raise ValueError("at most 8 stream-layout candidates")
        o.tune_streams = 1 if cands else 0
        o.n_stream_decays = len(cands)
        for i, v in enumerate(cands):
            o.stream_decays[i] = int(v)
        res = (TuneResult * 32)()
        n = C.c_int32()
        _check(lib().mcpet_autotune(self.ctx, C.byref(o), res, 32, C.byref(n)), self.ctx)
        return [r.as_dict() for r in res[:n.value]]

    def sync(self):
        _check(lib().mcpet_sync(self.ctx), self.ctx)

    def stats(self):
        s = Stats()
        _check(lib().mcpet_get_stats(self.ctx, C.byref(s)), self.ctx)
        return s.as_dict()

    def detector(self):
        d = Detector()
        _check(lib().mcpet_get_detector(self.ctx, C.byref(d)), self.ctx)
        return d

    def fetch(self, sinograms=True, images=True, spectrum=True, psf=True, dose=True, singles=True):
        return _fetch_tallies(lambda t: _check(lib().mcpet_fetch(self.ctx, C.byref(t)), self.ctx), self.cfg, self.phantom.nvox,
                              sinograms, images, spectrum, psf, dose, singles)

    def tally_rebinned(self, grid):
        """From the next run on, also tally trues and scatter directly on the (D, H, W) rebinning
        grid (mcpet_tally_rebinned; the fused estimate then reads them); None turns it off."""
        D, H, W = grid if grid is not None else (0, 0, 0)
        _check(lib().mcpet_tally_rebinned(self.ctx, D, H, W), self.ctx)
        self.tally_grid = tuple(grid) if grid is not None else None

    def rebin(self, which, grid):
        """Tally `which` ("trues", "scatter", ...) summed on the device into a (D, H, W) grid."""
        D, H, W = grid
        out = np.zeros((D, H, W), np.float32)
        _check(lib().mcpet_rebin_sinogram(self.ctx, TALLY_INDEX[which], D, H, W, _fptr(out)), self.ctx)
        return out

    def project(self, grid, fetch=True):
        """(activity integral, attenuation factor) of the phantom along the lines of response of
        each (D, H, W) grid bin (mcpet_project_maps); fetch False: computed and kept on the
        device only (for mcpet_surrogate_estimate), returns None."""
        D, H, W = grid
        if not fetch:
            _check(lib().mcpet_project_maps(self.ctx, D, H, W, None, None), self.ctx)
            return None
        act = np.zeros((D, H, W), np.float32)
        att = np.zeros((D, H, W), np.float32)
        _check(lib().mcpet_project_maps(self.ctx, D, H, W, _fptr(act), _fptr(att)), self.ctx)
        return act, att

    def close(self):
        if self.ctx:
            lib().mcpet_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


LDS_SET = 8


def _run_params(kernel=KERNEL_AUTO, block_size=0, blocks_per_cu=0, zero_tallies=True, seed=0, acquisition_time_ps=0,
                phase_policy=0, phase_cost=None, count_events=False, use_graph=False, lds_layout=None):
    return RunParams(kernel, block_size, blocks_per_cu, 1 if zero_tallies else 0, seed, acquisition_time_ps, phase_policy,
                     (C.c_int32 * NPHASES)(*(phase_cost or [0] * NPHASES)), 1 if count_events else 0,
                     1 if use_graph else 0, 0 if lds_layout is None else (LDS_SET | int(lds_layout)))


def _fetch_tallies(fetch_fn, cfg, nv, sinograms=True, images=True, spectrum=True, psf=True, dose=True, singles=True):
    """Allocate the result arrays for cfg, let fetch_fn fill a Tallies struct, return them."""
    out = {}
    t = Tallies()
    if sinograms:
        out["trues"] = np.zeros(cfg.nbins, np.int32)
        out["scatter"] = np.zeros(cfg.nbins, np.int32)
        t.trues = out["trues"].ctypes.data_as(C.POINTER(C.c_int32))
        t.scatter = out["scatter"].ctypes.data_as(C.POINTER(C.c_int32))
    if images:
        out["image_trues"] = np.zeros(nv, np.int32)
        out["image_scatter"] = np.zeros(nv, np.int32)
        t.image_trues = out["image_trues"].ctypes.data_as(C.POINTER(C.c_int32))
        t.image_scatter = out["image_scatter"].ctypes.data_as(C.POINTER(C.c_int32))
    if spectrum and cfg.ne > 0:
        out["spectrum"] = np.zeros(cfg.ne, np.int32)
        t.spectrum = out["spectrum"].ctypes.data_as(C.POINTER(C.c_int32))
    psf_arr = None
    if psf and cfg.detector.tally_psf_sinogram in (0, 1) and cfg.detector.psf_size > 0:
        psf_arr = np.zeros(cfg.detector.psf_size, PSF_DTYPE)
        t.psf = psf_arr.ctypes.data
        t.psf_capacity = cfg.detector.psf_size
    if dose:
        out["material_edep"] = np.zeros((MAX_MATERIALS, 2), np.uint64)
        t.material_edep = out["material_edep"].ctypes.data_as(C.POINTER(C.c_uint64))
        if cfg.voxel_dose and cfg.roi_x_max > -1:
            n_roi = (cfg.roi_x_max - cfg.roi_x_min + 1) * (cfg.roi_y_max - cfg.roi_y_min + 1) * \
                    (cfg.roi_z_max - cfg.roi_z_min + 1)
            out["voxel_edep"] = np.zeros((n_roi, 2), np.uint64)
            t.voxel_edep = out["voxel_edep"].ctypes.data_as(C.POINTER(C.c_uint64))
    singles_arr = None
    if cfg.detector.tally_psf_sinogram == TALLY_SINGLES:
        if sinograms:
            out["randoms"] = np.zeros(cfg.nbins, np.int32)
            out["spurious"] = np.zeros(cfg.nbins, np.int32)
            t.randoms = out["randoms"].ctypes.data_as(C.POINTER(C.c_int32))
            t.spurious = out["spurious"].ctypes.data_as(C.POINTER(C.c_int32))
        if singles:
            singles_arr = np.zeros(cfg.singles_capacity, SINGLE_DTYPE)
            t.singles = singles_arr.ctypes.data
            t.singles_buffer = cfg.singles_capacity
    fetch_fn(t)
    if singles_arr is not None:
        out["singles"] = singles_arr[:t.singles_count].copy()
    if cfg.detector.tally_psf_sinogram == TALLY_SINGLES:
        out["singles_count"] = t.singles_count
        out["singles_requested"] = t.singles_requested
        out["coincidences"] = dict(zip(("trues", "scatter", "randoms", "spurious", "multiples"),
                                       [int(x) for x in t.coincidences]))
    if psf_arr is not None:
        out["psf"] = psf_arr[:t.psf_count].copy()
    out["total_histories"] = t.total_histories
    out["psf_count"] = t.psf_count
    out["psf_requested"] = t.psf_requested
    out["sum_trues"] = t.sum_trues
    out["sum_scatter"] = t.sum_scatter
    out["final_seed"] = t.final_seed
    return out


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64), ("steps", C.c_int64),
                ("queued", C.c_int32), ("running", C.c_int32), ("budget_bytes", C.c_size_t),
                ("peak_admitted_bytes", C.c_size_t), ("peak_running", C.c_int32), ("latency_p50_ms", C.c_double),
                ("latency_p99_ms", C.c_double), ("latency_max_ms", C.c_double), ("busy_ms", C.c_double),
                ("failed", C.c_int64)]


REQ_STATES = {0: "queued", 1: "running", 2: "done", 3: "rejected", 4: "unknown", 5: "failed"}


class Server:
    """Serving loop (mcpet_server_*): continuous batching of simulation requests under an HBM
    budget; each step advances every admitted request by streams_per_step decay streams in
    one batched launch. A request's result equals its standalone Simulation bit for bit."""

    def __init__(self, tables, budget_bytes, max_running=32, streams_per_step=1 << 20):
        self.tables = tables
        self.s = C.c_void_p()
        self._meta = {}
        st = lib().mcpet_server_create(C.byref(tables.t), int(budget_bytes), int(max_running), int(streams_per_step),
                                       C.byref(self.s))
        if st != 0:
            raise McpetError(st, lib().mcpet_last_error(None).decode(errors="replace"))

    def _check(self, st):
        if st != 0:
            raise McpetError(st, lib().mcpet_server_last_error(self.s).decode(errors="replace"))

    def submit(self, cfg, phantom, seed=0):
        rid = C.c_int64()
        st = lib().mcpet_server_submit(self.s, C.byref(cfg), C.byref(phantom.ph), int(seed), C.byref(rid))
        self._meta[rid.value] = {"kind": "full", "cfg": cfg, "nvox": phantom.nvox}
        self._check(st)
        return rid.value

    def step(self):
        r, c = C.c_int32(), C.c_int32()
        self._check(lib().mcpet_server_step(self.s, C.byref(r), C.byref(c)))
        return r.value, c.value

    def drain(self):
        self._check(lib().mcpet_server_drain(self.s))

    def poll(self, rid):
        v = C.c_int32()
        self._check(lib().mcpet_server_poll(self.s, int(rid), C.byref(v)))
        return REQ_STATES[v.value]

    def fetch(self, rid, **kw):
        m = self._meta.get(rid)
        if m is None or m["kind"] != "full":    # the library reports the error
            self._check(lib().mcpet_server_fetch(self.s, int(rid), C.byref(Tallies())))
        out = _fetch_tallies(lambda t: self._check(lib().mcpet_server_fetch(self.s, int(rid), C.byref(t))), m["cfg"],
                             m["nvox"], singles=False, **kw)
        self._meta.pop(rid)
        return out

    def submit_rebinned(self, cfg, phantom, grid, seed=0, maps=False):
        """A request whose result is its trues and scatter rebinned on the device onto grid (and,
        with maps, the phantom's line integrals on that grid)."""
        rid = C.c_int64()
        D, H, W = grid
        st = lib().mcpet_server_submit_rebinned(self.s, C.byref(cfg), C.byref(phantom.ph), int(seed), D, H, W,
                                                int(bool(maps)), C.byref(rid))
        self._meta[rid.value] = {"kind": "rebinned", "grid": tuple(grid), "maps": bool(maps)}
        self._check(st)
        return rid.value

    def fetch_rebinned(self, rid):
        """(trues, scatter, decays, maps) of a finished rebinned request; maps is (activity
        integral, attenuation factor) or None."""
        m = self._meta.get(rid)
        if m is None or m["kind"] != "rebinned":    # the library reports the error
            self._check(lib().mcpet_server_fetch_rebinned(self.s, int(rid), None, None, None, None, None))
        grid = m["grid"]
        t = np.zeros(grid, np.float32)
        s = np.zeros(grid, np.float32)
        maps = (np.zeros(grid, np.float32), np.zeros(grid, np.float32)) if m["maps"] else None
        d = C.c_int64()
        self._check(lib().mcpet_server_fetch_rebinned(self.s, int(rid), _fptr(t), _fptr(s), _fptr(maps[0]) if maps else None,
                                                      _fptr(maps[1]) if maps else None, C.byref(d)))
        self._meta.pop(rid)
        return t, s, d.value, maps

    def submit_estimate(self, cfg, phantom, net, seed=0):
        """A request answered on the device by the scatter network `net` (a Surrogate) when its
        run completes (mcpet_server_submit_estimate); net must outlive the request."""
        rid = C.c_int64()
        st = lib().mcpet_server_submit_estimate(self.s, C.byref(cfg), C.byref(phantom.ph), int(seed), net.s, C.byref(rid))
        self._meta[rid.value] = {"kind": "estimate", "grid": tuple(net.grid)}
        self._check(st)
        return rid.value

    def fetch_estimate(self, rid):
        """(scatter estimate per decay, rebinned scatter counts, decays) of a finished estimate request."""
        m = self._meta.get(rid)
        if m is None or m["kind"] != "estimate":    # the library reports the error
            self._check(lib().mcpet_server_fetch_estimate(self.s, int(rid), None, None, None))
        e = np.zeros(m["grid"], np.float32)
        c = np.zeros(m["grid"], np.float32)
        d = C.c_int64()
        self._check(lib().mcpet_server_fetch_estimate(self.s, int(rid), _fptr(e), _fptr(c), C.byref(d)))
        self._meta.pop(rid)
        return e, c, d.value

    def stats(self):
        x = ServerStats()
        self._check(lib().mcpet_server_get_stats(self.s, C.byref(x)))
        return {k: getattr(x, k) for k, _ in x._fields_}

    def close(self):
        if self.s:
            lib().mcpet_server_destroy(self.s)
            self.s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """n phantoms sharing scanner, tallies and tables, simulated in one persistent launch
    (mcpet_batch_*). Phantom k's results equal a Simulation of phantoms[k] bit for bit."""

    def __init__(self, cfg, phantoms, tables):
        self.cfg = cfg
        self.phantoms = list(phantoms)
        self.tables = tables
        arr = (Phantom * len(self.phantoms))(*[p.ph for p in self.phantoms])
        self.b = C.c_void_p()
        st = lib().mcpet_batch_create(C.byref(cfg), arr, len(self.phantoms), C.byref(tables.t), C.byref(self.b))
        if st != 0:
            raise McpetError(st, lib().mcpet_last_error(None).decode(errors="replace"))

    def _check(self, st):
        if st != 0:
            raise McpetError(st, lib().mcpet_batch_last_error(self.b).decode(errors="replace"))

    def run(self, **kw):
        self._check(lib().mcpet_batch_run(self.b, C.byref(_run_params(**kw))))

    def sync(self):
        self._check(lib().mcpet_batch_sync(self.b))

    def fetch(self, k, **kw):
        return _fetch_tallies(lambda t: self._check(lib().mcpet_batch_fetch(self.b, k, C.byref(t))), self.cfg,
                              self.phantoms[k].nvox, **kw)

    def stats(self, k):
        s = Stats()
        self._check(lib().mcpet_batch_get_stats(self.b, k, C.byref(s)))
        return s.as_dict()

    def rebin(self, k, which, grid):
        D, H, W = grid
        out = np.zeros((D, H, W), np.float32)
        self._check(lib().mcpet_batch_rebin_sinogram(self.b, k, TALLY_INDEX[which], D, H, W, _fptr(out)))
        return out

    def project(self, k, grid):
        """Simulation.project of phantom k."""
        D, H, W = grid
        act = np.zeros((D, H, W), np.float32)
        att = np.zeros((D, H, W), np.float32)
        self._check(lib().mcpet_batch_project_maps(self.b, k, D, H, W, _fptr(act), _fptr(att)))
        return act, att

    def autotune(self, fraction=0.05):
        res = (TuneResult * 32)()
        n = C.c_int32()
        self._check(lib().mcpet_batch_autotune(self.b, float(fraction), res, 32, C.byref(n)))
        return [r.as_dict() for r in res[:n.value]]

    def close(self):
        if self.b:
            lib().mcpet_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_outputs(directory, cfg, phantom, tables, tallies, time_elapsed=0.0):
    """Write the reference output files for a fetched result dict (see Simulation.fetch)."""
    t = Tallies()
    keep = []

    def ip(name):
        a = tallies.get(name)
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int32))

    t.trues, t.scatter = ip("trues"), ip("scatter")
    t.image_trues, t.image_scatter, t.spectrum = ip("image_trues"), ip("image_scatter"), ip("spectrum")
    if tallies.get("psf") is not None:
        p = np.ascontiguousarray(tallies["psf"])
        keep.append(p)
        t.psf = p.ctypes.data
        t.psf_capacity = len(p)
        t.psf_count = len(p)
    if tallies.get("voxel_edep") is not None:
        v = np.ascontiguousarray(tallies["voxel_edep"], dtype=np.uint64)
        keep.append(v)
        t.voxel_edep = v.ctypes.data_as(C.POINTER(C.c_uint64))
    t.total_histories = int(tallies.get("total_histories", 0))
    err = C.create_string_buffer(1024)
    _check(lib().mcpet_write_outputs(os.fsencode(directory), C.byref(cfg), C.byref(phantom.ph), C.byref(tables.t),
                                     C.byref(t), float(time_elapsed), err, 1024), err=err)


class DoseReport(C.Structure):
    _fields_ = [("total_energy_kev", C.c_double), ("max_dose", C.c_double), ("max_dose_std_dev", C.c_double),
                ("mass", C.c_double * MAX_MATERIALS), ("edep", C.c_double * MAX_MATERIALS),
                ("dose", C.c_double * MAX_MATERIALS), ("dose_std_dev", C.c_double * MAX_MATERIALS),
                ("peak_dose", C.c_double * MAX_MATERIALS), ("peak_std_dev", C.c_double * MAX_MATERIALS),
                ("peak_voxel", (C.c_int32 * 3) * MAX_MATERIALS), ("voxels", C.c_uint32 * MAX_MATERIALS)]


def voxel_dose_report(cfg, phantom, voxel_edep, total_histories):
    """The ROI dose summary of the reference's voxel dose report (mcpet_voxel_dose_report) as a
    dict: total_energy_kev, max_dose, max_dose_std_dev and per material slot (lists of 15) mass,
    edep, dose, dose_std_dev, peak_dose, peak_std_dev, peak_voxel, voxels."""
    v = np.ascontiguousarray(voxel_edep, dtype=np.uint64)
    r = DoseReport()
    _check(lib().mcpet_voxel_dose_report(C.byref(cfg), C.byref(phantom.ph), v.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         int(total_histories), C.byref(r)))
    out = {k: getattr(r, k) for k in ("total_energy_kev", "max_dose", "max_dose_std_dev")}
    for k in ("mass", "edep", "dose", "dose_std_dev", "peak_dose", "peak_std_dev", "voxels"):
        out[k] = list(getattr(r, k))
    out["peak_voxel"] = [tuple(r.peak_voxel[m]) for m in range(MAX_MATERIALS)]
    return out


def selftest_ranecu(seed, stream_ids, n, as_double=False):
    ids = np.ascontiguousarray(stream_ids, dtype=np.int64)
    if as_double:
        out = np.zeros((len(ids), n), np.float64)
        _check(lib().mcpet_selftest_ranecu(seed, ids.ctypes.data_as(C.POINTER(C.c_int64)), len(ids), n, None,
                                           out.ctypes.data_as(C.POINTER(C.c_double)), 1))
    else:
        out = np.zeros((len(ids), n), np.float32)
        _check(lib().mcpet_selftest_ranecu(seed, ids.ctypes.data_as(C.POINTER(C.c_int64)), len(ids), n,
                                           out.ctypes.data_as(C.POINTER(C.c_float)), None, 0))
    return out


def stream_seeds(seed, stream_ids):
    """Host-side RANECU seeds at the start of each stream (same skip-ahead as the kernels)."""
    ids = np.ascontiguousarray(stream_ids, dtype=np.int64)
    sx = np.zeros(len(ids), np.int32)
    sy = np.zeros(len(ids), np.int32)
    _check(lib().mcpet_stream_seeds(seed, ids.ctypes.data_as(C.POINTER(C.c_int64)), len(ids),
                                    sx.ctypes.data_as(C.POINTER(C.c_int32)), sy.ctypes.data_as(C.POINTER(C.c_int32))))
    return sx, sy


def selftest_fmath(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = len(x)
    od = np.zeros((7, n), np.float64)
    of = np.zeros((8, n), np.float32)
    _check(lib().mcpet_selftest_fmath(x.ctypes.data_as(C.POINTER(C.c_double)), y.ctypes.data_as(C.POINTER(C.c_double)),
                                      n, od.ctypes.data_as(C.POINTER(C.c_double)), of.ctypes.data_as(C.POINTER(C.c_float))))
    return od, of


def selftest_sort(keys, passes=8, noalloc=False):
    """Device radix sort of uint64 keys (low 8*passes bits); returns the stable sorted order.
    noalloc: the allocation-free entry point (static device scratch, at most 2^20 keys)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    perm = np.zeros(len(k), np.uint32)
    fn = lib().mcpet_selftest_sort_noalloc if noalloc else lib().mcpet_selftest_sort
    _check(fn(k.ctypes.data_as(C.POINTER(C.c_uint64)), len(k), int(passes), perm.ctypes.data_as(C.POINTER(C.c_uint32))))
    return perm


def selftest_gemm_bf16(a, b, split=1):
    """C = a @ b on the device's BF16 MFMA GEMM (FP32 accumulation)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    M, K = a.shape
    N = b.shape[1]
    c = np.zeros((M, N), np.float32)
    _check(lib().mcpet_selftest_gemm_bf16(_fptr(a), _fptr(b), M, N, K, int(split), _fptr(c)))
    return c


class Surrogate:
    """The scatter surrogate's 3D U-Net on the device (mcpet_surrogate_*): BF16 MFMA
    convolutions with FP32 accumulation, FP32 parameters and Adam state."""

    def __init__(self, grid, in_channels=2, width=16, seed=1):
        self.grid = tuple(int(v) for v in grid)
        self.in_channels = int(in_channels)
        self.width = int(width)
        self.s = C.c_void_p()
        D, H, W = self.grid
        _check(lib().mcpet_surrogate_create(D, H, W, self.in_channels, self.width, int(seed), C.byref(self.s)))
        self.n_params = int(lib().mcpet_surrogate_param_count(self.s))
        self.gflop = 0.0

    def _check(self, st):
        if st != 0:
            raise McpetError(st, lib().mcpet_surrogate_last_error(self.s).decode(errors="replace"))

    def _x(self, x):
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim == 4:
            x = x[None]
        if x.shape[1:] != (self.in_channels,) + self.grid:
            raise ValueError(f"input shape {x.shape} does not match (batch, {self.in_channels}, {self.grid})")
        return x

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self._check(lib().mcpet_surrogate_get_params(self.s, _fptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("parameter vector size mismatch")
        self._check(lib().mcpet_surrogate_set_params(self.s, _fptr(p)))

    def forward(self, x):
        x = self._x(x)
        y = np.zeros((x.shape[0],) + self.grid, np.float32)
        g = C.c_double()
        self._check(lib().mcpet_surrogate_forward(self.s, _fptr(x), x.shape[0], _fptr(y), C.byref(g)))
        self.gflop = g.value
        return y

    def forward_ms(self):
        """Device time (ms) of the network in the last forward() call (its last sample)."""
        ms = C.c_double()
        self._check(lib().mcpet_surrogate_forward_ms(self.s, C.byref(ms)))
        return ms.value

    def estimate(self, sim):
        """Fused device estimate from the last run of Simulation `sim` (mcpet_surrogate_estimate):
        (scatter per decay, rebinned scatter counts, decays)."""
        est = np.zeros(self.grid, np.float32)
        cnt = np.zeros(self.grid, np.float32)
        d = C.c_int64()
        self._check(lib().mcpet_surrogate_estimate(self.s, sim.ctx, _fptr(est), _fptr(cnt), C.byref(d)))
        return est, cnt, d.value

    def _target(self, x, t):
        t = np.ascontiguousarray(t, np.float32).reshape((x.shape[0],) + self.grid)
        return t

    def gradients(self, x, target):
        x = self._x(x)
        t = self._target(x, target)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        self._check(lib().mcpet_surrogate_gradients(self.s, _fptr(x), _fptr(t), x.shape[0], C.byref(loss), _fptr(g)))
        return loss.value, g

    def train_step(self, x, target, lr=1e-3):
        x = self._x(x)
        t = self._target(x, target)
        loss = C.c_float()
        self._check(lib().mcpet_surrogate_train_step(self.s, _fptr(x), _fptr(t), x.shape[0], float(lr), C.byref(loss)))
        return loss.value

    def close(self):
        if self.s:
            lib().mcpet_surrogate_destroy(self.s)
            self.s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def selftest_copy(nbytes, iters=10):
    ms = C.c_double()
    _check(lib().mcpet_selftest_copy(int(nbytes), int(iters), C.byref(ms)))
    return ms.value


class Single(C.Structure):
    _fields_ = [("time_ps", C.c_uint64), ("event", C.c_uint64), ("energy", C.c_float), ("z", C.c_float),
                ("phi", C.c_float), ("scatter", C.c_int16), ("kind", C.c_int16)]


SINGLE_DTYPE = np.dtype([("time_ps", "<u8"), ("event", "<u8"), ("energy", "<f4"), ("z", "<f4"), ("phi", "<f4"),
                         ("scatter", "<i2"), ("kind", "<i2")])
TALLY_SINGLES = 3


def struct_sizes():
    """sizeof of the C structs as compiled into the library (see mcpet_struct_sizes)."""
    out = (C.c_size_t * 16)()
    n = lib().mcpet_struct_sizes(out, 16)
    return list(out)[:n]


def version():
    return lib().mcpet_version().decode()
