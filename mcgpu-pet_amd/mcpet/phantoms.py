"""Seeded synthetic phantoms and simulation configurations.

Workloads (all phantoms are generated here; only air and water cross-section files
exist, so tissues are water or air tables at their own densities, one material slot
per tissue):
  water_cube   -- 9x9x9 voxels of 1 cm water, uniform 18F background with three voxels
                  at 100x activity in the central plane (sample positions), 1e7 decays.
  head         -- 128x128x96 voxels of 2 mm: seed-placed ellipsoids for scalp, fat,
                  skull (bone 1.85 g/cm3), CSF, grey and white matter, sinus air, in air;
                  18F activity grey 4 : white 1 : other tissue 0.5 : air 0; 1e9 decays in 10 s.
  scatter_phantoms -- 128x128x64 voxels of 2 mm, random bodies, inclusions and lesions:
                  the scatter surrogate's training and evaluation phantoms.
  sample_cube  -- the reference's own sample phantom (phantom_9x9x9cm.vox), written
                  byte for byte by write_sample_vox.
"""
import os

import numpy as np

from . import Config, MAX_MATERIALS, NAME_LEN, phantom_from_arrays, sinogram_sizes

F18_MEAN_LIFE = 9502.0          # s: 109.77 min half-life / ln 2

AIR, WATER = "air_5-515keV.mcgpu.gz", "water_5-515keV.mcgpu.gz"


def write_sample_vox(path):
    """The reference sample phantom (sample_simulation/phantom_9x9x9cm.vox): a 5x5x5 cm water
    cube with 50 Bq voxels, 1000/2000/3000 Bq inserts and one air cavity in the central
    plane, inside 9x9x9 cm of air. Same bytes as the reference file."""
    lines = ["[SECTION VOXELS HEADER v.2008-04-13]\n", "9 9 9   No. OF VOXELS IN X,Y,Z\n",
             "1.0 1.0 1.0   VOXEL SIZE (cm) ALONG X,Y,Z\n",
             " 1                  COLUMN NUMBER WHERE MATERIAL ID IS LOCATED\n",
             " 2                  COLUMN NUMBER WHERE THE MASS DENSITY IS LOCATED\n",
             " 1                  BLANK LINES AT END OF X,Y-CYCLES (1=YES,0=NO)\n",
             "[END OF VXH SECTION]  # MCGPU-PET voxel format: Material  Density  Activity\n"]
    air, water = "1 0.0012 0.0\n", "2 1.0 %d\n"
    edge = 2
    for k in range(9):
        for j in range(9):
            for i in range(9):
                if i < edge or j < edge or k < edge or i >= 9 - edge or j >= 9 - edge or k >= 9 - edge:
                    lines.append(air)
                elif (i, j, k) == (3, 3, 4):
                    lines.append(water % 1000)
                elif (i, j, k) == (5, 3, 4):
                    lines.append(water % 2000)
                elif (i, j, k) == (5, 5, 4):
                    lines.append(water % 3000)
                elif (i, j, k) == (3, 5, 4):
                    lines.append(air)
                else:
                    lines.append(water % 50)
            lines.append("\n")
        lines.append("\n")
    lines.append("\n")
    with open(path, "w") as f:
        f.write("".join(lines))


def write_vox(path, material, density, activity, voxel_size):
    """Write a .vox file (MCGPU-PET format: material, density, activity per voxel, x fastest).
    Values are printed with 9 significant digits so they read back to the same float32."""
    material = np.asarray(material)
    nz, ny, nx = material.shape
    with open(path, "w") as f:
        f.write("[SECTION VOXELS HEADER v.2008-04-13]\n")
        f.write(f"{nx} {ny} {nz}   No. OF VOXELS IN X,Y,Z\n")
        f.write("%.9g %.9g %.9g   VOXEL SIZE (cm) ALONG X,Y,Z\n" % tuple(voxel_size))
        f.write(" 1                  COLUMN NUMBER WHERE MATERIAL ID IS LOCATED\n")
        f.write(" 2                  COLUMN NUMBER WHERE THE MASS DENSITY IS LOCATED\n")
        f.write(" 1                  BLANK LINES AT END OF X,Y-CYCLES (1=YES,0=NO)\n")
        f.write("[END OF VXH SECTION]  # MCGPU-PET voxel format: Material  Density  Activity\n")
        m = material.reshape(nz, ny, nx)
        d = np.asarray(density, np.float32).reshape(nz, ny, nx)
        a = np.asarray(activity, np.float32).reshape(nz, ny, nx)
        for k in range(nz):
            for j in range(ny):
                f.write("".join("%d %.9g %.9g\n" % (m[k, j, i], d[k, j, i], a[k, j, i]) for i in range(nx)))
                f.write("\n")
            f.write("\n")


def make_config(*, seed=12345, threads_per_block=64, acquisition_time_s=1.0, mean_life=F18_MEAN_LIFE,
                radius=-40.0, height=25.0, center=(0.0, 0.0, 0.0), psf_size=0, tally_psf_sinogram=2, tally_type=0,
                material_dose=False, voxel_dose=False, roi=None, e_resol=0.12, e_low=350000.0, e_high=600000.0,
                fovz=25.0, nrows=80, ncrystals=336, nangles=168, nrad=147, nzs=159, res=128, ne=700, mrd=79, span=11,
                materials=(), bg_mode=False, gpu_id=0, positron_fraction=0.0, prompt_gamma_fraction=0.0,
                prompt_gamma_energy=0.0, singles_capacity=0, coincidence_window_ps=0, stream_decays=0):
    """A Config equivalent to a MCGPU-PET.in file with these values (plus the library's
    extensions: decay branching, prompt gamma and the singles pipeline, tally_psf_sinogram=3;
    stream_decays > 0: the balanced decay-stream layout)."""
    c = Config()
    c.seed = seed
    c.gpu_id = gpu_id
    c.threads_per_block = threads_per_block
    c.fact_object = 1.0
    c.acquisition_time_ps = int(acquisition_time_s * 1.0e12 + 0.5)
    c.mean_life = mean_life
    c.psf_file = b"MCGPU_PET.psf"
    c.detector.center[:] = list(center)
    c.detector.radius = radius
    c.detector.height = height
    c.detector.psf_size = psf_size
    c.detector.tally_psf_sinogram = tally_psf_sinogram
    c.detector.tally_type = tally_type
    c.material_dose = 1 if material_dose else 0
    c.voxel_dose = 1 if voxel_dose else 0
    c.dose_file = b"mc-gpu_dose.dat"
    if voxel_dose and roi is not None:
        (c.roi_x_min, c.roi_x_max, c.roi_y_min, c.roi_y_max, c.roi_z_min, c.roi_z_max) = roi
    else:
        c.roi_x_min = c.roi_y_min = c.roi_z_min = 32500
        c.roi_x_max = c.roi_y_max = c.roi_z_max = -32500
    c.e_resol, c.e_low, c.e_high, c.fovz = e_resol, e_low, e_high, fovz
    c.nrows, c.ncrystals, c.nangles, c.nrad, c.nzs = nrows, ncrystals, nangles, nrad, nzs
    c.res, c.ne, c.mrd, c.span = res, ne, mrd, span
    sinogram_sizes(c)
    for i, m in enumerate(list(materials)[:MAX_MATERIALS]):
        c.material_files[i].value = os.fsencode(m)[:NAME_LEN - 1]
    c.bg_mode = 1 if bg_mode else 0
    c.positron_fraction = positron_fraction
    c.prompt_gamma_fraction = prompt_gamma_fraction
    c.prompt_gamma_energy = prompt_gamma_energy
    c.singles_capacity = int(singles_capacity)
    c.coincidence_window_ps = int(coincidence_window_ps)
    c.stream_decays = int(stream_decays)
    return c


def config_dict(c):
    """Plain-dict view of a Config (the oracle's parameter set)."""
    return {
        "seed": c.seed, "threads_per_block": c.threads_per_block, "acquisition_time_ps": c.acquisition_time_ps,
        "mean_life": c.mean_life, "bg_mode": c.bg_mode, "det_center": list(c.detector.center),
        "det_radius": c.detector.radius, "det_height": c.detector.height, "psf_size": c.detector.psf_size,
        "tally_psf_sinogram": c.detector.tally_psf_sinogram, "tally_type": c.detector.tally_type,
        "e_resol": c.e_resol, "e_low": c.e_low, "e_high": c.e_high, "fovz": c.fovz, "nrows": c.nrows,
        "ncrystals": c.ncrystals, "nangles": c.nangles, "nrad": c.nrad, "nzs": c.nzs, "nbins": c.nbins, "ne": c.ne,
        "mrd": c.mrd, "span": c.span, "nsinos": c.nsinos, "material_dose": c.material_dose,
        "roi": (c.roi_x_min, c.roi_x_max, c.roi_y_min, c.roi_y_max, c.roi_z_min, c.roi_z_max) if c.voxel_dose
        else (32500, -32500, 32500, -32500, 32500, -32500),
        "positron_fraction": c.positron_fraction, "prompt_gamma_fraction": c.prompt_gamma_fraction,
        "prompt_gamma_energy": c.prompt_gamma_energy, "singles_capacity": c.singles_capacity,
        "coincidence_window_ps": c.coincidence_window_ps, "stream_decays": c.stream_decays,
    }


def water_cube(n_decays=1.0e7, n=9, voxel_cm=1.0, hot=((3, 3, 4), (5, 3, 4), (5, 5, 4)), hot_factor=100.0,
               acquisition_time_s=1.0, mean_life=F18_MEAN_LIFE):
    """All-water cube; background a0 per voxel, hot voxels hot_factor*a0, E[decays] = n_decays."""
    mat = np.full((n, n, n), 2, np.int32)
    dens = np.ones((n, n, n), np.float32)
    units = n ** 3 - len(hot) + hot_factor * len(hot)
    frac = mean_life * (1.0 - np.exp(-acquisition_time_s / mean_life))     # E[decays] per Bq
    a0 = n_decays / (units * frac)
    act = np.full((n, n, n), a0, np.float32)
    for (i, j, k) in hot:
        act[k, j, i] = hot_factor * a0
    return mat, dens, act, (voxel_cm,) * 3


HEAD_TISSUES = [  # slot (1-based), name, material file, density g/cm3, activity weight
    (1, "air", AIR, 0.0012, 0.0),
    (2, "soft tissue", WATER, 1.03, 0.5),
    (3, "fat", WATER, 0.95, 0.5),
    (4, "bone", WATER, 1.85, 0.5),
    (5, "grey matter", WATER, 1.04, 4.0),
    (6, "white matter", WATER, 1.04, 1.0),
    (7, "CSF", WATER, 1.007, 0.5),
    (8, "sinus air", AIR, 0.0012, 0.0),
]


def _ellipsoid(shape, vs, c, r):
    nz, ny, nx = shape
    z = (np.arange(nz, dtype=np.float32) + 0.5) * vs
    y = (np.arange(ny, dtype=np.float32) + 0.5) * vs
    x = (np.arange(nx, dtype=np.float32) + 0.5) * vs
    return (((x[None, None, :] - c[0]) / r[0]) ** 2 + ((y[None, :, None] - c[1]) / r[1]) ** 2 +
            ((z[:, None, None] - c[2]) / r[2]) ** 2) <= 1.0


def head(seed=20261202, shape=(96, 128, 128), voxel_cm=0.2, n_decays=1.0e9, acquisition_time_s=10.0,
         mean_life=F18_MEAN_LIFE):
    """Head-size phantom: nested seed-jittered ellipsoids (scalp > fat > skull > CSF > grey >
    white), two ventricles (CSF) and a frontal sinus (air). Returns material slots 1..8."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    cx, cy, cz = nx * voxel_cm / 2, ny * voxel_cm / 2, nz * voxel_cm / 2
    j = lambda s: 1.0 + s * (rng.random() - 0.5)
    R = np.array([8.6 * j(0.04), 10.6 * j(0.04), 8.2 * j(0.04)], np.float32)
    c0 = np.array([cx + rng.normal(0, 0.1), cy + rng.normal(0, 0.1), cz + rng.normal(0, 0.1)], np.float32)
    mat = np.ones(shape, np.int32)
    layers = [(2, 0.0), (3, 0.35), (4, 0.55), (7, 1.15), (5, 1.35), (6, 2.6)]  # slot, inset (cm)
    for slot, inset in layers:
        m = _ellipsoid(shape, voxel_cm, c0, R - inset)
        mat[m] = slot
    for side in (-1, 1):   # lateral ventricles
        cv = c0 + np.array([side * 1.1 * j(0.1), 0.6 * j(0.2), 0.8 * j(0.2)], np.float32)
        mat[_ellipsoid(shape, voxel_cm, cv, np.array([0.6, 2.2, 0.9], np.float32) * j(0.1))] = 7
    cs = c0 + np.array([0.0, R[1] - 1.6, -2.2], np.float32)  # frontal sinus
    mat[_ellipsoid(shape, voxel_cm, cs, np.array([1.4, 0.7, 0.9], np.float32) * j(0.1))] = 8
    dens = np.zeros(shape, np.float32)
    w = np.zeros(shape, np.float64)
    for slot, _, _, d, a in HEAD_TISSUES:
        dens[mat == slot] = d
        w[mat == slot] = a
    frac = mean_life * (1.0 - np.exp(-acquisition_time_s / mean_life))
    act = (w * (n_decays / (w.sum() * frac))).astype(np.float32)
    return mat, dens, act, (voxel_cm,) * 3


TORSO_TISSUES = [  # slot (1-based), name, material file, density g/cm3, activity weight
    (1, "air", AIR, 0.0012, 0.0),
    (2, "soft tissue", WATER, 1.03, 1.0),
    (3, "fat", WATER, 0.95, 0.5),
    (4, "bone", WATER, 1.85, 0.8),
    (5, "lung", WATER, 0.30, 0.25),
    (6, "liver", WATER, 1.06, 4.0),
    (7, "lesion", WATER, 1.03, 12.0),
]


def torso(seed=20261202, shape=(200, 256, 256), voxel_cm=0.25, n_decays=5.0e9, acquisition_time_s=10.0,
          mean_life=F18_MEAN_LIFE):
    """Torso-size phantom (64 x 64 x 50 cm at 2.5 mm): an elliptic body (34 x 22 cm, 46 cm long)
    of soft tissue under 1.5 cm of fat, two lungs (0.3 g/cm3), spine and ribs (bone 1.85),
    liver, and two hot lesions, with seed-jittered sizes and positions. Lung, bone and liver
    use the water tables at their densities (only air and water tables exist). Activity
    weights as TORSO_TISSUES, scaled to n_decays expected decays. Returns material slots 1..7."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    vs = voxel_cm
    cx, cy, cz = nx * vs / 2, ny * vs / 2, nz * vs / 2
    j = lambda s: 1.0 + s * (rng.random() - 0.5)
    z = (np.arange(nz, dtype=np.float32) + 0.5) * vs
    y = (np.arange(ny, dtype=np.float32) + 0.5) * vs
    x = (np.arange(nx, dtype=np.float32) + 0.5) * vs
    X, Y = np.meshgrid(x, y)                                  # [ny, nx]
    ax, ay = 17.0 * j(0.06), 11.0 * j(0.06)
    body2 = ((X - cx) / ax) ** 2 + ((Y - cy) / ay) ** 2       # < 1 inside the body's outline
    inner2 = ((X - cx) / (ax - 1.5)) ** 2 + ((Y - cy) / (ay - 1.5)) ** 2
    ring2 = ((X - cx) / (ax - 2.0)) ** 2 + ((Y - cy) / (ay - 2.0)) ** 2
    z0, z1 = cz - 23.0, cz + 23.0
    mat = np.ones(shape, np.int32)
    inz = (z > z0) & (z < z1)
    mat[inz] = np.where(body2 < 1.0, 3, 1)[None]             # fat everywhere inside the outline
    mat[inz] = np.where(inner2 < 1.0, 2, mat[inz])           # soft tissue under it
    thorax = (z > cz + 2.0) & (z < z1 - 2.0)
    for side in (-1, 1):                                      # lungs
        c = np.array([cx + side * 7.5 * j(0.1), cy + 0.5 * j(0.5), cz + 12.0 * j(0.05)], np.float32)
        mat[_ellipsoid(shape, vs, c, np.array([5.0, 6.5, 9.5], np.float32) * j(0.08))] = 5
    for zr in np.arange(cz + 3.0, z1 - 2.0, 2.5):             # ribs: 0.75 cm bands of the ring
        k = (z > zr) & (z < zr + 0.75) & thorax
        band = (ring2 < 1.0) & (ring2 > 0.80)
        mat[k] = np.where(band, 4, mat[k])
    spine = ((X - cx) ** 2 + (Y - (cy - ay + 4.0)) ** 2) < (1.6 * j(0.1)) ** 2
    mat[inz] = np.where(spine, 4, mat[inz])
    cl = np.array([cx - 6.0 * j(0.2), cy + 1.0 * j(0.5), cz - 2.0 * j(0.2)], np.float32)
    mat[_ellipsoid(shape, vs, cl, np.array([7.5, 6.0, 6.5], np.float32) * j(0.08))] = 6   # liver
    for c in (cl + np.array([2.0, -1.0, 1.0], np.float32), np.array([cx + 7.0, cy, cz + 12.0], np.float32)):
        mat[_ellipsoid(shape, vs, c * np.float32(j(0.01)), np.array([1.0, 1.0, 1.0], np.float32))] = 7
    dens = np.zeros(shape, np.float32)
    w = np.zeros(shape, np.float64)
    for slot, _, _, d, a in TORSO_TISSUES:
        m = mat == slot
        dens[m] = d
        w[m] = a
    frac = mean_life * (1.0 - np.exp(-acquisition_time_s / mean_life))
    act = (w * (n_decays / (w.sum() * frac))).astype(np.float32)
    return mat, dens, act, (voxel_cm,) * 3


def torso_materials(materials_dir):
    return [os.path.join(materials_dir, f) for (_, _, f, _, _) in TORSO_TISSUES]


def random_phantoms(count=48, seed=20261202, voxel_cm=0.2, histories=(1.0e6, 1.0e8), xy=(32, 160), z=(32, 96),
                    acquisition_time_s=10.0, mean_life=F18_MEAN_LIFE):
    """A batch of varied phantoms (the many-phantom workload): nx, ny ~ U[xy], nz ~ U[z] at
    voxel_cm, air background with 2-6 random ellipsoids of the head tissues (material slots
    1..8 as HEAD_TISSUES, each with its density and activity weight), and a photon-history
    budget log-uniform in `histories` (2 per decay), i.e. proportional to the phantom's total
    activity. Returns a list of (mat, dens, act, voxel_size)."""
    rng = np.random.default_rng(seed)
    out = []
    frac = mean_life * (1.0 - np.exp(-acquisition_time_s / mean_life))
    for _ in range(count):
        nx, ny = (int(v) for v in rng.integers(xy[0], xy[1] + 1, 2))
        nz = int(rng.integers(z[0], z[1] + 1))
        shape = (nz, ny, nx)
        size = np.array([nx, ny, nz], np.float32) * voxel_cm
        mat = np.ones(shape, np.int32)
        # a body ellipsoid of soft tissue, then inclusions of other tissues
        mat[_ellipsoid(shape, voxel_cm, size / 2, size * rng.uniform(0.3, 0.48, 3).astype(np.float32))] = 2
        for _ in range(int(rng.integers(2, 7))):
            c = (size * rng.uniform(0.25, 0.75, 3)).astype(np.float32)
            r = (size * rng.uniform(0.05, 0.2, 3)).astype(np.float32)
            mat[_ellipsoid(shape, voxel_cm, c, r)] = int(rng.integers(2, 9))
        dens = np.zeros(shape, np.float32)
        w = np.zeros(shape, np.float64)
        for slot, _, _, d, a in HEAD_TISSUES:
            dens[mat == slot] = d
            w[mat == slot] = a
        if w.sum() == 0:
            w[mat == 2] = 1.0
        n_decays = 0.5 * float(np.exp(rng.uniform(np.log(histories[0]), np.log(histories[1]))))
        act = (w * (n_decays / (w.sum() * frac))).astype(np.float32)
        out.append((mat, dens, act, (voxel_cm,) * 3))
    return out


def scatter_phantoms(count, seed, shape=(64, 128, 128), voxel_cm=0.2, n_decays=1.0e9, acquisition_time_s=1.0,
                     mean_life=F18_MEAN_LIFE):
    """Phantoms of the scatter-surrogate workload: 128x128x64 voxels of 2 mm (the activity
    and attenuation maps the surrogate's training draws from). Each is a body ellipsoid of
    soft tissue (semi-axes 30-48 % of the field, offset up to 10 %) with 1-4 inclusions of
    other tissues (bone, lung-like air, fat, brain tissue) and 0-3 hot lesions (3-10x the
    local activity); tissue activity weights are the head weights times a log-normal factor
    (sigma 0.5). Activity scaled to n_decays expected decays in acquisition_time_s. Returns a
    list of (mat, dens, act, voxel_size)."""
    rng = np.random.default_rng(seed)
    frac = mean_life * (1.0 - np.exp(-acquisition_time_s / mean_life))
    nz, ny, nx = shape
    size = np.array([nx, ny, nz], np.float32) * voxel_cm
    out = []
    for _ in range(count):
        mat = np.ones(shape, np.int32)
        c0 = (size / 2 * (1.0 + rng.uniform(-0.1, 0.1, 3))).astype(np.float32)
        body = (size * rng.uniform(0.3, 0.48, 3)).astype(np.float32)
        body_mask = _ellipsoid(shape, voxel_cm, c0, body)
        mat[body_mask] = 2
        for _ in range(int(rng.integers(1, 5))):
            c = (c0 + body * rng.uniform(-0.6, 0.6, 3)).astype(np.float32)
            r = (body * rng.uniform(0.1, 0.4, 3)).astype(np.float32)
            mat[_ellipsoid(shape, voxel_cm, c, r) & body_mask] = int(rng.choice([3, 4, 5, 6, 7, 8]))
        dens = np.zeros(shape, np.float32)
        w = np.zeros(shape, np.float64)
        for slot, _, _, d, a in HEAD_TISSUES:
            dens[mat == slot] = d
            w[mat == slot] = a * float(np.exp(rng.normal(0.0, 0.5)))
        for _ in range(int(rng.integers(0, 4))):
            c = (c0 + body * rng.uniform(-0.5, 0.5, 3)).astype(np.float32)
            lesion = _ellipsoid(shape, voxel_cm, c, np.full(3, rng.uniform(0.4, 1.2), np.float32)) & body_mask
            w[lesion] *= rng.uniform(3.0, 10.0)
        if w.sum() == 0:
            w[mat == 2] = 1.0
        act = (w * (n_decays / (w.sum() * frac))).astype(np.float32)
        out.append((mat, dens, act, (voxel_cm,) * 3))
    return out


def head_materials(materials_dir):
    return [os.path.join(materials_dir, f) for (_, _, f, _, _) in HEAD_TISSUES]


def build_phantom(mat, dens, act, vs):
    return phantom_from_arrays(mat, dens, act, vs)
