"""The phantom seen along the lines of response of the surrogate's grid (mcpet_project_maps; no
reference counterpart). Pinned by closed forms: on a box phantom of water with activity in a
sub-box, a line's activity integral is the activity times its chord through the sub-box and
its attenuation factor exp(-mu L) with L its chord through the box; the crystal pairs of each
grid bin come from the oracle's coincidence binning (K:1749-1788), not from the library."""
import gzip

import numpy as np
import pytest

import mcpet
from conftest import AIR, WATER
from harness import Scenario
from mcpet import phantoms

E511 = 510998.95
GRID = (32, 56, 56)
VOX = 0.4
SHAPE = (32, 40, 64)             # z, y, x voxels: a 25.6 x 16 x 12.8 cm box
ACT_BOX = ((8, 40), (10, 30))    # activity in x voxels [8, 40), y voxels [10, 30), every z
A0 = 1000.0


def mu_water_511():
    """1 / total mean free path of water at 511 keV (the first table of the material file:
    energy, Rayleigh, Compton, photoelectric and total mean free paths; linear in energy
    between its grid points)."""
    with gzip.open(WATER, "rt") as f:
        lines = f.read().splitlines()
    n = int(lines[lines.index("#[NUMBER OF DATA VALUES]") + 1].strip("# "))
    rows = [l for l in lines if l.strip() and not l.startswith("#")][:n]
    t = np.array([[float(v) for v in l.split()[:5]] for l in rows])
    return float(np.interp(E511, t[:, 0], 1.0 / t[:, 4]))


def chord(ax, ay, bx, by, x0, x1, y0, y1):
    """Length of the segments A-B inside [x0, x1] x [y0, y1] (float64, Liang-Barsky)."""
    dx, dy = bx - ax, by - ay
    t0, t1 = np.zeros_like(ax), np.ones_like(ax)
    for o, d, lo, hi in ((ax, dx, x0, x1), (ay, dy, y0, y1)):
        with np.errstate(divide="ignore", invalid="ignore"):
            ta, tb = (lo - o) / d, (hi - o) / d
        par = np.abs(d) < 1e-12
        inside = (o > lo) & (o < hi)
        t0 = np.where(par, np.where(inside, t0, 2.0), np.maximum(t0, np.minimum(ta, tb)))
        t1 = np.where(par, np.where(inside, t1, -1.0), np.minimum(t1, np.maximum(ta, tb)))
    return np.maximum(t1 - t0, 0.0) * np.hypot(dx, dy)


def box_scenario():
    mat = np.full(SHAPE, 2, np.int32)
    dens = np.ones(SHAPE, np.float32)
    act = np.zeros(SHAPE, np.float32)
    (xa, xb), (ya, yb) = ACT_BOX
    act[:, ya:yb, xa:xb] = A0
    cfg = phantoms.make_config(seed=5, threads_per_block=64, nrad=168, materials=[AIR, WATER])
    return Scenario(cfg, mat, dens, act, (VOX,) * 3, [AIR, WATER])


def expected_maps(sc, grid):
    """Closed-form maps of box_scenario on the grid, and a mask of the plane groups that lie
    clear of the box's axial faces."""
    cfg = sc.cfg
    o = sc.oracle("libm")
    D, H, W = grid
    N, nang, nrad = cfg.ncrystals, cfg.nangles, cfg.nrad
    gh, gw, gz = -(-nang // H), -(-nrad // W), -(-cfg.nzs // D)
    k1, k2, grp = [], [], []
    for a in range(N):
        for b in range(a + 1, N):
            s = o.sinogram_bin(a, 40, b, 40)          # a ring pair in the segment-0 sinograms
            if s < 0:
                continue
            ith, ir = (s // nrad) % nang, s % nrad
            k1.append(a)
            k2.append(b)
            grp.append((ith // gh) * W + ir // gw)
    k1, k2, grp = np.array(k1), np.array(k2), np.array(grp)
    bx, by, bz = SHAPE[2] * VOX, SHAPE[1] * VOX, SHAPE[0] * VOX
    cx, cy, cz, R = bx / 2, by / 2, bz / 2, abs(cfg.detector.radius)
    phi = lambda k: 2.0 * np.pi * (k + 0.5) / N - np.pi       # centre of crystal k (binning of K:1755)
    ax, ay = cx + R * np.cos(phi(k1)), cy + R * np.sin(phi(k1))
    bxx, byy = cx + R * np.cos(phi(k2)), cy + R * np.sin(phi(k2))
    l_box = chord(ax, ay, bxx, byy, 0.0, bx, 0.0, by)
    (xa, xb), (ya, yb) = ACT_BOX
    l_act = chord(ax, ay, bxx, byy, xa * VOX, xb * VOX, ya * VOX, yb * VOX)
    mu = mu_water_511()
    n = np.bincount(grp, minlength=H * W).astype(float)
    act_in = np.bincount(grp, A0 * l_act, H * W) / np.maximum(n, 1)
    att_in = np.bincount(grp, np.exp(-mu * l_box), H * W) / np.maximum(n, 1)
    # plane group d at the centre ring sum s = min(d gz + (gz - 1) / 2, NZS - 1): ring s / 2,
    # whose axial centre is ((s / 2 + 0.5) / NROWS - 0.5) FOVZ from the detector centre
    s = np.minimum(np.arange(D) * gz + 0.5 * (gz - 1), cfg.nzs - 1)
    z = cz + ((0.5 * s + 0.5) / cfg.nrows - 0.5) * cfg.fovz
    inside = (z > 0) & (z < bz)
    clear = np.abs(z) > 0.05
    clear &= np.abs(z - bz) > 0.05
    act = np.where(inside[:, None], act_in[None, :], 0.0).reshape(D, H, W)
    att = np.where(inside[:, None], att_in[None, :], 1.0).reshape(D, H, W)
    empty = (n == 0).reshape(H, W)
    return act, att, clear, empty, mu, inside


@pytest.mark.gpu
def test_projection_matches_the_closed_form_on_a_box_phantom():
    sc = box_scenario()
    act_x, att_x, clear, empty, mu, inside = expected_maps(sc, GRID)
    with mcpet.Simulation(sc.cfg, sc.phantom, sc.tables) as sim:
        assert sim.project(GRID, fetch=False) is None                    # computed and kept
        act, att = sim.project(GRID)
    with mcpet.Simulation(sc.cfg, sc.phantom, sc.tables) as sim:
        act2, att2 = sim.project(GRID)
    step = 0.5 * VOX
    assert 0.09 < mu < 0.1
    assert np.array_equal(act, act2) and np.array_equal(att, att2)       # deterministic
    assert inside.any() and (~inside).any()
    a, ax_, t, tx = act[clear], act_x[clear], att[clear], att_x[clear]
    # midpoint samples every `step`: within one step of the chord (activity, two sub-box faces)
    # and half a step of it (attenuation, the box's faces)
    assert np.abs(a - ax_).max() <= A0 * step * 1.02, np.abs(a - ax_).max()
    assert np.abs(t - tx).max() <= mu * 0.5 * step * 1.02 + 1e-6, np.abs(t - tx).max()
    # most bins are far closer than the bound; lines that miss the box see nothing
    assert np.median(np.abs(a - ax_)[ax_ > 0]) < 0.25 * A0 * step
    assert act_x.max() > 10 * A0 and np.all(act[:, empty] == 0) and np.all(att[:, empty] == 1)
    assert np.all(act[~inside] == 0) and np.all(att[~inside] == 1)


@pytest.mark.gpu
def test_batch_projection_equals_the_simulation_projection():
    """A batch's phantoms are projected as each would be alone (the box and two random
    phantoms of head tissues), bit for bit."""
    from mcpet import surrogate as S
    names = S._materials()
    sc = box_scenario()
    cfg = phantoms.make_config(seed=5, threads_per_block=64, nrad=168, materials=names)
    grid = (16, 28, 28)
    phs = [sc.phantom] + [mcpet.phantom_from_arrays(*sp) for sp in phantoms.random_phantoms(2, 9)]
    tab = mcpet.load_materials(names, np.max([p.density_max() for p in phs], axis=0))
    with mcpet.Batch(cfg, phs, tab) as b:
        got = [b.project(k, grid) for k in range(3)]
    for k in (0, 2):
        with mcpet.Simulation(cfg, phs[k], tab) as sim:
            alone = sim.project(grid)
        assert np.array_equal(got[k][0], alone[0]) and np.array_equal(got[k][1], alone[1]), k
    assert got[2][0].max() > 0 and got[2][1].min() < 1
    for p in phs[1:]:
        p.close()
    tab.close()


@pytest.mark.gpu
def test_projection_rejects_bad_grids():
    sc = box_scenario()
    with mcpet.Simulation(sc.cfg, sc.phantom, sc.tables) as sim:
        for g in [(0, 4, 4), (257, 4, 4), (4, 0, 4)]:
            with pytest.raises(mcpet.McpetError):
                sim.project(g)
