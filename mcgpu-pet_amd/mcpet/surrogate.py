"""Learned scatter surrogate: a low-count Monte Carlo run, rebinned on the device, and the
phantom's activity and attenuation seen along the same lines of response (mcpet_project_maps)
turned into a scatter estimate by the 3D U-Net (mcpet_surrogate_*), with a consistency check
that falls back to Monte Carlo. The reference only runs Monte Carlo (MCGPU-PET.cu:450-495); this module
has no counterpart there.

Grid (fixed before any training or measurement): the 168-angle x 168-radial x 1293-sinogram
michelogram (NRAD 168, NANGLES 168, NZS 159, SPAN 11, MRD 79) rebinned single-slice to 159
ring-sum planes, then summed in groups of 5 planes, 3 angles and 3 radial bins:
32 x 56 x 56 = 100352 bins (about 8 mm axially, 3.2 degrees, 6.3 mm radially per bin; the finest
grid on which a 1 %-noise Monte Carlo reference of each evaluation phantom stays affordable).

NRMSE(estimate, reference) = ||estimate - reference||_2 / ||reference||_2 over that grid, both
per simulated decay. The reference is a high-count Monte Carlo run whose own Poisson noise
floor sqrt(sum counts) / ||counts|| is reported beside every figure.

Time to 5 % NRMSE:
  Monte Carlo: the decays N5 = sum(r) / (sum(r^2) 0.05^2) at which a Monte Carlo estimate's
  expected NRMSE reaches 5 % (r = reference per decay; Poisson counts), simulated and timed
  (run, wait, rebin), its NRMSE measured against the reference.
  Surrogate: a Monte Carlo run of N_LOW decays, the device rebinning of its trues and
  scatter, the U-Net forward; if the consistency check fails (or its NRMSE exceeds 5 %) the
  time of the Monte Carlo fallback is added.
"""
import json
import os
import time

import numpy as np

from . import Batch, Simulation, Surrogate, load_materials, phantom_from_arrays, phantoms

GRID = (32, 56, 56)
N_LOW = 5.0e6              # decays: 1e7 photon histories
TARGET = 0.05
WIDTH = 32
IN_CHANNELS = 4                 # low-count scatter, low-count trues, activity integrals, attenuation factors
DECAYS_PER_SECOND = 1.0e9       # activity of the workload's phantoms: E[decays] = 1e9 per second
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(DATA_DIR, f"scatter_unet_w{WIDTH}.npy")


def stream_decays(n_decays):
    """Decays per stream of the balanced stream layout for a run of n_decays: about 2^20
    streams beyond the one every emitting voxel gets, at least 4 decays each."""
    return int(max(4, np.ceil(n_decays / 2.0 ** 20)))


def config(materials, n_decays, seed=20261202):
    """The head scanner (168 x 168 x 1293 sinogram) and an acquisition of n_decays expected
    decays for the workload's phantoms (1e9 decays per second), balanced stream layout (the
    same rule for the surrogate's input and for Monte Carlo)."""
    return phantoms.make_config(seed=seed, threads_per_block=1, acquisition_time_s=n_decays / DECAYS_PER_SECOND, nrad=168,
                                materials=materials, stream_decays=stream_decays(n_decays))


def nrmse(est, ref):
    return float(np.linalg.norm(est - ref) / np.linalg.norm(ref))


def poisson_floor(counts):
    """Expected NRMSE of a Poisson estimate with these expected counts: sqrt(sum) / ||counts||."""
    counts = np.asarray(counts, np.float64)
    return float(np.sqrt(counts.sum()) / np.linalg.norm(counts))


def decays_for(ref_per_decay, target=TARGET):
    r = np.asarray(ref_per_decay, np.float64)
    return float(r.sum() / ((r * r).sum() * target * target))


def normalize(scatter_low, trues_low, n_low, maps=None):
    """Network input and the factor that turns the network's output into scatter per decay.
    Channels: scatter and trues scaled so that the trues channel has mean 1; with maps =
    (activity integrals, attenuation factors) also the activity integrals divided by their
    maximum and the attenuation factors. The arithmetic is the device path's
    (mcpet_surrogate_estimate): counts times float32(voxels / total trues), the total exact in
    float64; activity times float32(1 / maximum); factor float32(total trues / (voxels * decays))."""
    s = np.asarray(scatter_low, np.float32)
    t = np.asarray(trues_low, np.float32)
    tot = float(np.sum(t, dtype=np.float64))
    nvox = t.size
    scale = np.float32(nvox / tot) if tot > 0 else np.float32(0.0)
    factor = np.float32(tot / (nvox * float(n_low))) if tot > 0 and n_low > 0 else np.float32(0.0)
    ch = [s * scale, t * scale]
    if maps is not None:
        act = np.asarray(maps[0], np.float32)
        att = np.asarray(maps[1], np.float32)
        mx = float(act.max()) if act.size else 0.0
        ch += [act * np.float32(1.0 / mx) if mx > 0 else act.copy(), att.copy()]
    return np.stack(ch), factor


def consistency_z(pred_per_decay, scatter_low, n_low, block=(4, 4, 4), min_mean=20.0):
    """Chi-square z-score of the observed low-count scatter against the prediction, over
    blocks of the grid with at least min_mean expected counts: z = (chi2 - dof) / sqrt(2 dof).
    Large z: the prediction disagrees with the data it was made from (fall back)."""
    mu = np.asarray(pred_per_decay, np.float64) * n_low
    n = np.asarray(scatter_low, np.float64)
    D, H, W = mu.shape
    b = block
    mu_b = mu[:D // b[0] * b[0], :H // b[1] * b[1], :W // b[2] * b[2]].reshape(D // b[0], b[0], H // b[1], b[1], W // b[2],
                                                                                   b[2]).sum(axis=(1, 3, 5))
    n_b = n[:D // b[0] * b[0], :H // b[1] * b[1], :W // b[2] * b[2]].reshape(D // b[0], b[0], H // b[1], b[1], W // b[2],
                                                                                b[2]).sum(axis=(1, 3, 5))
    m = mu_b >= min_mean
    dof = int(m.sum())
    if dof == 0:
        return float("inf")
    chi2 = float((((n_b - mu_b)[m]) ** 2 / mu_b[m]).sum())
    return (chi2 - dof) / np.sqrt(2.0 * dof)


def param_count(in_channels, width):
    """Parameters of the U-Net (mcpet_surrogate_create's layer list)."""
    c = width
    layers = [(in_channels, c, 3), (c, c, 3), (c, 2 * c, 3), (2 * c, 2 * c, 3), (2 * c, 4 * c, 3), (4 * c, 4 * c, 3),
              (6 * c, 2 * c, 3), (3 * c, c, 3), (c, 1, 1)]
    return sum(co * ci * k ** 3 + co for ci, co, k in layers)


def network_shape(n_params):
    """(in_channels, width) of the U-Net with n_params parameters (2 or 4 input channels,
    width 1..256); None if none fits."""
    for width in range(1, 257):
        for c in (IN_CHANNELS, 2):
            if param_count(c, width) == n_params:
                return c, width
    return None


class ScatterEstimator:
    """The trained U-Net and the estimate / fallback policy. in_channels None: 2 or 4, from the
    size of the weights (IN_CHANNELS without weights); width None: from the size of the weights
    as well."""

    def __init__(self, weights=WEIGHTS, width=WIDTH, grid=GRID, z_max=5.0, in_channels=None):
        if isinstance(weights, str):
            weights = np.load(weights, allow_pickle=False)
        if width is None:
            shape = network_shape(np.size(weights)) if weights is not None else (IN_CHANNELS, WIDTH)
            if shape is None:
                raise ValueError(f"{np.size(weights)} weights fit no U-Net of width 1-256")
            width = shape[1]
        self.width = int(width)
        if in_channels is None:
            in_channels = IN_CHANNELS
            if weights is not None:
                fits = [c for c in (2, 4) if param_count(c, width) == np.size(weights)]
                if not fits:
                    raise ValueError(f"{np.size(weights)} weights fit no 2- or 4-channel network of width {width}")
                in_channels = fits[0]
        self.in_channels = int(in_channels)
        self.net = Surrogate(grid, in_channels=self.in_channels, width=width, seed=1)
        if weights is not None:
            self.net.set_params(weights)
        self.grid = tuple(grid)
        self.z_max = z_max

    def predict(self, scatter_low, trues_low, n_low, maps=None):
        """Scatter per decay (float32) from rebinned low-count scatter and trues of n_low decays
        and (for a 4-channel network) the phantom's maps on the grid (Simulation.project)."""
        if (maps is not None) != (self.in_channels == 4):
            raise ValueError(f"a {self.in_channels}-channel network " + ("needs" if self.in_channels == 4 else "takes no")
                             + " maps")
        x, factor = normalize(scatter_low, trues_low, n_low, maps)
        y = self.net.forward(x)[0]
        return np.maximum(y, np.float32(0.0)) * factor

    def estimate(self, sim, seed=0):
        """Scatter per decay on the grid from a run of `sim` (a Simulation of the phantom with
        config(materials, N_LOW)): the run, then the fused device estimate (rebinning, maps,
        normalisation, network). Returns (estimate, info); info["consistent"] False means: use
        Monte Carlo. Equal to predict() on the rebinned tallies and maps, bit for bit."""
        t0 = time.perf_counter()
        if getattr(sim, "tally_grid", None) != self.grid:   # the run tallies on the network's grid
            sim.tally_rebinned(self.grid)
        sim.run(seed=seed)
        est, s, decays = self.net.estimate(sim)
        t1 = time.perf_counter()
        z = consistency_z(est, s, decays)
        return est, {"decays": decays, "z": z, "consistent": bool(z <= self.z_max), "seconds": t1 - t0,
                     "scatter_counts": float(s.sum())}

    def close(self):
        self.net.close()


class ScatterService:
    """Scatter estimates for many phantoms through the serving loop (mcpet_server_*): each
    request's low-count run is an estimate request (mcpet_server_submit_estimate), batched with
    the other requests' runs and answered on the device by the U-Net when it completes; the
    estimate is checked (consistency z-score) and, if it fails, the request is resubmitted as a
    rebinned Monte Carlo request of N5 decays (N5 from the prediction) whose scatter is the
    answer.

    Spot checks (spot_check_every = k > 0): every k-th request's estimate is also compared with
    a Monte Carlo run of N5 decays; the surrogate's own error is estimated as
    sqrt(max(0, NRMSE(estimate, MC)^2 - floor(MC)^2)) (the Monte Carlo run's Poisson floor
    removed). Above TARGET the request is answered by that Monte Carlo run, and after
    `disable_after` consecutive failed spot checks the surrogate is switched off: later
    requests are all answered by Monte Carlo (a prediction that merely copies its noisy input
    passes the consistency test but not a spot check). A request's answer does not depend
    on the other requests in flight."""

    def __init__(self, tables, estimator, budget_bytes=16 << 30, max_running=16, streams_per_step=1 << 22,
                 materials=None, spot_check_every=0, disable_after=2):
        from . import Server
        self.server = Server(tables, budget_bytes, max_running=max_running, streams_per_step=streams_per_step)
        self.est = estimator
        self.names = materials if materials is not None else _materials()
        self.jobs = {}          # service id -> dict
        self.by_rid = {}        # server request id -> service id
        self.next_id = 1
        self.latencies = []
        self.fallbacks = 0
        self.spot_every = int(spot_check_every)
        self.disable_after = int(disable_after)
        self.spot_checks = []   # estimated surrogate errors of the spot checks, in completion order
        self.failed_in_row = 0
        self.disabled = False

    def submit(self, phantom, seed=1):
        sid = self.next_id
        self.next_id += 1
        rid = self.server.submit_estimate(config(self.names, N_LOW), phantom, self.est.net, seed=seed)
        spot = self.spot_every > 0 and sid % self.spot_every == 0
        self.jobs[sid] = {"phantom": phantom, "seed": seed, "stage": "surrogate", "t0": time.perf_counter(),
                          "result": None, "info": None, "spot": spot}
        self.by_rid[rid] = sid
        return sid

    def _monte_carlo(self, sid, job, n5, stage):
        job["stage"] = stage
        job["info"]["n5"] = n5
        rid = self.server.submit_rebinned(config(self.names, n5), job["phantom"], self.est.grid, seed=job["seed"])
        self.by_rid[rid] = sid

    def _complete(self, rid):
        sid = self.by_rid.pop(rid)
        job = self.jobs[sid]
        if job["stage"] == "surrogate":
            est, s, decays = self.server.fetch_estimate(rid)
            z = consistency_z(est, s, decays)
            job["info"] = {"decays": decays, "z": z, "consistent": bool(z <= self.est.z_max)}
            n5 = decays_for(est) if est.sum() > 0 else 1e9
            if not job["info"]["consistent"] or self.disabled:
                # fall back: Monte Carlo to the decays a 5 % estimate needs
                self.fallbacks += 1
                self._monte_carlo(sid, job, n5, "monte_carlo")
                return
            job["result"] = est
            if job["spot"]:
                self._monte_carlo(sid, job, n5, "spot_check")
                return
        else:
            _, s, decays, _ = self.server.fetch_rebinned(rid)
            mc = s.astype(np.float64) / decays
            job["info"]["mc_decays"] = decays
            if job["stage"] == "spot_check":
                e = nrmse(job["result"], mc)
                floor = poisson_floor(s)
                err = float(np.sqrt(max(0.0, e * e - floor * floor)))
                job["info"].update({"spot_nrmse": e, "spot_floor": floor, "spot_error": err})
                self.spot_checks.append(err)
                if err > TARGET:
                    job["result"] = mc
                    job["info"]["spot_failed"] = True
                    self.fallbacks += 1
                    self.failed_in_row += 1
                    if self.failed_in_row >= self.disable_after:
                        self.disabled = True
                else:
                    self.failed_in_row = 0
            else:
                job["result"] = mc
        job["seconds"] = time.perf_counter() - job["t0"]
        self.latencies.append(job["seconds"])

    def step(self):
        self.server.step()
        for rid in list(self.by_rid):
            state = self.server.poll(rid)
            if state == "done":
                self._complete(rid)
            elif state in ("failed", "rejected"):
                self._fail(rid)

    def _fail(self, rid):
        """A request the serving loop could not answer: the job ends with no result and the
        library's error in its info."""
        sid = self.by_rid.pop(rid)
        job = self.jobs[sid]
        from . import McpetError
        try:
            self.server.fetch_rebinned(rid) if job["stage"] != "surrogate" else self.server.fetch_estimate(rid)
            err = "unknown"
        except McpetError as e:
            err = str(e)
        job["result"] = None
        job["info"] = dict(job["info"] or {}, error=err)
        job["stage"] = "failed"
        job["seconds"] = time.perf_counter() - job["t0"]

    def drain(self):
        while self.by_rid:
            self.step()

    def result(self, sid):
        job = self.jobs.pop(sid)
        return job["result"], job["info"]

    def stats(self):
        lat = np.array(self.latencies) * 1e3
        out = {"completed": len(self.latencies), "fallbacks": self.fallbacks, "spot_checks": list(self.spot_checks),
               "disabled": self.disabled, "server": self.server.stats()}
        if len(lat):
            out.update({"latency_p50_ms": float(np.percentile(lat, 50)), "latency_p99_ms": float(np.percentile(lat, 99))})
        return out

    def close(self):
        self.server.close()


def monte_carlo(sim, grid=GRID, seed=0):
    """Scatter per decay on the grid from a Monte Carlo run of `sim` (its configured
    acquisition); (estimate, decays, seconds, counts)."""
    t0 = time.perf_counter()
    sim.run(seed=seed)
    s = sim.rebin("scatter", grid)
    decays = sim.stats()["decays"]
    dt = time.perf_counter() - t0
    return s.astype(np.float64) / decays, decays, dt, s


# ------------------------------------------------------------------ training data and training

def _materials():
    here = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    return phantoms.head_materials(os.path.join(here, "tests", "golden", "sample_simulation", "materials"))


def family_phantoms(family, n, seed):
    """n phantoms of a family with activity scaled to DECAYS_PER_SECOND expected decays per
    second: "scatter" (scatter_phantoms, 128x128x64), "batch" (random_phantoms, 32-160 x 32-160 x
    32-96 voxels, random ellipsoids), "head" (seed-jittered head phantoms, 128x128x96)."""
    if family == "scatter":
        return phantoms.scatter_phantoms(n, seed, n_decays=DECAYS_PER_SECOND, acquisition_time_s=1.0)
    if family == "head":
        return [phantoms.head(seed=seed + k, n_decays=DECAYS_PER_SECOND, acquisition_time_s=1.0) for k in range(n)]
    if family == "batch":
        frac = phantoms.F18_MEAN_LIFE * (1.0 - np.exp(-1.0 / phantoms.F18_MEAN_LIFE))
        out = []
        for mat, dens, act, vs in phantoms.random_phantoms(n, seed):
            act = (act.astype(np.float64) * (DECAYS_PER_SECOND / (act.sum(dtype=np.float64) * frac))).astype(np.float32)
            out.append((mat, dens, act, vs))
        return out
    raise ValueError(f"unknown phantom family {family}")


def generate(n_phantoms, seed, n_target=5.0e8, n_low=N_LOW, inputs_per_phantom=4, chunk=16, families=("scatter",),
             in_channels=IN_CHANNELS, log=print):
    """Training pairs from n_phantoms phantoms split evenly over `families` (family_phantoms,
    seeds seed, seed + 1, ...): per phantom one target run of n_target decays (scatter per
    decay) and inputs_per_phantom independent runs of n_low decays (normalised scatter +
    trues, with in_channels 4 also the phantom's maps). Returns X [n, in_channels, D, H, W],
    Y [n, D, H, W] (targets in the inputs' normalisation) and the phantom index of each sample
    (phantoms interleaved over the families)."""
    names = _materials()
    per = [n_phantoms // len(families) + (1 if i < n_phantoms % len(families) else 0) for i in range(len(families))]
    lists = [family_phantoms(f, k, seed + i) for i, (f, k) in enumerate(zip(families, per))]
    specs = [sp for group in zip(*[l + [None] * (max(per) - len(l)) for l in lists]) for sp in group if sp is not None]
    X, Y, idx = [], [], []
    for c0 in range(0, n_phantoms, chunk):
        part = specs[c0:c0 + chunk]
        phs = [phantom_from_arrays(*sp) for sp in part]
        dmax = np.max([p.density_max() for p in phs], axis=0)
        tab = load_materials(names, dmax)
        targets = []
        with Batch(config(names, n_target), phs, tab) as b:
            b.run(seed=int(seed * 1000 + c0 + 1))
            b.sync()
            for k in range(len(phs)):
                d = b.stats(k)["decays"]
                targets.append(b.rebin(k, "scatter", GRID).astype(np.float64) / d)
        with Batch(config(names, n_low), phs, tab) as b:
            maps = [b.project(k, GRID) if in_channels == 4 else None for k in range(len(phs))]
            for r in range(inputs_per_phantom):
                b.run(seed=int(seed * 1000 + c0 + 101 + r))
                b.sync()
                for k in range(len(phs)):
                    d = b.stats(k)["decays"]
                    x, factor = normalize(b.rebin(k, "scatter", GRID), b.rebin(k, "trues", GRID), d, maps[k])
                    X.append(x)
                    Y.append((targets[k] / np.float64(factor)).astype(np.float32))
                    idx.append(c0 + k)
        for p in phs:
            p.close()
        tab.close()
        log(f"generate: phantoms {c0}..{c0 + len(part) - 1} done")
    return np.stack(X), np.stack(Y), np.array(idx)


def train(X, Y, idx, width=WIDTH, epochs=150, batch=8, lr=5e-4, warmup=200, val_fraction=0.125, seed=1, log=print,
          checkpoint=None):
    """Adam on the MSE, learning rate warmed up linearly over `warmup` steps then decayed on a
    cosine; the parameters of the epoch with the lowest validation loss (phantoms held out by
    index) are kept, and written to `checkpoint` (a .npy path) whenever they improve."""
    in_channels = X.shape[1]
    rng = np.random.default_rng(seed)
    phantom_ids = np.unique(idx)
    n_val = max(1, int(len(phantom_ids) * val_fraction))
    val_ids = set(phantom_ids[-n_val:].tolist())
    val = np.array([i in val_ids for i in idx])
    tr = np.flatnonzero(~val)
    va = np.flatnonzero(val)
    net = Surrogate(GRID, in_channels=in_channels, width=width, seed=seed)
    history = []
    best = (np.inf, None, -1)
    steps = epochs * max(1, len(tr) // batch)
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = rng.permutation(tr)
        tl = []
        for j in range(0, len(perm) - batch + 1, batch):
            b = perm[j:j + batch]
            cur = lr * min(1.0, (step + 1) / warmup) * 0.5 * (1.0 + np.cos(np.pi * step / steps))
            tl.append(net.train_step(X[b], Y[b], lr=max(cur, lr * 0.01)))
            step += 1
            if ep == 0 and step % 10 == 0:
                log(f"train: step {step} loss {tl[-1]:.5f}")
        vl = float(np.mean([net.gradients(X[va[j:j + batch]], Y[va[j:j + batch]])[0]
                            for j in range(0, len(va), batch)]))
        vn = float(np.mean([nrmse(net.forward(X[i:i + 1])[0], Y[i]) for i in va]))
        history.append({"epoch": ep, "train_loss": float(np.mean(tl)), "val_loss": vl, "val_nrmse": vn})
        if vl < best[0]:
            best = (vl, net.params(), ep)
            if checkpoint:
                np.save(checkpoint, best[1].astype(np.float32))
        if ep < 5 or ep % 5 == 0 or ep == epochs - 1:
            log(f"train: epoch {ep} train {np.mean(tl):.5f} val {vl:.5f} val nrmse {vn:.4f} "
                f"({time.perf_counter() - t0:.1f} s)")
    net.close()
    return best[1], {"best_epoch": best[2], "best_val_loss": best[0], "history": history,
                     "train_samples": int(len(tr)), "val_samples": int(len(va)), "seconds": time.perf_counter() - t0}


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train the scatter surrogate on simulator data")
    ap.add_argument("--phantoms", type=int, default=96)
    ap.add_argument("--families", default="scatter", help="comma-separated phantom families (scatter, batch, head)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--n-target", type=float, default=5.0e8)
    ap.add_argument("--inputs-per-phantom", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--width", type=int, default=WIDTH)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--in-channels", type=int, default=IN_CHANNELS, choices=(2, 4))
    ap.add_argument("--out", default="surrogate_out")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    log = lambda m: print(m, flush=True)
    t0 = time.perf_counter()
    fams = tuple(a.families.split(","))
    X, Y, idx = generate(a.phantoms, a.seed, n_target=a.n_target, inputs_per_phantom=a.inputs_per_phantom, families=fams,
                         in_channels=a.in_channels, log=log)
    gen_s = time.perf_counter() - t0
    log(f"generate: {len(X)} samples in {gen_s:.1f} s")
    w, info = train(X, Y, idx, width=a.width, epochs=a.epochs, lr=a.lr, log=log,
                    checkpoint=os.path.join(a.out, f"scatter_unet_w{a.width}.npy"))
    np.save(os.path.join(a.out, f"scatter_unet_w{a.width}.npy"), w.astype(np.float32))
    info.update({"phantoms": a.phantoms, "families": fams, "phantom_seed": a.seed, "n_target": a.n_target, "n_low": N_LOW,
                 "inputs_per_phantom": a.inputs_per_phantom, "grid": GRID, "width": a.width, "in_channels": a.in_channels,
                 "generate_seconds": gen_s})
    with open(os.path.join(a.out, f"scatter_unet_w{a.width}.json"), "w") as f:
        json.dump(info, f)
    log(f"saved {a.out}/scatter_unet_w{a.width}.npy (best epoch {info['best_epoch']})")


if __name__ == "__main__":
    main()
