"""Learned scatter surrogate (no reference counterpart): the BF16 MFMA GEMM that carries every
convolution, checked against float64 products of the BF16-rounded operands."""
import numpy as np
import pytest

import mcpet


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16
    return b.astype(np.uint32).view(np.float32)


def test_bf16_rounding_helper():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415927, 65504.0], np.float32)
    r = bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125)   # ties to even, then up
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K,split", [(16, 16, 32, 1), (32, 64, 32, 1), (37, 100, 77, 1), (5, 3000, 1000, 4),
                                         (64, 40, 2000, 7)])
def test_mfma_gemm_matches_float64_of_bf16_operands(M, N, K, split):
    rng = np.random.default_rng(M * 1000 + N + K)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    c = mcpet.selftest_gemm_bf16(a, b, split)
    ref = bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)
    err = np.abs(c - ref).max() / (np.abs(bf16(a)).astype(np.float64) @ np.abs(bf16(b)).astype(np.float64)).max()
    assert err < 1e-5, err


# ------------------------------------------------------------------ rebinning grid

def test_sinogram_planes_match_the_binning_of_every_ring_pair():
    """plane_of_sinogram[izm] is the ring sum iz1 + iz2 of every ring pair the oracle's
    binning (K:1749-1788) puts into sinogram izm, for every accepted pair of rings."""
    from harness import water_cube_scenario
    sc = water_cube_scenario(n_decays=1e3)
    o = sc.oracle("libm")
    cfg = sc.cfg
    planes = mcpet.sinogram_planes(cfg)
    nang, nrad, nrows = cfg.nangles, cfg.nrad, cfg.nrows
    seen = np.zeros(cfg.nsinos, bool)
    for iz1 in range(nrows):
        for iz2 in range(nrows):
            b = o.sinogram_bin(10, iz1, 10 + nang, iz2)
            if abs(iz1 - iz2) > cfg.mrd:
                assert b < 0
                continue
            izm = b // (nang * nrad)
            assert planes[izm] == iz1 + iz2, (iz1, iz2, izm)
            seen[izm] = True
    assert seen.all()        # every sinogram holds some ring pair
    assert planes.min() == 0 and planes.max() == 2 * nrows - 2


# ------------------------------------------------------------------ U-Net reference (torch, CPU)

SPEC = lambda cin, c: [(cin, c, 3, 1), (c, c, 3, 1), (c, 2 * c, 3, 2), (2 * c, 2 * c, 3, 1), (2 * c, 4 * c, 3, 2),
                       (4 * c, 4 * c, 3, 1), (6 * c, 2 * c, 3, 1), (3 * c, c, 3, 1), (c, 1, 1, 1)]


def n_params(cin, c):
    return sum(co * ci * k ** 3 + co for ci, co, k, _ in SPEC(cin, c))


def unet_reference(params, x, target, cin, c, mode="bf16", grad=True):
    """Loss, parameter gradients and output of the U-Net by torch autograd on the CPU.
    mode "bf16": every convolution (forward and both backward products) rounds its two operands
    to BF16 and accumulates in float64, as the device GEMM does (a custom autograd function
    with torch's conv3d_input / conv3d_weight); "bf16f32": the same rounding with float32
    accumulation and float32 activations (the device's precision, in another summation order);
    "f32": the function without rounding, in float32; "f64": torch's own conv3d and autograd
    in float64."""
    import torch
    dt = torch.float32 if mode in ("f32", "bf16f32") else torch.float64
    rb = (lambda t: t.to(torch.bfloat16).to(dt)) if mode in ("bf16", "bf16f32") else (lambda t: t)

    class Conv(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xx, w, b, stride, pad):
            ctx.save_for_backward(xx, w)
            ctx.stride, ctx.pad = stride, pad
            return torch.nn.functional.conv3d(rb(xx), rb(w), b, stride=stride, padding=pad)

        @staticmethod
        def backward(ctx, gy):
            xx, w = ctx.saved_tensors
            gx = torch.nn.grad.conv3d_input(xx.shape, rb(w), rb(gy), stride=ctx.stride, padding=ctx.pad)
            gw = torch.nn.grad.conv3d_weight(rb(xx), w.shape, rb(gy), stride=ctx.stride, padding=ctx.pad)
            return gx, gw, gy.sum(dim=(0, 2, 3, 4)), None, None

    p = torch.tensor(params, dtype=dt, requires_grad=grad)
    ws, off = [], 0
    for ci, co, k, s in SPEC(cin, c):
        w = p[off:off + co * ci * k ** 3].view(co, ci, k, k, k)
        off += co * ci * k ** 3
        ws.append((w, p[off:off + co], s, k // 2))
        off += co
    if mode == "f64":
        conv = lambda i, t: torch.nn.functional.conv3d(t, ws[i][0], ws[i][1], stride=ws[i][2], padding=ws[i][3])
    else:
        conv = lambda i, t: Conv.apply(t, ws[i][0], ws[i][1], ws[i][2], ws[i][3])
    relu = torch.relu
    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    with torch.enable_grad() if grad else torch.no_grad():
        X = torch.tensor(x, dtype=dt)
        a1 = relu(conv(0, X))
        b1 = relu(conv(1, a1))
        d1 = relu(conv(2, b1))
        b2 = relu(conv(3, d1))
        d2 = relu(conv(4, b2))
        b3 = relu(conv(5, d2))
        e2 = relu(conv(6, torch.cat([up(b3), b2], 1)))
        e1 = relu(conv(7, torch.cat([up(e2), b1], 1)))
        out = X[:, 0] + conv(8, e1)[:, 0]
        loss = ((out - torch.tensor(target, dtype=dt)) ** 2).mean()
        if grad:
            loss.backward()
    g = p.grad.numpy().astype(np.float64) if grad else None
    return float(loss.detach()), g, out.detach().numpy()


def _layers(cin, c):
    o, out = 0, []
    for ci, co, k, _ in SPEC(cin, c):
        nw = co * ci * k ** 3
        out.append((slice(o, o + nw), slice(o + nw, o + nw + co)))
        o += nw + co
    return out


def test_unet_reference_gradient_by_finite_differences():
    """torch's float64 autograd of the network against central differences of its loss, on
    three parameters of every layer."""
    rng = np.random.default_rng(3)
    cin, c, grid = 2, 2, (4, 4, 4)
    p = rng.standard_normal(n_params(cin, c)) * 0.3
    x = rng.random((2, cin) + grid)
    t = rng.random((2,) + grid)
    _, g, _ = unet_reference(p, x, t, cin, c, "f64")
    for sw, sb in _layers(cin, c):
        for idx in (sw.start, (sw.start + sw.stop) // 2, sb.start):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[idx] += h
            pm[idx] -= h
            fd = (unet_reference(pp, x, t, cin, c, "f64", grad=False)[0] -
                  unet_reference(pm, x, t, cin, c, "f64", grad=False)[0]) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 + 1e-5 * abs(fd), (idx, fd, g[idx])


def test_rounding_reference_backward_matches_torch_autograd():
    """The custom convolution backward (the BF16 reference's) equals torch's autograd without
    rounding (float32 vs float64), and with BF16 operands it moves every layer's gradient by a
    BF16-sized amount, not more."""
    rng = np.random.default_rng(4)
    cin, c, grid = 2, 3, (8, 4, 8)
    p = rng.standard_normal(n_params(cin, c)) * 0.3
    x = rng.random((2, cin) + grid)
    t = rng.random((2,) + grid)
    l64, g64, o64 = unet_reference(p, x, t, cin, c, "f64")
    l32, g32, o32 = unet_reference(p, x, t, cin, c, "f32")
    lb, gb, ob = unet_reference(p, x, t, cin, c, "bf16")
    assert abs(l32 - l64) <= 1e-5 * l64
    for sw, sb in _layers(cin, c):
        for sl in (sw, sb):
            ref = np.linalg.norm(g64[sl])
            assert np.linalg.norm(g32[sl] - g64[sl]) <= 1e-4 * ref
            rel = np.linalg.norm(gb[sl] - g64[sl]) / ref
            assert 1e-6 < rel < 0.25, rel


# ------------------------------------------------------------------ device U-Net (GPU tier)

def _layer_rel(a, b, cin, c):
    return [max(np.linalg.norm(a[s] - b[s]) / max(np.linalg.norm(b[s]), 1e-30) for s in (sw, sb)) for sw, sb in _layers(cin, c)]


@pytest.mark.gpu
@pytest.mark.parametrize("grid,cin,c,batch", [((8, 8, 8), 2, 4, 2), ((12, 8, 16), 2, 3, 1), ((4, 4, 4), 1, 2, 3),
                                              ((8, 12, 12), 4, 32, 2), ((8, 8, 8), 4, 56, 1), ((4, 8, 8), 4, 16, 1),
                                              ((16, 64, 64), 4, 32, 1)])
def test_device_unet_forward_and_gradients_match_bf16_reference(grid, cin, c, batch):
    """Output, loss and every layer's weight and bias gradients of the device network against
    the BF16-operand CPU reference (float64 accumulation). The remaining differences come
    from float32 accumulation order and from activations whose float32 and float64 values
    round to different BF16 neighbours: bounded by 2e-4 (output, loss) and 2e-3 (gradients,
    relative per layer). Widths 2-4 use one 16-row tile per product; widths 16, 32 and 56 with
    four input channels (the shipped network's layer shapes and a ~4 M-parameter one) take the
    2-, 3- and 4-tile convolution kernels, several row blocks per product and the epilogue split
    over the workgroup's waves, in the forward, input-gradient and weight-gradient modes; the
    stride-1 layers' weight gradients take the LDS-tiled product (32-channel blocks, partial
    blocks at widths 16 and 56, voxel tiles cut by the grid's edges). On the 16 x 64 x 64 grid
    (1024 four-row tiles per plane set) the full-resolution forward and input-gradient products
    run on their LDS tiles (`k_conv_tiled`), as on the shipped 32 x 56 x 56 grid."""
    rng = np.random.default_rng(sum(grid) + cin + c)
    with mcpet.Surrogate(grid, in_channels=cin, width=c, seed=7) as s:
        assert s.n_params == n_params(cin, c)
        p = s.params().astype(np.float64)
        x = rng.random((batch, cin) + grid).astype(np.float32)
        t = rng.random((batch,) + grid).astype(np.float32)
        # as created: He-normal weights, zero biases, zero output layer, so the output is input channel 0
        layers = _layers(cin, c)
        for sw, sb in layers:
            assert np.all(p[sb] == 0.0)
        assert np.all(p[layers[-1][0]] == 0.0) and np.array_equal(s.forward(x), x[:, 0])
        # give the biases and the output layer values to check every gradient
        for sw, sb in layers:
            p[sb] = rng.standard_normal(sb.stop - sb.start) * 0.05
        p[layers[-1][0]] = rng.standard_normal(layers[-1][0].stop - layers[-1][0].start) * 0.3
        s.set_params(p.astype(np.float32))
        with mcpet.Surrogate(grid, in_channels=cin, width=c, seed=1) as fresh:
            with pytest.raises(mcpet.McpetError):
                fresh.forward_ms()              # no forward yet
        y = s.forward(x)
        ms = s.forward_ms()                     # HIP events around the last sample's network
        assert 0.0 < ms < 1000.0
        loss, g = s.gradients(x, t)
        lr, gr, yr = unet_reference(p.astype(np.float32).astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                                    cin, c, "bf16")
    # the same rounding with float32 accumulation, in torch's order: how far an implementation
    # with the device's precision lands from the float64 reference through BF16 rounding flips
    l32, g32, y32 = unet_reference(p.astype(np.float32).astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                                   cin, c, "bf16f32")
    flip_y = np.linalg.norm(y32 - yr) / np.linalg.norm(yr)
    assert np.linalg.norm(y - yr) <= max(2e-4, 4 * flip_y) * np.linalg.norm(yr), flip_y
    assert abs(loss - lr) <= max(2e-4, 4 * abs(l32 - lr) / lr) * lr
    rel = _layer_rel(g.astype(np.float64), gr, cin, c)
    rel32 = _layer_rel(g32, gr, cin, c)
    assert all(r <= max(2e-3, 4 * r32) for r, r32 in zip(rel, rel32)), (rel, rel32)


@pytest.mark.gpu
def test_packaged_network_forward_matches_bf16_reference():
    """The shipped network (packaged weights, 4 input channels, width 32) on its full 32 x 56 x
    56 grid: the device forward equals the BF16-operand CPU forward (float64 accumulation) within
    the rounding-flip distance of a float32-accumulating implementation (at least 2e-4)."""
    from mcpet import surrogate as sg
    p = np.load(sg.WEIGHTS, allow_pickle=False).astype(np.float32)
    rng = np.random.default_rng(12)
    x = rng.random((1, sg.IN_CHANNELS) + sg.GRID).astype(np.float32)
    with mcpet.Surrogate(sg.GRID, in_channels=sg.IN_CHANNELS, width=sg.WIDTH, seed=1) as s:
        assert s.n_params == p.size == n_params(sg.IN_CHANNELS, sg.WIDTH)
        s.set_params(p)
        y = s.forward(x)
        assert np.array_equal(s.forward(x), y)           # deterministic
    ref = {m: unet_reference(p.astype(np.float64), x.astype(np.float64), np.zeros((1,) + sg.GRID), sg.IN_CHANNELS,
                             sg.WIDTH, m, grad=False)[2] for m in ("bf16", "bf16f32")}
    rel = np.linalg.norm(y - ref["bf16"]) / np.linalg.norm(ref["bf16"])
    flip = np.linalg.norm(ref["bf16f32"] - ref["bf16"]) / np.linalg.norm(ref["bf16"])
    print(f"packaged network forward: relative difference {rel:.2e} (float32-accumulation reference: {flip:.2e})",
          flush=True)
    assert rel <= max(2e-4, 4 * flip), (rel, flip)


def integer_network(cin, c, rng, taps=3):
    """Parameters for which the whole forward is exact integer arithmetic: inputs are small
    integers, every output channel of every layer has `taps` weights of +-1 at random (input
    channel, tap) positions and an integer bias in {0, 1}, all other weights 0. Every
    pre-activation is then an integer below 2^24 (exact in float32 in any summation order) and
    every BF16 rounding rounds an exact integer, so a correct implementation equals the float64
    reference bit for bit and any misplaced weight, tap or channel shows."""
    p = np.zeros(n_params(cin, c))
    off = 0
    for ci, co, k, _ in SPEC(cin, c):
        T = k ** 3
        w = np.zeros((co, ci * T))
        for o in range(co):
            pos = rng.choice(ci * T, size=min(taps, ci * T), replace=False)
            w[o, pos] = rng.choice([-1.0, 1.0], size=len(pos))
            w[o, pos[0]] = 1.0
        p[off:off + co * ci * T] = w.ravel()
        off += co * ci * T
        p[off:off + co] = rng.integers(0, 2, co)
        off += co
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("grid,c,batch", [((8, 12, 12), 32, 2), ((8, 8, 8), 56, 1), ((4, 8, 8), 16, 1),
                                          ((32, 56, 56), 32, 1)])
def test_device_unet_forward_is_exact_on_integer_networks(grid, c, batch):
    """Forward of the device network on an integer network (integer_network): equal to the
    float64 reference bit for bit, at the shipped width on its full grid and at widths 16 and 56,
    so every forward tile shape (k_conv<0, 1..4>), several row blocks per product and the
    epilogue's channels-last copies are pinned exactly."""
    cin = 4
    rng = np.random.default_rng(c * 7 + sum(grid))
    p = integer_network(cin, c, rng)
    x = rng.integers(0, 4, (batch, cin) + grid).astype(np.float32)
    with mcpet.Surrogate(grid, in_channels=cin, width=c, seed=3) as s:
        s.set_params(p.astype(np.float32))
        y = s.forward(x)
    _, _, yr = unet_reference(p, x.astype(np.float64), np.zeros((batch,) + grid), cin, c, "bf16", grad=False)
    assert np.abs(yr).max() < 2.0 ** 24 and np.all(yr == np.round(yr))
    assert np.array_equal(y.astype(np.float64), yr), np.abs(y - yr).max()


@pytest.mark.gpu
def test_device_training_reduces_loss_on_a_denoising_task():
    """Adam on a toy task of the surrogate's kind (recover a smooth field from a Poisson
    sample of it): the loss falls by more than half in 60 steps, and train_step reports the
    loss of the parameters it starts from (equal to gradients() there)."""
    rng = np.random.default_rng(11)
    grid = (8, 8, 8)
    z, y, x = np.meshgrid(*[np.arange(n) for n in grid], indexing="ij")

    def sample():
        field = np.zeros(grid)
        for _ in range(3):
            cz, cy, cx = rng.uniform(0, 8, 3)
            field += np.exp(-((z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2) / (2 * rng.uniform(1.5, 3) ** 2))
        field = field / field.mean()
        noisy = rng.poisson(field * 4) / 4.0
        return np.stack([noisy, field * 0 + noisy.mean()]).astype(np.float32), field.astype(np.float32)

    data = [sample() for _ in range(16)]
    X = np.stack([d[0] for d in data])
    T = np.stack([d[1] for d in data])
    with mcpet.Surrogate(grid, in_channels=2, width=4, seed=3) as s:
        l0, _ = s.gradients(X, T)
        first = s.train_step(X, T, lr=3e-3)
        assert abs(first - l0) <= 1e-6 * l0
        for _ in range(59):
            s.train_step(X, T, lr=3e-3)
        l1, _ = s.gradients(X, T)
    assert l1 < 0.5 * l0, (l0, l1)


@pytest.mark.gpu
def test_device_rebinning_equals_host_sum_of_the_fetched_sinograms():
    """The device rebinning of trues and scatter equals the sum of the fetched int32
    sinograms over the same plane, angle and radial groups (exact: integer counts)."""
    from harness import water_cube_scenario
    sc = water_cube_scenario(n_decays=3e5, tpb=64)
    cfg = sc.cfg
    planes = mcpet.sinogram_planes(cfg)
    with mcpet.Simulation(cfg, sc.phantom, sc.tables) as sim:
        sim.run()
        out = sim.fetch(images=False, psf=False, dose=False)
        for grid in ((32, 24, 24), (7, 5, 9), (159, 168, 147)):
            D, H, W = grid
            gz, gh, gw = -(-cfg.nzs // D), -(-cfg.nangles // H), -(-cfg.nrad // W)
            for which in ("trues", "scatter"):
                dev = sim.rebin(which, grid)
                sino = out[which].reshape(cfg.nsinos, cfg.nangles, cfg.nrad).astype(np.int64)
                red = np.zeros((D, cfg.nangles, cfg.nrad), np.int64)
                np.add.at(red, planes // gz, sino)
                red = np.pad(red, ((0, 0), (0, gh * H - cfg.nangles), (0, gw * W - cfg.nrad)))
                ref = red.reshape(D, H, gh, W, gw).sum(axis=(2, 4))
                assert np.array_equal(dev.astype(np.int64), ref), (which, grid)
                assert dev.sum() == sino.sum() > 0


@pytest.mark.gpu
def test_device_rebinning_with_more_than_256_radial_bins():
    """NRAD > 256 takes the LDS-atomic rebinning kernel (the row kernel holds one radial bin
    per thread): equal to the host sum as well."""
    from harness import water_cube_scenario
    sc = water_cube_scenario(n_decays=1e5, tpb=16, nrad=300)
    cfg = sc.cfg
    planes = mcpet.sinogram_planes(cfg)
    with mcpet.Simulation(cfg, sc.phantom, sc.tables) as sim:
        sim.run(seed=5)
        out = sim.fetch(images=False, psf=False, dose=False)
        grid = (16, 12, 30)
        D, H, W = grid
        gz, gh, gw = -(-cfg.nzs // D), -(-cfg.nangles // H), -(-cfg.nrad // W)
        dev = sim.rebin("trues", grid)
    sino = out["trues"].reshape(cfg.nsinos, cfg.nangles, cfg.nrad).astype(np.int64)
    red = np.zeros((D, cfg.nangles, cfg.nrad), np.int64)
    np.add.at(red, planes // gz, sino)
    red = np.pad(red, ((0, 0), (0, gh * H - cfg.nangles), (0, gw * W - cfg.nrad)))
    assert np.array_equal(dev.astype(np.int64), red.reshape(D, H, gh, W, gw).sum(axis=(2, 4))) and dev.sum() > 0
