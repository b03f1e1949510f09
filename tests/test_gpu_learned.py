"""GPU tests of the learned psi extrapolation: the BF16 MFMA convolution kernels (forward,
backward data, backward weights) against the CPU reference (tests/cnn_ref.py) with the same bf16
roundings, the whole network's gradient against the reference, training on seeded trajectories,
the gain over plain Lagrange extrapolation on held-out seeds, and the fallback rule in the loop.

Bounds: forward and backward kernels within 1e-5 relative (FP32 accumulation of bf16 products,
288-term sums; bf16 outputs within one bf16 rounding); the network gradient within 2e-2 relative
in norm (bf16 storage of activations and gradients between layers rounds differently where the
accumulation order differs)."""
import math

import numpy as np
import pytest

import cnn_ref as R

pytestmark = pytest.mark.gpu


def dev16(a):
    import torch
    return torch.from_numpy(R.bf16_bits(np.ravel(a)).view(np.int16).copy()).cuda()


def dev32(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host16(t):
    import torch
    torch.cuda.synchronize()
    return R.from_bits(t.cpu().numpy().view(np.uint16))


def vp(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def stream():
    import ctypes, torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_conv_kernels_match_cpu(dev):
    import cmm
    import torch
    rng = np.random.default_rng(1)
    N, H, W = 2, 6, 128
    x = R.bf16(rng.standard_normal((N, H, W, 32)))
    w = R.bf16(rng.standard_normal((32, 9, 32)) * 0.1)
    b = rng.standard_normal(32).astype(np.float32).astype(np.float64)
    L = cmm.lib()
    xin, wd, bd = dev16(x), dev16(w), dev32(b)
    ref = R.conv(x, w, b)
    out32 = torch.zeros(N * H * W * 32, dtype=torch.float32, device="cuda")
    cmm._check(L.cmm_cnn_forward(N, H, W, vp(xin), vp(wd), vp(bd), 2, vp(out32), stream()), "forward")
    got = out32.cpu().numpy().reshape(ref.shape)
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print("forward fp32 out: rel err", e)
    assert e <= 1e-5
    out16 = torch.zeros(N * H * W * 32, dtype=torch.int16, device="cuda")
    cmm._check(L.cmm_cnn_forward(N, H, W, vp(xin), vp(wd), vp(bd), 0, vp(out16), stream()), "forward relu")
    got = host16(out16).reshape(ref.shape)
    refr = np.maximum(ref, 0)
    assert np.all(np.abs(got - refr) <= 2.0 ** -8 * np.abs(refr) + 1e-5 * np.abs(ref).max())
    # backward data with the transposed, flipped weights; backward weights and bias
    g = R.bf16(rng.standard_normal((N, H, W, 32)))
    w32 = dev32(np.concatenate([w.ravel()]))
    wb = torch.zeros(32 * 9 * 32, dtype=torch.int16, device="cuda")
    wt = torch.zeros_like(wb)
    cmm._check(L.cmm_cnn_weights_bf16(1, vp(w32), vp(wb), vp(wt), stream()), "weights")
    assert np.array_equal(host16(wb), w.ravel())
    gin = torch.zeros(N * H * W * 32, dtype=torch.int16, device="cuda")
    # every device input is held in a variable while a call reads it (a temporary's memory goes
    # back to the allocator, which can hand it to the next tensor before the kernel runs)
    gd = dev16(g)
    cmm._check(L.cmm_cnn_backward_data(N, H, W, vp(gd), vp(wt), vp(gin), stream()), "backward data")
    refd = R.conv_back_data(g, w)
    got = host16(gin).reshape(refd.shape)
    e = np.abs(got - refd).max() / np.abs(refd).max()
    print("backward data: rel err", e)
    assert e <= 2.0 ** -8 + 1e-5
    # with the ReLU derivative of the layer below in the same pass: the separate mask's result
    act = dev16(R.bf16(rng.standard_normal((N, H, W, 32))))
    gin_r = torch.zeros_like(gin)
    cmm._check(L.cmm_cnn_backward_data_relu(N, H, W, vp(gd), vp(wt), vp(act), vp(gin_r), stream()),
               "backward data relu")
    cmm._check(L.cmm_cnn_relu_mask(N * H * W * 32, vp(act), vp(gin), stream()), "relu mask")
    assert np.array_equal(host16(gin_r), host16(gin))
    dw = torch.zeros(32 * 9 * 32, dtype=torch.float32, device="cuda")
    db = torch.zeros(32, dtype=torch.float32, device="cuda")
    cmm._check(L.cmm_cnn_backward_weights(N, H, W, vp(gd), vp(xin), vp(dw), vp(db), stream()), "backward w")
    rw, rb = R.conv_back_weights(g, x)
    ew = np.abs(dw.cpu().numpy().reshape(rw.shape) - rw).max() / np.abs(rw).max()
    eb = np.abs(db.cpu().numpy() - rb).max() / np.abs(rb).max()
    print("backward weights: rel err", ew, "bias", eb)
    assert ew <= 1e-5 and eb <= 1e-5


def test_conv_kernels_on_large_images_match_cpu(dev):
    """Images large enough for the strip-resident launch (2046 x 160 x 32 channels, where blocks
    stay over a 128-pixel strip and stage each next row group during the products, including a
    strip of one segment and a height not a multiple of the 4-row groups): forward with FP32 and
    ReLU outputs and the backward-data pass with the ReLU derivative, against the CPU reference
    with the small test's bounds."""
    import cmm
    import torch
    rng = np.random.default_rng(2)
    N, H, W = 1, 2046, 160
    x = R.bf16(rng.standard_normal((N, H, W, 32)))
    w = R.bf16(rng.standard_normal((32, 9, 32)) * 0.1)
    b = rng.standard_normal(32).astype(np.float32).astype(np.float64)
    L = cmm.lib()
    xin, wd, bd = dev16(x), dev16(w), dev32(b)
    ref = R.conv(x, w, b)
    out32 = torch.zeros(N * H * W * 32, dtype=torch.float32, device="cuda")
    cmm._check(L.cmm_cnn_forward(N, H, W, vp(xin), vp(wd), vp(bd), 2, vp(out32), stream()), "forward")
    got = out32.cpu().numpy().reshape(ref.shape)
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print("large forward fp32 out: rel err", e)
    assert e <= 1e-5
    out16 = torch.zeros(N * H * W * 32, dtype=torch.int16, device="cuda")
    cmm._check(L.cmm_cnn_forward(N, H, W, vp(xin), vp(wd), vp(bd), 0, vp(out16), stream()), "forward relu")
    got = host16(out16).reshape(ref.shape)
    refr = np.maximum(ref, 0)
    assert np.all(np.abs(got - refr) <= 2.0 ** -8 * np.abs(refr) + 1e-5 * np.abs(ref).max())
    g = R.bf16(rng.standard_normal((N, H, W, 32)))
    act = R.bf16(rng.standard_normal((N, H, W, 32)))
    w32 = dev32(w.ravel())
    wb = torch.zeros(32 * 9 * 32, dtype=torch.int16, device="cuda")
    wt = torch.zeros_like(wb)
    cmm._check(L.cmm_cnn_weights_bf16(1, vp(w32), vp(wb), vp(wt), stream()), "weights")
    gin = torch.zeros(N * H * W * 32, dtype=torch.int16, device="cuda")
    gd, ad = dev16(g), dev16(act)
    cmm._check(L.cmm_cnn_backward_data_relu(N, H, W, vp(gd), vp(wt), vp(ad), vp(gin), stream()),
               "backward data relu")
    refd = R.conv_back_data(g, w) * (act > 0)
    got = host16(gin).reshape(refd.shape)
    e = np.abs(got - refd).max() / np.abs(refd).max()
    print("large backward data relu: rel err", e)
    assert e <= 2.0 ** -8 + 1e-5


def test_network_gradient_matches_cpu(dev):
    import learned
    import torch
    rng = np.random.default_rng(3)
    net = learned.Net(seed=4)
    N, H, W = 2, 4, 128
    net.buffers(N, H, W)
    x = np.zeros((N, H, W, 32))
    x[..., :3] = R.bf16(rng.standard_normal((N, H, W, 3)))
    tgt = rng.standard_normal((N, H, W, 2)).astype(np.float32).astype(np.float64)
    net.act[0].copy_(dev16(x))
    net.tgt.copy_(torch.from_numpy(tgt.astype(np.float32).ravel()))
    net.forward()
    net.backward()
    torch.cuda.synchronize()
    p = net.p.cpu().numpy()
    ws = [R.bf16(p[l * learned.WPL:(l + 1) * learned.WPL].reshape(32, 9, 32)) for l in range(4)]
    bs = [p[4 * learned.WPL + 32 * l:4 * learned.WPL + 32 * (l + 1)].astype(np.float64) for l in range(4)]
    loss, dws, dbs, out = R.network(x, ws, bs, tgt, bf=True)
    gl = 0.5 * float(net.loss.item())
    g = net.g.cpu().numpy()
    ref = np.concatenate([d.ravel() for d in dws] + [d for d in dbs])
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    cos = float(g @ ref / (np.linalg.norm(g) * np.linalg.norm(ref)))
    print(f"loss gpu {gl:.6e} cpu {loss:.6e}; gradient rel err {rel:.3e}, cosine {cos:.6f}")
    for l in range(4):
        gl_ = g[l * learned.WPL:(l + 1) * learned.WPL]
        print(f"  layer {l}: rel err {np.linalg.norm(gl_ - dws[l].ravel()) / np.linalg.norm(dws[l]):.3e}")
    assert abs(gl - loss) <= 1e-3 * loss
    assert rel <= 2e-2 and cos >= 0.999


def test_training_gain_and_fallback(dev):
    """trains on seeded trajectories (map 128^2, psi 256^2, fine 512^2, psi recorded every 1/32, so
    the extrapolations span 1/16, where the Lagrange extrapolation's residual is large enough to
    learn), then compares the extrapolation errors with and without the network on held-out seeds:
    the corrected extrapolation to t + dt must be at least 5 % closer to the recorded psi (measured
    1.13x with these 6 training seeds and 300 iterations, 1.13x with 64 seeds and 3000).  Then the
    corrected time loop at dt = 1/16 with the fallback judge (1e-6): its extrapolations must be
    closer to the new Poisson psi than the plain ones while it is active."""
    import cmm
    import learned
    train_seeds, test_seeds = range(0, 6), range(64, 66)
    trajs = [learned.trajectory(s, dt=1.0 / 32, steps=64) for s in train_seeds]
    held = [learned.trajectory(s, dt=1.0 / 32, steps=64) for s in test_seeds]
    net = learned.Net(seed=0)
    rms = learned.fit_scales(net, trajs)
    print("scales (psi, d1, d2, r1, r12):", rms)
    hist = learned.train(net, trajs, iters=300, batch=4, lr=2e-3, log=print)
    assert all(np.isfinite(l) for _, l in hist)
    p1, l1, p12, l12 = learned.evaluate(net, held)
    print(f"held-out relative L2 extrapolation error to t+dt: plain {p1:.3e}, learned {l1:.3e} "
          f"(gain {p1 / l1:.3f}); to t+dt/2: plain {p12:.3e}, learned {l12:.3e} (gain {p12 / l12:.3f})")
    assert p1 / l1 >= 1.05 and p12 / l12 > 1.0
    # the corrected loop at twice the recording step, with the judge
    g = cmm.grid
    e = cmm.Euler(g(128), g(512), g(256), g(512), 1.0 / 16, initial_condition="random_spectrum", seed=64,
                  freq_cut_psi=128.0, stack_capacity=16)
    e.init()
    cor = learned.Corrector(net, e, g(256), tol=1e-6)
    for _ in range(16):
        cor.step(1.0 / 16)
    print(f"in the loop: learned deviations {np.round(np.log10(cor.dev_learned[:8]), 2)} ..., plain "
          f"{np.round(np.log10(cor.dev_plain[:8]), 2)} ..., fallback at step {cor.fallback_step}")
    assert len(cor.dev_plain) == 16
    assert cor.active or cor.fallback_step is not None
    n_active = len(cor.dev_learned)
    assert n_active >= 1 and np.mean(cor.dev_learned) < np.mean(cor.dev_plain[:n_active])
    e.close()
