"""Learned psi extrapolation: a 4-layer 3x3 convolutional network (32 channels, BF16 weights and
activations, FP32 accumulation on the matrix cores) that predicts the residual of the Lagrange
extrapolation of psi used at the RK sub-stages, to t_{n+1} and to t_n + dt/2, from the last three
psi snapshots.  The network runs and trains through the HIP kernels of libcmm.so (cmm_cnn_*);
torch only allocates device buffers.  This form has no counterpart in the reference.

The correction enters the map advection through cmm_euler_set_psi_correction: its value planes are
made Hermite fields by the spectral Hermite transform of the psi grid, and added to the combined
fields C1 and C12.  After each step the newest psi (the Poisson solve of the new map, the judge) is
compared with the corrected extrapolation to t_{n+1}; when their relative L2 deviation exceeds the
tolerance, the run falls back to plain extrapolation for the rest of the run.
"""
import ctypes as C
import math

import numpy as np

import cmm

CH = 32            # channels of every layer
LAYERS = 4
WPL = CH * 9 * CH  # weights per layer
# Lagrange weights for equal steps: three levels to t_{n+1} and to t_n + dt/2
L1_EQ = (3.0, -3.0, 1.0)
L12_EQ = (1.875, -1.25, 0.375)


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _st():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ck(rc, what):
    cmm._check(rc, what)


class Net:
    """Parameters (FP32 masters, bf16 copies), AdamW state and activation buffers for n_img
    images of H x W."""

    def __init__(self, seed=0, device="cuda"):
        import torch
        self.dev = device
        rng = np.random.default_rng(seed)
        w = np.zeros((LAYERS, CH, 9, CH), dtype=np.float32)
        cin_real = (3, CH, CH, CH)
        cout_real = (CH, CH, CH, 2)
        for l in range(LAYERS):
            std = math.sqrt(2.0 / (9 * cin_real[l])) * (0.1 if l == LAYERS - 1 else 1.0)
            w[l, :cout_real[l], :, :cin_real[l]] = rng.standard_normal((cout_real[l], 9, cin_real[l])) * std
        # parameters: weights of all layers, then biases
        p = np.concatenate([w.ravel(), np.zeros(LAYERS * CH, dtype=np.float32)])
        self.p = torch.from_numpy(p).to(device)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.wb = torch.zeros(LAYERS * WPL, dtype=torch.int16, device=device)
        self.wt = torch.zeros_like(self.wb)
        self.step_count = 0
        self.scales = torch.ones(3, dtype=torch.float32, device=device)
        self.rscales = torch.ones(2, dtype=torch.float32, device=device)
        self.n_img = self.H = self.W = 0
        self.refresh()

    def state(self):
        """the trained state as host arrays (parameters and normalisation scales)"""
        return {"p": self.p.cpu().numpy(), "scales": self.scales.cpu().numpy(), "rscales": self.rscales.cpu().numpy()}

    def load(self, st):
        import torch
        self.p.copy_(torch.from_numpy(np.asarray(st["p"], dtype=np.float32)))
        self.scales.copy_(torch.from_numpy(np.asarray(st["scales"], dtype=np.float32)))
        self.rscales.copy_(torch.from_numpy(np.asarray(st["rscales"], dtype=np.float32)))
        self.refresh()

    def refresh(self):
        _ck(cmm.lib().cmm_cnn_weights_bf16(LAYERS, _vp(self.p), _vp(self.wb), _vp(self.wt), _st()),
            "cmm_cnn_weights_bf16")

    def w(self, l):
        return C.c_void_p(self.wb.data_ptr() + 2 * l * WPL)

    def wt_(self, l):
        return C.c_void_p(self.wt.data_ptr() + 2 * l * WPL)

    def bias(self, l):
        return C.c_void_p(self.p.data_ptr() + 4 * (LAYERS * WPL + l * CH))

    def dw(self, l):
        return C.c_void_p(self.g.data_ptr() + 4 * l * WPL)

    def db(self, l):
        return C.c_void_p(self.g.data_ptr() + 4 * (LAYERS * WPL + l * CH))

    def buffers(self, n_img, H, W):
        import torch
        if (n_img, H, W) == (self.n_img, self.H, self.W):
            return
        npx = n_img * H * W
        self.act = [torch.zeros(npx * CH, dtype=torch.int16, device=self.dev) for _ in range(LAYERS)]
        self.out = torch.zeros(npx * CH, dtype=torch.float32, device=self.dev)
        self.ga = torch.zeros(npx * CH, dtype=torch.int16, device=self.dev)
        self.gb = torch.zeros_like(self.ga)
        self.tgt = torch.zeros(npx * 2, dtype=torch.float32, device=self.dev)
        self.loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self.n_img, self.H, self.W = n_img, H, W

    def forward(self):
        """act[0] holds the input; out the two predicted channels (of 32) per pixel"""
        L = cmm.lib()
        n, H, W = self.n_img, self.H, self.W
        for l in range(LAYERS):
            dst = self.out if l == LAYERS - 1 else self.act[l + 1]
            mode = 2 if l == LAYERS - 1 else 0
            _ck(L.cmm_cnn_forward(n, H, W, _vp(self.act[l]), self.w(l), self.bias(l), mode, _vp(dst), _st()),
                "cmm_cnn_forward")

    def backward(self):
        """gradients of 0.5 sum (out - tgt)^2 into self.g (overwritten); returns nothing, the loss
        (sum of squares) is in self.loss"""
        L = cmm.lib()
        n, H, W = self.n_img, self.H, self.W
        npx = n * H * W
        _ck(L.cmm_cnn_fill(self.g.numel(), _vp(self.g), 0.0, _st()), "cmm_cnn_fill")
        self.loss.zero_()
        _ck(L.cmm_cnn_mse(npx, _vp(self.out), _vp(self.tgt), _vp(self.ga), _vp(self.loss), _st()), "cmm_cnn_mse")
        g, gn = self.ga, self.gb
        for l in range(LAYERS - 1, -1, -1):
            _ck(L.cmm_cnn_backward_weights(n, H, W, _vp(g), _vp(self.act[l]), self.dw(l), self.db(l), _st()),
                "cmm_cnn_backward_weights")
            if l == 0:
                break
            _ck(L.cmm_cnn_backward_data_relu(n, H, W, _vp(g), self.wt_(l), _vp(self.act[l]), _vp(gn), _st()),
                "cmm_cnn_backward_data_relu")
            g, gn = gn, g

    def adamw(self, lr, gscale, wd=1e-4):
        self.step_count += 1
        _ck(cmm.lib().cmm_cnn_adamw(self.p.numel(), _vp(self.p), _vp(self.g), _vp(self.m), _vp(self.v), lr, 0.9,
                                    0.999, 1e-8, wd, gscale, self.step_count, _st()), "cmm_cnn_adamw")
        self.refresh()

    def set_input(self, i, psi0, psi1, psi2):
        """image i of the batch from three psi value planes (device float64 tensors or pointers)"""
        npx = self.H * self.W
        dst = C.c_void_p(self.act[0].data_ptr() + 2 * i * npx * CH)
        _ck(cmm.lib().cmm_cnn_input(npx, _ptr(psi0), _ptr(psi1), _ptr(psi2), _vp(self.scales), dst, _st()),
            "cmm_cnn_input")

    def set_target(self, i, psi0, psi1, psi2, psi_next, psi_half, L1=L1_EQ, L12=L12_EQ):
        npx = self.H * self.W
        l1 = (C.c_double * 3)(*L1)
        l12 = (C.c_double * 3)(*L12)
        dst = C.c_void_p(self.tgt.data_ptr() + 4 * i * npx * 2)
        _ck(cmm.lib().cmm_cnn_target(npx, _ptr(psi0), _ptr(psi1), _ptr(psi2), _ptr(psi_next), _ptr(psi_half),
                                     C.cast(l1, C.c_void_p), C.cast(l12, C.c_void_p), _vp(self.rscales), dst, _st()),
            "cmm_cnn_target")


def _ptr(x):
    if isinstance(x, C.c_void_p):
        return x
    if isinstance(x, int):
        return C.c_void_p(x)
    return C.c_void_p(x.data_ptr())


# ---- trajectories -----------------------------------------------------------------------------

def trajectory(seed, n_map=128, n_psi=256, n_fine=512, dt=1.0 / 256, steps=256, fcut=None):
    """psi value planes of a random-spectrum Euler run (seeded), one per step of length dt:
    a device tensor [steps + 1, n_psi^2] (row 0 after the warm-up).  fcut: the psi cut-off
    (default n_psi / 2)."""
    import torch
    g = cmm.grid
    e = cmm.Euler(g(n_map), g(n_fine), g(n_psi), g(n_fine), dt, initial_condition="random_spectrum", seed=seed,
                  freq_cut_psi=fcut if fcut else n_psi / 2, stack_capacity=16)
    e.init()
    out = torch.zeros((steps + 1, n_psi * n_psi), dtype=torch.float64, device="cuda")
    _ck(cmm.lib().cmm_euler_copy_psi_level(e.h, 0, 0, _vp(out[0])), "cmm_euler_copy_psi_level")
    for k in range(steps):
        e.step(dt)
        _ck(cmm.lib().cmm_euler_copy_psi_level(e.h, 0, 0, _vp(out[k + 1])), "cmm_euler_copy_psi_level")
    torch.cuda.synchronize()
    e.close()
    return out


def samples(traj, stride=2):
    """(k, k - 2, k - 4, k + 2, k + 1) row indices: three levels a full step (2 rows) apart, the
    next level and the half step between"""
    n = traj.shape[0]
    return [(k, k - stride, k - 2 * stride, k + stride, k + stride // 2) for k in range(2 * stride, n - stride)]


def fit_scales(net, trajs):
    """normalisation of the inputs (psi, first and second differences) and of the two residuals,
    from the training trajectories (root mean squares over a subsample)"""
    import torch
    acc = np.zeros(5)
    cnt = 0
    for tr in trajs:
        for (a, b, c, nx, hf) in samples(tr)[::16]:
            pa, pb, pc, pn, ph = (tr[i].cpu().numpy() for i in (a, b, c, nx, hf))
            r1 = pn - (L1_EQ[0] * pa + L1_EQ[1] * pb + L1_EQ[2] * pc)
            r12 = ph - (L12_EQ[0] * pa + L12_EQ[1] * pb + L12_EQ[2] * pc)
            for j, f in enumerate((pa, pa - pb, pa - 2 * pb + pc, r1, r12)):
                acc[j] += np.mean(f * f)
            cnt += 1
    rms = np.sqrt(acc / cnt)
    net.scales.copy_(torch.tensor(rms[:3], dtype=torch.float32))
    net.rscales.copy_(torch.tensor(rms[3:], dtype=torch.float32))
    return rms


def train(net, trajs, iters=1500, batch=8, lr=1e-3, seed=0, log=None, warmup=0.05, clip=2.0):
    """AdamW on the squared error of the scaled residuals over random samples of the trajectories.
    The learning rate rises linearly over the first `warmup` fraction of the iterations and then
    follows a cosine decay; a gradient whose norm exceeds `clip` times the running mean of the
    norms is scaled down to that bound (clip 0: no clipping).  Both keep a rare sample with a large
    residual from throwing the weights off late in training."""
    import torch
    H = W = int(round(math.sqrt(trajs[0].shape[1])))
    net.buffers(batch, H, W)
    rng = np.random.default_rng(seed)
    pool = [(t, s) for t, tr in enumerate(trajs) for s in samples(tr)]
    hist = []
    n_warm = max(1, int(warmup * iters))
    gmean = None
    for it in range(iters):
        pick = rng.choice(len(pool), size=batch, replace=False)
        for i, j in enumerate(pick):
            t, (a, b, c, nx, hf) = pool[j]
            tr = trajs[t]
            net.set_input(i, tr[a], tr[b], tr[c])
            net.set_target(i, tr[a], tr[b], tr[c], tr[nx], tr[hf])
        net.forward()
        net.backward()
        gscale = 1.0 / (batch * H * W)
        if clip > 0:
            gn = float(net.g.norm().item()) * gscale
            if gmean is not None and gn > clip * gmean:
                gscale *= clip * gmean / gn
                gn = clip * gmean
            gmean = gn if gmean is None else 0.98 * gmean + 0.02 * gn
        rate = lr * (it + 1) / n_warm if it < n_warm else \
            lr * 0.5 * (1 + math.cos(math.pi * (it - n_warm) / max(1, iters - n_warm)))
        net.adamw(rate, gscale)
        if it % 100 == 0 or it == iters - 1:
            l = float(net.loss.item()) / (batch * H * W * 2)
            hist.append((it, l))
            if log:
                log(f"train it {it}: mean squared scaled residual {l:.4e}")
    torch.cuda.synchronize()
    return hist


def evaluate(net, trajs):
    """relative L2 errors of the extrapolations to t_{n+1} and t_n + dt/2, plain and corrected,
    over every sample of the trajectories: (plain1, learned1, plain12, learned12)"""
    import torch
    H = W = int(round(math.sqrt(trajs[0].shape[1])))
    net.buffers(1, H, W)
    npx = H * W
    d1 = torch.zeros(npx, dtype=torch.float64, device="cuda")
    d12 = torch.zeros_like(d1)
    pred = torch.zeros_like(d1)
    nrm = torch.zeros(2, dtype=torch.float64, device="cuda")
    L = cmm.lib()
    acc = np.zeros(4)
    den = np.zeros(2)
    w1 = (C.c_double * 3)(*L1_EQ)
    w12 = (C.c_double * 3)(*L12_EQ)
    for tr in trajs:
        for (a, b, c, nx, hf) in samples(tr):
            net.set_input(0, tr[a], tr[b], tr[c])
            net.forward()
            _ck(L.cmm_cnn_correction(npx, _vp(net.out), _vp(net.rscales), _vp(d1), _vp(d12), _st()),
                "cmm_cnn_correction")
            for k, (w, d, ref) in enumerate(((w1, d1, tr[nx]), (w12, d12, tr[hf]))):
                for m, corr in enumerate((None, d)):
                    _ck(L.cmm_extrapolate(npx, _vp(tr[a]), _vp(tr[b]), _vp(tr[c]), C.cast(w, C.c_void_p),
                                          None if corr is None else _vp(corr), _vp(pred), _st()), "cmm_extrapolate")
                    _ck(L.cmm_dev_norms(npx, _vp(pred), _vp(ref), _vp(nrm), _st()), "cmm_dev_norms")
                    v = nrm.cpu().numpy()
                    acc[2 * k + m] += v[0]
                    if m == 0:
                        den[k] += v[1]
    return (math.sqrt(acc[0] / den[0]), math.sqrt(acc[1] / den[0]), math.sqrt(acc[2] / den[1]),
            math.sqrt(acc[3] / den[1]))


class Corrector:
    """Inference in the time loop: before each step the network corrects the engine's Lagrange
    extrapolations from its three newest psi levels; after it the corrected extrapolation to
    t_{n+1} is judged against the new Poisson psi, and beyond `tol` (relative L2) the run falls
    back to plain extrapolation."""

    def __init__(self, net, engine, g_psi, tol=1e-6):
        import torch
        self.net, self.e, self.g_psi, self.tol = net, engine, g_psi, tol
        n = g_psi.nx * g_psi.ny
        self.n = n
        net.buffers(1, g_psi.ny, g_psi.nx)
        self.d1 = torch.zeros(n, dtype=torch.float64, device="cuda")
        self.d12 = torch.zeros_like(self.d1)
        self.h1 = torch.zeros(4 * n, dtype=torch.float64, device="cuda")
        self.h12 = torch.zeros_like(self.h1)
        self.pred = torch.zeros(n, dtype=torch.float64, device="cuda")
        self.plain = torch.zeros_like(self.pred)
        self.nrm = torch.zeros(2, dtype=torch.float64, device="cuda")
        self.active = True
        self.fallback_step = None
        self.dev_learned, self.dev_plain = [], []

    def level(self, l):
        p = C.c_void_p()
        _ck(cmm.lib().cmm_euler_psi_level(self.e.h, l, C.byref(p)), "cmm_euler_psi_level")
        return p

    def step(self, dt):
        import torch
        L = cmm.lib()
        n = self.n
        # the engine's psi solve runs on its own stream: the whole device is synchronised before
        # its levels are read here and before the engine reads the corrections
        torch.cuda.synchronize()
        p0, p1, p2 = self.level(0), self.level(1), self.level(2)
        w1 = (C.c_double * 3)(*L1_EQ)
        _ck(L.cmm_extrapolate(n, p0, p1, p2, C.cast(w1, C.c_void_p), None, _vp(self.plain), _st()), "cmm_extrapolate")
        if self.active:
            self.net.set_input(0, p0, p1, p2)
            self.net.forward()
            _ck(L.cmm_cnn_correction(n, _vp(self.net.out), _vp(self.net.rscales), _vp(self.d1), _vp(self.d12),
                                     _st()), "cmm_cnn_correction")
            cmm.hermite_from_real(self.g_psi, self.d1, self.h1)
            cmm.hermite_from_real(self.g_psi, self.d12, self.h12)
            _ck(L.cmm_extrapolate(n, p0, p1, p2, C.cast(w1, C.c_void_p), _vp(self.d1), _vp(self.pred), _st()),
                "cmm_extrapolate")
            _ck(L.cmm_euler_set_psi_correction(self.e.h, _vp(self.h1), _vp(self.h12)), "cmm_euler_set_psi_correction")
        else:
            _ck(L.cmm_euler_set_psi_correction(self.e.h, None, None), "cmm_euler_set_psi_correction")
        torch.cuda.synchronize()
        info = self.e.step(dt)
        torch.cuda.synchronize()
        new = self.level(0)
        _ck(L.cmm_dev_norms(n, _vp(self.plain), new, _vp(self.nrm), _st()), "cmm_dev_norms")
        v = self.nrm.cpu().numpy()
        self.dev_plain.append(math.sqrt(v[0] / v[1]))
        if self.active:
            _ck(L.cmm_dev_norms(n, _vp(self.pred), new, _vp(self.nrm), _st()), "cmm_dev_norms")
            v = self.nrm.cpu().numpy()
            dl = math.sqrt(v[0] / v[1])
            self.dev_learned.append(dl)
            if dl > self.tol:
                self.active = False
                self.fallback_step = len(self.dev_plain)
                _ck(L.cmm_euler_set_psi_correction(self.e.h, None, None), "cmm_euler_set_psi_correction")
        return info
