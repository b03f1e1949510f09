"""Vlasov-Poisson sample outputs (cmm_euler_vp_sample; the reference's
sample_compute_and_write_vlasov, Cuda-CMM/src/simulation/cmm-simulation-host.cu:898-1030) on a
256^2 sample of the nonlinear Landau case (map 64^2, fine = psi = vort 128^2, dt 1/16, remap at
E_inc > 1e-4, which here remaps every step):

- at t = 0 against closed forms: mass 4 pi, kinetic energy 2 pi, momentum 0, E(x) = sin(x/2)
  (amplitude alpha/k = 0.5/0.5) up to sign, potential energy pi;
- after 16 steps (16 stacked maps) against the oracle: the positions through the engine's own
  current map and stack (oracle composition), f0 there (oracle sampling), and the measures and
  the field recomputed with numpy from the oracle's f (the column integral, the 1D Poisson solve
  phi_hat = I_hat / (N k^2) and E = i k phi_hat by numpy's real FFT, the Nyquist bin dropped).
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import rel_linf

pytestmark = pytest.mark.gpu


def numpy_measures(gs, f):
    n = gs.nx
    hx = (gs.x1 - gs.x0) / float(np.float32(n))
    hv = (gs.y1 - gs.y0) / float(np.float32(n))
    F = f.reshape(n, n)
    v = gs.y0 + np.arange(n) * hv
    mass = hx * hv * F.sum()
    ekin = 0.5 * hx * hv * (F * (v * v)[:, None]).sum()
    mom = hx * hv * (F * v[:, None]).sum()
    integral = hv * F.sum(axis=0) - 1.0
    k = 2 * np.pi / (hx * n) * np.arange(n // 2 + 1)
    ih = np.fft.rfft(integral) / n
    ph = np.zeros_like(ih)
    ph[1:] = ih[1:] / (k[1:] ** 2)
    E = np.fft.irfft(1j * k * ph, n=n) * n
    epot = 0.5 * hx * (E * E).sum()
    return dict(mass=mass, ekin=ekin, momentum=mom, epot=epot, e=E)


def test_vp_sample_closed_forms_and_oracle(dev):
    import cmm
    gc, gf = cmm.vp_grid(64), cmm.vp_grid(128)
    dt = 1.0 / 16
    e = cmm.Euler(gc, gf, gf, gf, dt, initial_condition="nonlin_landau_damping", incomp_threshold=1e-4,
                  stack_capacity=32, simulation_type="cmm_vlasov_poisson_1d")
    e.init()
    gs = cmm.vp_grid(256)
    s0 = e.vp_sample(gs)
    x = gs.x0 + np.arange(256) * (gs.x1 - gs.x0) / float(np.float32(256))
    print("t=0:", {k: s0[k] for k in ("mass", "ekin", "epot", "momentum", "etot")})
    assert abs(s0["mass"] - 4 * np.pi) < 1e-9 and abs(s0["ekin"] - 2 * np.pi) < 1e-9
    assert abs(s0["epot"] - np.pi) < 1e-9 and abs(s0["momentum"]) < 1e-9
    assert abs(s0["etot"] - (s0["ekin"] + s0["epot"])) < 1e-12
    assert np.abs(np.abs(s0["e"]) - np.abs(np.sin(x / 2))).max() < 1e-9
    for _ in range(16):
        info = e.step(dt)
    n_stack = info.stack_size
    assert n_stack >= 2
    s1 = e.vp_sample(gs)
    diag = e.vp_diagnostics()
    cx, cy = e.field("chix"), e.field("chiy")
    maps = [e.stack_map(m) for m in range(n_stack)]
    e.close()
    ogs = O.vp_grid(256)
    sx = np.concatenate([m[0] for m in maps])
    sy = np.concatenate([m[1] for m in maps])
    xy = O.compose_positions(O.vp_grid(64), ogs, cx, cy, sx, sy, n_stack)
    # positions are wrapped into the domain: a point within rounding of an edge may sit at either end
    def wrapped(a, b, L):
        d = np.abs(a - b)
        return np.minimum(d, np.abs(L - d)).max()
    dx = wrapped(s1["x"], xy[0::2], ogs.x1 - ogs.x0)
    dv = wrapped(s1["v"], xy[1::2], ogs.y1 - ogs.y0)
    f = O.sample_init(ogs, xy, 202)
    ef = rel_linf(s1["f"], f)
    ref = numpy_measures(ogs, f)
    print("after", n_stack, "maps: |dx|", dx, "|dv|", dv, "f rel", ef)
    assert dx <= 1e-11 and dv <= 1e-11 and ef <= 1e-12
    for k in ("mass", "ekin", "epot"):
        r = abs(s1[k] - ref[k]) / abs(ref[k])
        print(k, s1[k], ref[k], r)
        assert r <= 1e-11
    assert abs(s1["momentum"] - ref["momentum"]) <= 1e-11 * abs(ref["ekin"])
    assert np.abs(s1["e"] - ref["e"]).max() <= 1e-11 * np.abs(ref["e"]).max()
    assert abs(s1["etot"] - (s1["ekin"] + s1["epot"])) < 1e-12
    assert np.array_equal(s1["emodes"], np.asarray(diag[5:9]))
