"""GPU tests of the split-precision form: the map stored as identity + FP32 displacement, FP32
records of the Lagrange-combined psi fields and of the fine vorticity, FP64 arithmetic, FP64 psi
levels, FFTs and Poisson divide.  The form has no counterpart in the reference; its outputs are
compared with the FP64 oracle (operators) and with the FP64 engine, which itself matches the
oracle to 1e-12 (test_gpu_parity.py, test_gpu_engine.py), and each error is printed.

Bounds written in the tests:
- record conversion: the value planes (stored with an FP32 low part) within 4 U^2 of the
  displacement, the derivative planes within one FP32 rounding (U = 2^-24 relative);
- one advection step: map values within 8 dt U max|psi| / h_psi + 4 U h_c max|D'| (the rounding
  of the combined psi records' value plane, which the Hermite derivative amplifies by about 6/h,
  and of the map's derivative planes, weighted by h), derivative planes within 1e-6 (in the
  stencil's 1/eps scale, as the FP64 parity tests measure them);
- the check: |E_split - E_fp64| <= 8 U (max|D'| + h max|D''|)(1 + max|D'|) (FP32 rounding of the
  derivative planes, which enter the Jacobian with weight about one);
- psi from the split map and vorticity: 1e-6 relative per plane (spectral scale);
- the split stack (FP64 value planes, FP32 derivative planes): the values bit-identical, the
  derivatives one FP32 rounding; composed positions within the sum over layers of
  h/4 U (|dX/dx| + |dX/dy|) + h^2/16 U |d2X/dxdy| (the Hermite weights of the derivative planes sum
  to at most 1/4 per direction, 1/16 for the cross plane), each carried through the later layers
  by their gradient norm;
- the split engine on the double-shear-layer workload (map 128^2, psi 256^2, fine 512^2,
  dt 1/64) to T = 2: relative L2 vorticity error <= 1e-5 against the FP64 engine.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev_array, hermite_errors, host, rel_linf

pytestmark = pytest.mark.gpu

GC, GF, GP = 256, 512, 256
U = 2.0 ** -24


def f32(n):
    import torch
    return torch.zeros(n, dtype=torch.float32, device="cuda:0")


def kmax(n):
    return np.pi / (O.TWO_PI / n)


@pytest.fixture(scope="module")
def base():
    gc, gf, gp = O.grid(GC), O.grid(GF), O.grid(GP)
    return O.euler_run(gc, gf, gp, gf, dt=1.0, final_time=1.0, skip_remapping=1,
                       want=("chix", "chiy", "psi", "w_herm"))


def node_coords(n):
    h = O.TWO_PI / float(np.float32(n))
    x = (np.arange(n) * h)[None, :].repeat(n, 0).ravel()
    y = (np.arange(n) * h)[:, None].repeat(n, 1).ravel()
    return x, y


def displacement_planes(cx, cy, n):
    """planes of X - identity (as the split records hold them)"""
    x, y = node_coords(n)
    N = n * n
    dx = cx.copy()
    dy = cy.copy()
    dx[:N] -= x
    dx[N:2 * N] -= 1.0
    dy[:N] -= y
    dy[2 * N:3 * N] -= 1.0
    return dx, dy


def test_split_record_roundtrip(dev, base):
    import cmm
    g = cmm.grid(GC)
    N = GC * GC
    m = f32(12 * N)
    cx, cy = dev_array(base["chix"]), dev_array(base["chiy"])
    cmm.map_to_split(g, cx, cy, m)
    bx, by = dev_array(np.zeros(4 * N)), dev_array(np.zeros(4 * N))
    cmm.map_from_split(g, m, bx, by)
    dx, dy = displacement_planes(base["chix"], base["chiy"], GC)
    for name, got, ref, d in (("X", host(bx), base["chix"], dx), ("Y", host(by), base["chiy"], dy)):
        for p in range(4):
            err = np.max(np.abs(got[p * N:(p + 1) * N] - ref[p * N:(p + 1) * N]))
            # value planes keep a low part (about 2^-48 relative), derivative planes one FP32 rounding
            rel = U * U * 4 if p == 0 else 1.01 * U
            bound = rel * np.max(np.abs(d[p * N:(p + 1) * N])) + 4e-16 * np.max(np.abs(ref[p * N:(p + 1) * N]))
            print(f"{name} plane {p}: roundtrip error {err:.3e} (bound {bound:.3e})")
            assert err <= bound


@pytest.mark.parametrize("case", ["deformed_dt1", "identity_dt1/32", "identity_2nd_order"])
def test_split_advect_vs_oracle(dev, base, case):
    import cmm
    gc, gp = cmm.grid(GC), cmm.grid(GP)
    ogc, ogp = O.grid(GC), O.grid(GP)
    psi = base["psi"]
    n4 = 4 * GP * GP
    N = GC * GC
    dt = 1.0 if case == "deformed_dt1" else 1.0 / 32
    muo = 0 if case == "identity_2nd_order" else 1
    t = np.array([-2 * dt, -dt, 0.0, dt])
    dts = np.full(4, dt)
    oc = O.advect_coeffs(t, dts, 0, 3, 30, muo, 1e-3)
    c = cmm.advect_coeffs(t, dts, 0, 3, 30, muo, 1e-3)
    if case == "deformed_dt1":
        mx, my = base["chix"], base["chiy"]
    else:
        mx, my = O.identity_map(ogc)
    rx, ry = O.map_advect(ogc, ogp, mx, my, psi, oc)
    lv = [dev_array(psi[i * n4:(i + 1) * n4]) for i in range(3)]
    m_in, m_out = f32(12 * N), f32(12 * N)
    cmm.map_to_split(gc, dev_array(mx), dev_array(my), m_in)
    cmm.advect_split(gc, gp, m_in, lv, c, m_out)
    ox, oy = dev_array(np.zeros(4 * N)), dev_array(np.zeros(4 * N))
    cmm.map_from_split(gc, m_out, ox, oy)
    gx, gy = host(ox), host(oy)
    dx, dy = displacement_planes(rx, ry, GC)
    for name, g, r, d in (("X", gx, rx, dx), ("Y", gy, ry, dy)):
        dscale = np.max(np.abs(d[:N]))
        ev = np.max(np.abs(g[:N] - r[:N]))
        errs = hermite_errors(g, r, N, 1.0 / 1e-3)
        # FP32 rounding of the combined psi records' value plane, which reaches the velocity through
        # the basis derivative (about 6 U |psi| / h), and of the map's derivative planes (weight h)
        hc = O.TWO_PI / GC
        bound = 8 * dt * U * np.max(np.abs(psi[:GP * GP])) / (O.TWO_PI / GP) + 4 * U * hc * np.max(np.abs(d[N:]))
        print(f"split advect {case} {name}: value abs error {ev:.3e} (bound {bound:.3e}, max |X - x| {dscale:.3e}), "
              f"derivative planes {errs[1:]}")
        assert ev <= bound
        assert max(errs[1:]) <= 1e-6


def test_split_incomp(dev, base):
    import cmm
    gc, gf = cmm.grid(GC), cmm.grid(GF)
    ogc, ogf = O.grid(GC), O.grid(GF)
    N = GC * GC
    # the deformed map and a map one small step from the identity
    psi = base["psi"]
    n4 = 4 * GP * GP
    dt = 1.0 / 32
    t = np.array([-2 * dt, -dt, 0.0, dt])
    oc = O.advect_coeffs(t, np.full(4, dt), 0, 3, 30, 1, 1e-3)
    ix, iy = O.identity_map(ogc)
    sx, sy = O.map_advect(ogc, O.grid(GP), ix, iy, psi, oc)
    for name, mx, my in (("deformed", base["chix"], base["chiy"]), ("small step", sx, sy)):
        ref = O.incomp_error(ogc, ogf, mx, my)
        m = f32(12 * N)
        cmm.map_to_split(gc, dev_array(mx), dev_array(my), m)
        got = cmm.incompressibility_check_split(gf, gc, m)
        dx, dy = displacement_planes(mx, my, GC)
        # FP32 rounding of the derivative planes, which enter the Jacobian with weight about one
        # (the value planes keep a low part)
        hc = O.TWO_PI / GC
        dmax = max(np.max(np.abs(v[N:3 * N])) for v in (dx, dy))
        dxy = max(np.max(np.abs(v[3 * N:])) for v in (dx, dy)) * hc
        bound = 8 * U * (dmax + dxy) * (1 + dmax) + 1e-13
        print(f"split check {name}: E_split {got:.6e} E_fp64 {ref:.6e} diff {abs(got - ref):.3e} (bound {bound:.3e})")
        assert abs(got - ref) <= bound


def test_split_psi(dev, base):
    import cmm
    gc, gf, gp = cmm.grid(GC), cmm.grid(GF), cmm.grid(GP)
    ogc, ogf, ogp = O.grid(GC), O.grid(GF), O.grid(GP)
    N, NF = GC * GC, GF * GF
    ref = O.psi_from_vort(ogc, ogf, ogf, ogp, base["chix"], base["chiy"], base["w_herm"], 0, 128.0)
    m, w = f32(12 * N), f32(4 * NF)
    cmm.map_to_split(gc, dev_array(base["chix"]), dev_array(base["chiy"]), m)
    cmm.hermite_to_split(gf, dev_array(base["w_herm"]), w)
    out = dev_array(np.zeros(4 * GP * GP))
    cmm.evaluate_stream_hermite_split(gc, gf, gf, gp, m, w, out, 0, 128.0)
    errs = hermite_errors(host(out), ref, GP * GP, kmax(GP))
    print("split psi per-plane errors", errs)
    assert max(errs) <= 1e-6


def torus_err(a, b):
    d = np.abs(a - b)
    return np.minimum(d, O.TWO_PI - d)


def run_c1(precision, steps, n_part=1024):
    import cmm
    e = cmm.Euler(cmm.grid(128), cmm.grid(512), cmm.grid(256), cmm.grid(512), 1.0 / 64, freq_cut_psi=128.0,
                  initial_condition="double_shear_layer", incomp_threshold=1e-4, stack_capacity=64,
                  precision=precision)
    e.add_particles(cmm.uniform_particles(n_part, 0), 30)
    e.init()
    infos = [e.step(1.0 / 64) for _ in range(steps)]
    out = {"einc": np.array([i.incomp for i in infos]), "stack": infos[-1].stack_size,
           "w": e.sample(cmm.grid(512))["w"], "psi": e.field("psi")[:4 * 256 * 256],
           "w_coarse": e.field("w_coarse"), "part": e.particles(0)}
    e.close()
    return out


def test_split_engine_double_shear_layer(dev):
    """The double-shear-layer workload (map 128^2, psi 256^2, fine 512^2, dt 1/64, RK3, 4th-order
    stencil, remap at 1e-4) to T = 2 in both forms: per-output errors of the split form."""
    ref = run_c1("fp64", 128)
    got = run_c1("split", 128)
    rl2 = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    n = 256 * 256
    ew = (rel_linf(got["w"], ref["w"]), rl2(got["w"], ref["w"]))
    ep = (rel_linf(got["psi"][:n], ref["psi"][:n]), rl2(got["psi"][:n], ref["psi"][:n]))
    ex = torus_err(got["part"], ref["part"])
    print(f"remaps fp64 {ref['stack']} split {got['stack']}")
    print(f"omega (512^2 through the stack): rel Linf {ew[0]:.3e} rel L2 {ew[1]:.3e}")
    print(f"psi: rel Linf {ep[0]:.3e} rel L2 {ep[1]:.3e}")
    print(f"particles: max {ex.max():.3e} rms {np.sqrt(np.mean(ex ** 2)):.3e}")
    print(f"E_inc: max abs diff {np.abs(got['einc'] - ref['einc']).max():.3e}")
    assert got["stack"] == ref["stack"]
    assert ew[1] <= 1e-5
    assert ep[1] <= 1e-5
    assert ex.max() <= 1e-5


def test_split_stack_composition(dev):
    """positions through a stack of sheared maps: the plane-major FP64 stack against its FP64 node
    records (bit for bit) and its split records (within the FP32 rounding bound)"""
    import cmm
    import torch
    nc, nf, nm = 128, 512, 12
    gm, gt = cmm.grid(nc), cmm.grid(nf)
    N = nc * nc
    h = O.TWO_PI / float(np.float32(nc))
    xs = np.arange(nc) * h
    X, Y = np.meshgrid(xs, xs)
    X, Y = X.ravel(), Y.ravel()
    st = np.empty((nm, 8, N))
    bound, gains = [], []
    for m in range(nm):
        a, b, ph = 0.02 * (1 + m % 4), 0.013 * (1 + m % 3), 0.41 * m
        st[m, 0] = X + a * np.sin(Y + ph) + b * np.sin(X + Y)
        st[m, 1] = 1.0 + b * np.cos(X + Y)
        st[m, 2] = a * np.cos(Y + ph) + b * np.cos(X + Y)
        st[m, 3] = -b * np.sin(X + Y)
        st[m, 4] = Y + a * np.sin(X - ph)
        st[m, 5] = a * np.cos(X - ph)
        st[m, 6] = 1.0
        st[m, 7] = 0.0
        ex = h / 4 * U * (np.abs(st[m, 1]).max() + np.abs(st[m, 2]).max()) + h * h / 16 * U * np.abs(st[m, 3]).max()
        ey = h / 4 * U * (np.abs(st[m, 5]).max() + np.abs(st[m, 6]).max()) + h * h / 16 * U * np.abs(st[m, 7]).max()
        bound.append(max(ex, ey))
        gains.append(max((np.abs(st[m, 1]) + np.abs(st[m, 2])).max(), (np.abs(st[m, 5]) + np.abs(st[m, 6])).max()))
    # the layers run newest first: the error made in layer m is carried by layers m-1 .. 0
    tot = sum(bound[m] * float(np.prod(gains[:m])) for m in range(nm))
    stack = torch.from_numpy(st.ravel()).cuda()
    # both record formats: their bytes against the planes, then the compositions
    packed = torch.zeros(nm * 6 * N, dtype=torch.float64, device="cuda")
    cmm.stack_pack(gm, stack, nm, "split", packed)
    rec64 = torch.zeros(nm * 8 * N, dtype=torch.float64, device="cuda")
    cmm.stack_pack(gm, stack, nm, "fp64", rec64)
    pb = packed.cpu().numpy().view(np.uint8).reshape(nm, N, 48)
    r64 = rec64.cpu().numpy().reshape(nm, N, 8)
    for m in range(nm):
        v = pb[m, :, :16].copy().view(np.float64)
        d = pb[m, :, 16:40].copy().view(np.float32)
        assert np.array_equal(v[:, 0], st[m, 0]) and np.array_equal(v[:, 1], st[m, 4])
        assert np.array_equal(d, st[m, [1, 2, 3, 5, 6, 7]].T.astype(np.float32))
        assert np.array_equal(r64[m], st[m].T)
    cx = torch.empty(4 * N, dtype=torch.float64, device="cuda")
    cy = torch.empty_like(cx)
    cmm.map_init_identity(gm, cx, cy)
    xy64 = torch.empty(2 * nf * nf, dtype=torch.float64, device="cuda")
    xys = torch.empty_like(xy64)
    xyr = torch.empty_like(xy64)
    cmm.apply_map_stack(gt, gm, cx, cy, stack, nm, xy64)
    cmm.apply_map_stack_packed(gt, gm, cx, cy, packed, nm, "split", xys)
    cmm.apply_map_stack_packed(gt, gm, cx, cy, rec64, nm, "fp64", xyr)
    # FP64 records: the same values in the same arithmetic, so the same positions bit for bit
    assert torch.equal(xyr, xy64)
    d = np.abs(xys.cpu().numpy() - xy64.cpu().numpy())
    d = np.minimum(d, O.TWO_PI - d)
    print(f"split stack composition, {nm} maps: max position difference {d.max():.3e}, bound {tot:.3e}")
    assert d.max() <= tot
    cmm.apply_map_stack_packed(gt, gm, cx, cy, packed, 0, "split", xys)
    cmm.apply_map_stack(gt, gm, cx, cy, stack, 0, xy64)
    assert torch.equal(xys, xy64)
