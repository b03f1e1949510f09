"""GPU parity: every HIP kernel of the map-advection path against the CPU oracle on identical
seeded inputs (C-ABI calls through cmm.py), at the fixture sizes C256 / F512 / psi256.

Tolerance: relative L-infinity (max |gpu - oracle| / max |oracle|) <= 1e-12 for the values of
omega, psi, the map and positions.  Derivative planes of Hermite fields are held to the same 1e-12
measured in the derivative's own scale (gpu_util.hermite_errors): max|f| * (pi/h)^m for
spectral Hermite fields, where FFT rounding in mode k grows by k^m, and max|X| / eps^m for a map
built by the footpoint stencil, where footpoint rounding grows by 1/eps^m.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev_array, hermite_errors, host, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12


def kmax(n):
    return np.pi / (O.TWO_PI / n)

GC, GF, GP = 256, 512, 256


def grids():
    import cmm
    return cmm.grid(GC), cmm.grid(GF), cmm.grid(GP)


def ogrids():
    return O.grid(GC), O.grid(GF), O.grid(GP)


def random_field(n, seed):
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n) * n
    F = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.exp(-(k[None, :] ** 2 + k[:, None] ** 2) / 200.0)
    return np.real(np.fft.ifft2(F)) * n


@pytest.fixture(scope="module")
def base():
    """Oracle state after one strong step from the 4_nodes condition: map, psi history, fine omega."""
    gc, gf, gp = ogrids()
    r = O.euler_run(gc, gf, gp, gf, dt=1.0, final_time=1.0, skip_remapping=1,
                    want=("chix", "chiy", "psi", "w_herm"))
    return r


def test_hermite_from_real(dev):
    import cmm
    for n, seed in ((256, 1), (512, 2), (64, 3)):
        f = random_field(n, seed)
        ref = O.hermite_from_real(O.grid(n), f.ravel())
        out = dev_array(np.zeros(4 * n * n))
        cmm.hermite_from_real(cmm.grid(n), dev_array(f.ravel()), out)
        errs = hermite_errors(host(out), ref, n * n, kmax(n))
        print(f"n={n} per-plane errors {errs}")
        assert max(errs) <= TOL


def test_translate_ic_empty_stack(dev):
    import cmm
    gc, gf, gp = grids()
    ogc, ogf, ogp = ogrids()
    cx, cy = O.identity_map(ogc)
    xy = O.compose_positions(ogc, ogf, cx, cy, None, None, 0)
    samp = O.sample_init(ogf, xy, 0)
    ref = O.hermite_from_real(ogf, samp)
    W = dev_array(np.zeros(4 * GF * GF))
    cmm.translate_initial_condition_through_map_stack(gc, gf, dev_array(cx), dev_array(cy), None, 0, 0, W)
    errs = hermite_errors(host(W), ref, GF * GF, kmax(GF))
    print("translate_ic per-plane errors", errs)
    assert max(errs) <= TOL


def test_compose_and_sample_with_stack(dev, base):
    import cmm
    gc, gf, gp = grids()
    ogc, ogf, ogp = ogrids()
    # current map and a stack of two maps, all equal to the strongly deformed map
    mx, my = base["chix"], base["chiy"]
    stack = np.concatenate([mx, my, mx, my])
    # the oracle takes the stack as separate X and Y arrays of maps
    sx = np.concatenate([mx, mx])
    sy = np.concatenate([my, my])
    ref_xy = O.compose_positions(ogc, ogf, mx, my, sx, sy, 2)
    xy = dev_array(np.zeros(2 * GF * GF))
    cmm.apply_map_stack(gf, gc, dev_array(mx), dev_array(my), dev_array(stack), 2, xy)
    # positions are wrapped into [0, 2 pi) after each layer: one within rounding of the boundary
    # can come out on either side, so they are compared as points of the torus
    d = np.mod(host(xy) - ref_xy + np.pi, 2 * np.pi) - np.pi
    e = np.max(np.abs(d)) / np.max(np.abs(ref_xy))
    print("compose rel Linf", e)
    assert e <= TOL
    ref_s = O.sample_init(ogf, ref_xy, 0)
    out = dev_array(np.zeros(GF * GF))
    cmm.sample_init(gc, gf, dev_array(mx), dev_array(my), dev_array(stack), 2, 0, out)
    e = rel_linf(host(out), ref_s)
    print("sample_init rel Linf", e)
    assert e <= TOL


def _shear_maps(n, a, b):
    """Hermite planes of A(x, y) = (x + a sin y, y) and B(x, y) = (x, y + b sin x) on an n^2 grid."""
    h = 2 * np.pi / float(np.float32(n))
    X, Y = np.meshgrid(np.arange(n) * h, np.arange(n) * h)
    one, zero = np.ones_like(X), np.zeros_like(X)
    ax = np.concatenate([p.ravel() for p in (X + a * np.sin(Y), one, a * np.cos(Y), zero)])
    ay = np.concatenate([p.ravel() for p in (Y, zero, one, zero)])
    bx = np.concatenate([p.ravel() for p in (X, one, zero, zero)])
    by = np.concatenate([p.ravel() for p in (Y + b * np.sin(X), b * np.cos(X), one, zero)])
    return ax, ay, bx, by


def test_forward_composition_and_points(dev, base):
    """Forward positions through a forward stack (oldest first, current last) and forwarded points
    with every map selection, against the oracle."""
    import cmm
    gc, gf, gp = grids()
    ogc, ogf, ogp = ogrids()
    ax, ay, bx, by = _shear_maps(GC, 0.3, 0.3)
    mx, my = base["chix"], base["chiy"]
    stack = np.concatenate([mx, my, ax, ay])
    sx, sy = np.concatenate([mx, ax]), np.concatenate([my, ay])
    ref_xy = O.forward_positions(ogc, ogf, bx, by, sx, sy, 2)
    xy = dev_array(np.zeros(2 * GF * GF))
    cmm.apply_map_stack_forward(gf, gc, dev_array(bx), dev_array(by), dev_array(stack), 2, xy)
    d = np.mod(host(xy) - ref_xy + np.pi, 2 * np.pi) - np.pi
    e = np.max(np.abs(d)) / np.max(np.abs(ref_xy))
    print("forward compose rel Linf", e)
    assert e <= TOL
    rng = np.random.default_rng(11)
    pts = rng.uniform(0.0, 2 * np.pi, size=(5000, 2)).ravel()
    pts[:4] = [0.0, 1.0, 1.0, 0.0]  # on the boundary: NaN
    zoom = cmm.Grid(64, 64, 1.0, 5.0, 0.5, 4.5)
    for gb, bounds in ((gc, (0.0, 2 * np.pi, 0.0, 2 * np.pi)), (zoom, (1.0, 5.0, 0.5, 4.5))):
        for first, cur in ((0, 1), (1, 1), (2, 1), (0, 0), (2, 0), (5, 1)):
            ref = O.forward_points(ogc, bounds, bx, by, sx, sy, 2, first, cur, pts)
            out = dev_array(np.zeros_like(pts))
            cmm.apply_map_stack_points_forward(gb, gc, dev_array(bx), dev_array(by), dev_array(stack), 2, first, cur,
                                               dev_array(pts), out)
            got = host(out)
            assert np.array_equal(np.isnan(got), np.isnan(ref)), (first, cur)
            ok = ~np.isnan(ref)
            dd = np.mod(got[ok] - ref[ok] + np.pi, 2 * np.pi) - np.pi
            print("forward points", bounds, first, cur, "nan", int((~ok).sum()), "max diff", np.max(np.abs(dd)))
            assert np.max(np.abs(dd)) <= TOL * 2 * np.pi


def test_psi_from_vort(dev, base):
    import cmm
    gc, gf, gp = grids()
    ogc, ogf, ogp = ogrids()
    for fcut in (128.0, 40.5, 1000.0):
        ref = O.psi_from_vort(ogc, ogf, ogf, ogp, base["chix"], base["chiy"], base["w_herm"], 0, fcut)
        out = dev_array(np.zeros(4 * GP * GP))
        cmm.evaluate_stream_hermite(gc, gf, gf, gp, dev_array(base["chix"]), dev_array(base["chiy"]),
                                    dev_array(base["w_herm"]), out, 0, fcut)
        errs = hermite_errors(host(out), ref, GP * GP, kmax(GP))
        print(f"psi fcut={fcut} per-plane errors {errs}")
        assert max(errs) <= TOL


def test_map_advect_rk3(dev, base):
    import cmm
    gc, gf, gp = grids()
    ogc, ogf, ogp = ogrids()
    psi = base["psi"]
    n4 = 4 * GP * GP
    for dt, start in ((1.0, "deformed"), (1.0 / 32, "identity")):
        t = np.array([-2 * dt, -dt, 0.0, dt])
        dts = np.full(4, dt)
        oc = O.advect_coeffs(t, dts, 0, 3, 30, 1, 1e-3)
        c = cmm.advect_coeffs(t, dts, 0, 3, 30, 1, 1e-3)
        if start == "identity":
            mx, my = O.identity_map(ogc)
        else:
            mx, my = base["chix"], base["chiy"]
        rx, ry = O.map_advect(ogc, ogp, mx, my, psi, oc)
        lv = [dev_array(psi[i * n4:(i + 1) * n4]) for i in range(3)]
        ox = dev_array(np.zeros(4 * GC * GC))
        oy = dev_array(np.zeros(4 * GC * GC))
        cmm.advect_using_stream_hermite(gc, gp, dev_array(mx), dev_array(my), lv, c, ox, oy)
        gx, gy = host(ox), host(oy)
        n = GC * GC
        for name, g, r in (("X", gx, rx), ("Y", gy, ry)):
            errs = hermite_errors(g, r, n, 1.0 / 1e-3)
            print(f"advect dt={dt} {start} {name} per-plane errors {errs}")
            assert max(errs) <= TOL


@pytest.mark.parametrize("ratio", [1, 2])
def test_map_advect_schemes(dev, base, ratio):
    """Every time scheme (Euler, AB2, RK2, RK3, RK3-mod, RK4, RK4-mod) and footpoint stencil order
    (2, 4, 6) on the deformed map, with psi on the map grid and on a 2x finer grid, at a small
    (1/32) and a large (1/2) step; each without a workspace (levels interpolated one by one) and
    with one (Lagrange-weighted psi combinations formed first)."""
    import torch
    import cmm
    gc, gf, _ = grids()
    ogc, ogf, _ = ogrids()
    npsi = GC * ratio
    ogp, gp = O.grid(npsi), cmm.grid(npsi)
    if ratio == 1:
        psi = base["psi"]
    else:
        p1 = O.psi_from_vort(ogc, ogf, ogf, ogp, base["chix"], base["chiy"], base["w_herm"], 0, 128.0)
        psi = np.concatenate([p1, 0.9 * p1, 0.8 * p1])
    n4 = 4 * npsi * npsi
    lv = [dev_array(psi[i * n4:(i + 1) * n4]) for i in range(3)]
    mx, my = dev_array(base["chix"]), dev_array(base["chiy"])
    n = GC * GC
    for dt in (1.0 / 32, 0.5):
        t = np.array([-2 * dt, -dt, 0.0, dt])
        dts = np.full(4, dt)
        for tin in (10, 20, 21, 30, 31, 40, 41):
            for muo in (0, 1, 2):
                oc = O.advect_coeffs(t, dts, 0, 3, tin, muo, 1e-3)
                c = cmm.advect_coeffs(t, dts, 0, 3, tin, muo, 1e-3)
                rx, ry = O.map_advect(ogc, ogp, base["chix"], base["chiy"], psi, oc)
                wb = cmm.advect_workspace_bytes(gc, gp, c)
                # C1, C12 (RK3/RK4 variants) and the copy of level 0 (all but Euler), 4 planes each,
                # and the interleaved map (8 values per map node)
                nf = 3 if tin in (30, 40, 41) else 2 if tin == 21 else 1
                assert wb == (0 if tin in (20, 31) else nf * 32 * npsi * npsi + 64 * GC * GC)
                for work in (None, torch.zeros(max(wb, 8) // 8, dtype=torch.float64, device="cuda")):
                    ox = dev_array(np.zeros(4 * n))
                    oy = dev_array(np.zeros(4 * n))
                    cmm.advect_using_stream_hermite(gc, gp, mx, my, lv, c, ox, oy, work=work)
                    e = max(max(hermite_errors(host(ox), rx, n, 1.0 / 1e-3)),
                            max(hermite_errors(host(oy), ry, n, 1.0 / 1e-3)))
                    print(f"advect r={ratio} dt={dt} tin={tin} order={muo} work={work is not None}: {e:.2e}")
                    assert e <= TOL, (ratio, dt, tin, muo, work is not None, e)


@pytest.mark.parametrize("ratio", [1, 2])
def test_map_advect_forward(dev, base, ratio):
    """The forward map (direction +1): every scheme and stencil order advects the deformed map
    forward with the reference's forward schemes, against the oracle; then the invertibility check
    of the backward and forward maps of the same step against the oracle's."""
    import cmm
    gc, gf, _ = grids()
    ogc, ogf, _ = ogrids()
    npsi = GC * ratio
    ogp, gp = O.grid(npsi), cmm.grid(npsi)
    if ratio == 1:
        psi = base["psi"]
    else:
        p1 = O.psi_from_vort(ogc, ogf, ogf, ogp, base["chix"], base["chiy"], base["w_herm"], 0, 128.0)
        psi = np.concatenate([p1, 0.9 * p1, 0.8 * p1])
    n4 = 4 * npsi * npsi
    lv = [dev_array(psi[i * n4:(i + 1) * n4]) for i in range(3)]
    mx, my = dev_array(base["chix"]), dev_array(base["chiy"])
    n = GC * GC
    dt = 1.0 / 32
    t = np.array([-2 * dt, -dt, 0.0, dt])
    dts = np.full(4, dt)
    for tin in (10, 20, 21, 30, 31, 40, 41):
        for muo in (0, 1, 2):
            oc = O.advect_coeffs(t, dts, 0, 3, tin, muo, 1e-3, direction=1)
            c = cmm.advect_coeffs(t, dts, 0, 3, tin, muo, 1e-3, direction=1)
            rx, ry = O.map_advect(ogc, ogp, base["chix"], base["chiy"], psi, oc)
            ox = dev_array(np.zeros(4 * n))
            oy = dev_array(np.zeros(4 * n))
            cmm.advect_using_stream_hermite(gc, gp, mx, my, lv, c, ox, oy)
            e = max(max(hermite_errors(host(ox), rx, n, 1.0 / 1e-3)), max(hermite_errors(host(oy), ry, n, 1.0 / 1e-3)))
            print(f"forward r={ratio} tin={tin} order={muo}: {e:.2e}")
            assert e <= TOL, (ratio, tin, muo, e)
    # invertibility of this step's backward and forward maps from the identity
    ix, iy = O.identity_map(ogc)
    cb = O.advect_coeffs(t, dts, 0, 3, 30, 1, 1e-3)
    cf = O.advect_coeffs(t, dts, 0, 3, 30, 1, 1e-3, direction=1)
    bx, by = O.map_advect(ogc, ogp, ix, iy, psi, cb)
    fx, fy = O.map_advect(ogc, ogp, ix, iy, psi, cf)
    ref = O.lib().cmmo_invert_error(ogc, ogc, ogc, O.ptr(bx), O.ptr(by), O.ptr(fx), O.ptr(fy))
    got = cmm.invertibility_check(gc, gc, dev_array(bx), dev_array(by), gc, dev_array(fx), dev_array(fy))
    print("invertibility", got, ref)
    # a distance |x - X_b(X_f(x))| ~ 1e-6 between positions of size ~1: rounding is absolute, ~1e-15
    assert ref > 0 and abs(got - ref) <= 1e-10 * ref + 1e-13


def test_incompressibility(dev, base):
    import cmm
    gc, gf, gp = grids()
    ogc, ogf, ogp = ogrids()
    ref = O.incomp_error(ogc, ogf, base["chix"], base["chiy"])
    got = cmm.incompressibility_check(gf, gc, dev_array(base["chix"]), dev_array(base["chiy"]))
    print("incomp", got, ref)
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
    # check grids that refine the map grid by 1, 3, 4 and 8 (tiles of 1, 9, 16 points inside one
    # map cell; four tiles per cell for 8) and by 1.5 (one lane per check point)
    for nchk in (256, 768, 1024, 2048, 384):
        ref = O.incomp_error(ogc, O.grid(nchk), base["chix"], base["chiy"])
        got = cmm.incompressibility_check(cmm.grid(nchk), gc, dev_array(base["chix"]), dev_array(base["chiy"]))
        print("incomp", nchk, got, ref)
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
    cx, cy = O.identity_map(ogc)
    # identity map: det = 1 up to the rounding of the basis partition of unity (fused multiply-adds)
    assert cmm.incompressibility_check(gf, gc, dev_array(cx), dev_array(cy)) <= 1e-12


def test_particles_rk3(dev, base):
    import cmm
    gc, gf, gp = grids()
    ogc, ogf, ogp = ogrids()
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, O.TWO_PI, 2 * 1000)
    psi = base["psi"]
    n4 = 4 * GP * GP
    Lp = [0.375, 0.75, -0.125]
    ref = O.particles_advect(ogp, psi, 3, Lp, 0.05, 30, pos)
    d = dev_array(pos)
    cmm.particles_advect(gp, [dev_array(psi[i * n4:(i + 1) * n4]) for i in range(3)], 3, Lp, 0.05, 30, d)
    e = rel_linf(host(d), ref)
    print("particles rel Linf", e)
    assert e <= TOL


@pytest.fixture(scope="module")
def vp_base():
    """Oracle Vlasov-Poisson state after 4 steps of nonlinear Landau damping (map 64^2, fine 128^2)."""
    g64, g128 = O.vp_grid(64), O.vp_grid(128)
    return O.euler_run(g64, g128, g128, g128, dt=0.1, final_time=0.4, ic=202, skip_remapping=1, sim=1,
                       want=("chix", "chiy", "psi", "w_herm"))


def test_vp_psi_from_f(dev, vp_base):
    """Vlasov-Poisson psi on the GPU against the oracle: vort grid equal to the psi grid and half of
    it (zero padding), mollifiers 0/4/8, through the deformed map."""
    import cmm
    gm, gf = cmm.vp_grid(64), cmm.vp_grid(128)
    ogm, ogf = O.vp_grid(64), O.vp_grid(128)
    mx, my, W = vp_base["chix"], vp_base["chiy"], vp_base["w_herm"]
    for nv in (128, 64):
        for molly in (0, 4, 8):
            ref = O.vp_psi_from_f(ogm, O.vp_grid(nv), ogf, ogf, mx, my, W, molly)
            out = dev_array(np.zeros(4 * 128 * 128))
            cmm.evaluate_potential_from_density_hermite(gm, cmm.vp_grid(nv), gf, gf, dev_array(mx), dev_array(my),
                                                        dev_array(W), out, molly)
            errs = hermite_errors(host(out), ref, 128 * 128, 128 / 4.0)
            print(f"vp psi vort {nv} molly {molly} per-plane errors {errs}")
            assert max(errs) <= TOL
    # a vort grid finer than the psi grid is refused, as in the reference (error 1013)
    with pytest.raises(cmm.CmmError):
        cmm.evaluate_potential_from_density_hermite(gm, gf, gf, cmm.vp_grid(64), dev_array(mx), dev_array(my),
                                                    dev_array(W), dev_array(np.zeros(4 * 64 * 64)), 0)


def test_spectral_operators(dev):
    """Hermite planes of F, i_laplacian(F) and laplacian(F) of a random field on the GPU against
    the oracle (derivative planes in their own scale)."""
    import cmm
    for n in (64, 256):
        f = random_field(n, 11)
        for op in (0, 1, 2):
            ref = O.spectral_from_real(O.grid(n), f.ravel(), op)
            out = dev_array(np.zeros(4 * n * n))
            cmm.spectral_from_real(cmm.grid(n), dev_array(f.ravel()), op, out)
            errs = hermite_errors(host(out), ref, n * n, kmax(n))
            print(f"spectral n={n} op={op} per-plane errors {errs}")
            assert max(errs) <= TOL
