"""Pins of the CPU oracle by what the mathematics fixes (closed forms, invariants, brute force),
chosen so that a dropped term, a wrong sign or index, or a transposed operand fails.

Covers: Hermite evaluation / velocity / map evaluation / determinant, Lagrange and stencil
coefficients, FFT and C2R semantics (numpy.fft), fourier_hermite and the psi solve, map advection
with zero and linear psi (backward and forward), the forward scheme's order and invertibility,
composition, initial conditions new to this engine, particles.
"""
import numpy as np
import pytest

import oracle_lib as O

TWO_PI = O.TWO_PI


def herm_from_poly(n, f, fx, fy, fxy, L=TWO_PI):
    """Hermite planes of a smooth function sampled on the grid (exact derivative planes)."""
    h = L / float(np.float32(n))
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    return np.concatenate([g(X, Y).ravel() for g in (f, fx, fy, fxy)]), h


def test_hermite_reproduces_bicubic_locally():
    # a bicubic polynomial in the cell-local coordinates is reproduced exactly inside one cell
    n = 16
    rng = np.random.default_rng(0)
    c = rng.standard_normal((4, 4))
    g = O.grid(n, b=(0.0, 1.0, 0.0, 1.0))
    h = 1.0 / float(np.float32(n))

    def p(x, y, dx=0, dy=0):
        s = 0.0
        for i in range(4):
            for j in range(4):
                if i < dx or j < dy:
                    continue
                fi = np.prod(range(i - dx + 1, i + 1)) if dx else 1
                fj = np.prod(range(j - dy + 1, j + 1)) if dy else 1
                s = s + c[i, j] * fi * fj * x ** (i - dx) * y ** (j - dy)
        return s

    H, _ = herm_from_poly(n, lambda X, Y: p(X, Y), lambda X, Y: p(X, Y, 1, 0), lambda X, Y: p(X, Y, 0, 1),
                          lambda X, Y: p(X, Y, 1, 1), L=1.0)
    for (x, y) in ((0.3 * h + 5 * h, 0.71 * h + 7 * h), (2 * h + 0.5 * h, 3 * h + 0.99 * h)):
        v = O.lib().cmmo_hermite_eval(g, O.ptr(H), x, y)
        assert abs(v - p(x, y)) < 1e-12 * max(1, abs(p(x, y)))
        u = np.zeros(2)
        O.lib().cmmo_hermite_grad(g, O.ptr(H), 1, O.C.c_double(x), O.C.c_double(y), O.ptr(u))
        assert abs(u[0] - p(x, y, 0, 1)) < 1e-10          # u = (d/dy, -d/dx)
        assert abs(u[1] + p(x, y, 1, 0)) < 1e-10


def test_hermite_tensor_formula_brute_force():
    # direct tensor-product Hermite formula on a 4x4 periodic grid with random data
    n = 4
    rng = np.random.default_rng(1)
    H = rng.standard_normal(4 * n * n)
    g = O.grid(n, b=(0.0, 1.0, 0.0, 1.0))
    h = 1.0 / float(np.float32(n))
    x, y = 0.37, 0.81
    ix, iy = int(np.floor(x / h)), int(np.floor(y / h))
    dx, dy = x / h - ix, y / h - iy
    h00 = lambda t: 2 * t ** 3 - 3 * t ** 2 + 1
    h01 = lambda t: -2 * t ** 3 + 3 * t ** 2
    h10 = lambda t: t ** 3 - 2 * t ** 2 + t
    h11 = lambda t: t ** 3 - t ** 2
    v = 0.0
    for cx, wx, dwx in ((ix, h00(dx), h10(dx)), ((ix + 1) % n, h01(dx), h11(dx))):
        for cy, wy, dwy in ((iy, h00(dy), h10(dy)), ((iy + 1) % n, h01(dy), h11(dy))):
            k = cy * n + cx
            v += wx * wy * H[k] + dwx * h * wy * H[n * n + k] + wx * dwy * h * H[2 * n * n + k] \
                + dwx * dwy * h * h * H[3 * n * n + k]
    got = O.lib().cmmo_hermite_eval(g, O.ptr(H), O.C.c_double(x), O.C.c_double(y))
    assert abs(got - v) < 1e-14


def test_identity_map_and_determinant():
    g = O.grid(64)
    cx, cy = O.identity_map(g)
    xo, yo = O.C.c_double(), O.C.c_double()
    for (x, y) in ((0.1, 0.2), (-3.3, 7.9), (TWO_PI * 2.5, -0.01)):
        O.lib().cmmo_diffeo_eval(g, O.ptr(cx), O.ptr(cy), O.C.c_double(x), O.C.c_double(y), O.C.byref(xo), O.C.byref(yo))
        assert abs(xo.value - x) < 1e-12 and abs(yo.value - y) < 1e-12   # wrap jumps restore x
    assert O.incomp_error(g, O.grid(128), cx, cy) < 1e-13


def test_area_preserving_shear_det():
    # X = x + a sin(y), Y = y has det 1; Hermite interpolation keeps it to interpolation accuracy
    n, a = 128, 0.3
    g = O.grid(n)
    Hx, _ = herm_from_poly(n, lambda X, Y: X + a * np.sin(Y), lambda X, Y: 1 + 0 * X, lambda X, Y: a * np.cos(Y),
                           lambda X, Y: 0 * X)
    Hy, _ = herm_from_poly(n, lambda X, Y: Y, lambda X, Y: 0 * X, lambda X, Y: 1 + 0 * X, lambda X, Y: 0 * X)
    e = O.incomp_error(g, O.grid(256), Hx, Hy)
    assert e < 1e-6
    # a wrong sign in the cross term would break it: X = x + a sin(y) + a sin(x) is not area preserving
    Hx2, _ = herm_from_poly(n, lambda X, Y: X + a * np.sin(Y) + a * np.sin(X), lambda X, Y: 1 + a * np.cos(X),
                            lambda X, Y: a * np.cos(Y), lambda X, Y: 0 * X)
    assert O.incomp_error(g, O.grid(256), Hx2, Hy) > 0.1


def test_lagrange_and_stencil_coefficients():
    dt = 0.1
    t = np.array([-2 * dt, -dt, 0.0, dt])
    c = O.advect_coeffs(t, np.full(4, dt), 0, 3, 30, 1, 1e-3)
    assert np.allclose(list(c.L1)[:3], [3, -3, 1], atol=1e-13)
    assert np.allclose(list(c.L12)[:3], [1.875, -1.25, 0.375], atol=1e-13)
    # 4th-order stencil: weights of the value sum to 1; derivative weights vanish on constants
    assert abs(sum(list(c.c1)[:8]) - 1) < 1e-15
    assert abs(sum(list(c.cx)[:8])) < 1e-9 and abs(sum(list(c.cy)[:8])) < 1e-9 and abs(sum(list(c.cxy)[:8])) < 1e-3
    # and reproduce d/dx of a linear function at the footpoint offsets (x = +-eps, +-2 eps)
    eps = 1e-3
    offs_x = [((k >> 2) + 1) * eps * (1 - 2 * ((k >> 1) & 1)) for k in range(8)]
    offs_y = [((k >> 2) + 1) * eps * (1 - 2 * (((k + 1) >> 1) & 1)) for k in range(8)]
    assert abs(sum(w * o for w, o in zip(list(c.cx)[:8], offs_x)) - 1) < 1e-12
    assert abs(sum(w * o for w, o in zip(list(c.cy)[:8], offs_y)) - 1) < 1e-12
    assert abs(sum(w * ox * oy for w, ox, oy in zip(list(c.cxy)[:8], offs_x, offs_y)) - 1) < 1e-9


@pytest.mark.parametrize("n", [8, 64, 256])
def test_fft_against_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    F = O.rfft2(a)
    assert np.max(np.abs(F - np.fft.rfft2(a))) < 1e-12 * n * n
    # C2R with arbitrary (non-Hermitian) input incl. imaginary DC / Nyquist bins: numpy's semantics
    G = rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))
    assert np.max(np.abs(O.irfft2(G, n) - np.fft.irfft2(G, s=(n, n)) * n * n)) < 1e-11 * n * n


def test_fourier_hermite_of_trig_field():
    # f = sin(2x) cos(3y) + cos(x): derivative planes exact
    n = 64
    H_exact, h = herm_from_poly(n, lambda X, Y: np.sin(2 * X) * np.cos(3 * Y) + np.cos(X),
                                lambda X, Y: 2 * np.cos(2 * X) * np.cos(3 * Y) - np.sin(X),
                                lambda X, Y: -3 * np.sin(2 * X) * np.sin(3 * Y),
                                lambda X, Y: -6 * np.cos(2 * X) * np.sin(3 * Y))
    H = O.hermite_from_real(O.grid(n), H_exact[:n * n])
    assert np.max(np.abs(H - H_exact)) < 1e-12


def test_psi_solve_sign_and_cutoff():
    # omega = cos(2x) + cos(y) sampled through the identity map: psi = -(cos(2x)/4 + cos(y));
    # with freq_cut 1.5 the k=2 mode is removed
    n = 64
    g = O.grid(n)
    cx, cy = O.identity_map(g)
    W, h = herm_from_poly(n, lambda X, Y: np.cos(2 * X) + np.cos(Y), lambda X, Y: -2 * np.sin(2 * X),
                          lambda X, Y: -np.sin(Y), lambda X, Y: 0 * X)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    psi = O.psi_from_vort(g, g, g, g, cx, cy, W, 0, 32.0)
    assert np.max(np.abs(psi[:n * n] - (-(np.cos(2 * X) / 4 + np.cos(Y))).ravel())) < 1e-12
    assert np.max(np.abs(psi[n * n:2 * n * n] - (np.sin(2 * X) / 2).ravel())) < 1e-12
    psi = O.psi_from_vort(g, g, g, g, cx, cy, W, 0, 1.5)
    assert np.max(np.abs(psi[:n * n] - (-np.cos(Y)).ravel())) < 1e-12
    # psi grid smaller than the vorticity grid (spectrum move)
    gp = O.grid(32)
    psi = O.psi_from_vort(g, g, g, gp, cx, cy, W, 0, 32.0)
    xp = np.arange(32) * (TWO_PI / 32.0)
    Xp, Yp = np.meshgrid(xp, xp)
    assert np.max(np.abs(psi[:32 * 32] - (-(np.cos(2 * Xp) / 4 + np.cos(Yp))).ravel())) < 1e-12


def psi_linear(n, a, b):
    """Hermite psi = a*y - b*x is not periodic; use psi = a sin(y) - b sin(x) near the origin
    region instead: here we build u = (psi_y, -psi_x) = (a, b) constant via psi = a*y - b*x with the
    Hermite planes set directly (values are never wrapped in the advection)."""
    H, h = herm_from_poly(n, lambda X, Y: a * Y - b * X, lambda X, Y: -b + 0 * X, lambda X, Y: a + 0 * X,
                          lambda X, Y: 0 * X)
    return H


def test_advection_zero_psi_keeps_identity():
    n = 32
    g = O.grid(n)
    cx, cy = O.identity_map(g)
    psi = np.zeros(3 * 4 * n * n)
    dt = 0.1
    c = O.advect_coeffs(np.array([-2 * dt, -dt, 0, dt]), np.full(4, dt), 0, 3, 30, 1, 1e-3)
    nx, ny = O.map_advect(g, g, cx, cy, psi, c)
    # derivative planes are footpoint differences over eps (eps^2 for the cross plane), so rounding
    # of the footpoint positions (~1e-15) grows by 1/eps per derivative
    nn = n * n
    for new, old in ((nx, cx), (ny, cy)):
        for p, tol in enumerate((1e-13, 1e-10, 1e-10, 1e-8)):
            assert np.max(np.abs(new[p * nn:(p + 1) * nn] - old[p * nn:(p + 1) * nn])) < tol * TWO_PI


def test_advection_uniform_velocity_is_translation():
    # psi planes encode the constant velocity u = (a, b) inside the domain interior: the footpoint
    # moves back by dt*(a, b) and the new map is the translated identity in the interior nodes
    n = 32
    g = O.grid(n)
    a, b = 0.3, -0.2
    P = psi_linear(n, a, b)
    psi = np.concatenate([P, P, P])
    cx, cy = O.identity_map(g)
    dt = 0.05
    for tin in (10, 21, 30, 40):
        L = {10: 1, 21: 2, 30: 3, 40: 4}[tin]
        t = np.array([-(L - 1 - i) * dt for i in range(L)] + [dt])
        c = O.advect_coeffs(t, np.full(L + 1, dt), 0, L, tin, 1, 1e-3)
        psiL = np.concatenate([P] * L)
        nx, ny = O.map_advect(g, g, cx, cy, psiL, c)
        nn = n * n
        inner = np.zeros((n, n), bool)
        inner[4:-4, 4:-4] = True
        inner = inner.ravel()
        assert np.max(np.abs(nx[:nn][inner] - (cx[:nn][inner] - dt * a))) < 1e-12
        assert np.max(np.abs(ny[:nn][inner] - (cy[:nn][inner] - dt * b))) < 1e-12
        assert np.max(np.abs(nx[nn:2 * nn][inner] - 1)) < 1e-9
        assert np.max(np.abs(nx[2 * nn:3 * nn][inner])) < 1e-9 and np.max(np.abs(nx[3 * nn:][inner])) < 1e-6


def test_composition_of_translations():
    n = 32
    gm = O.grid(n)
    gt = O.grid(64)
    cx, cy = O.identity_map(gm)
    shifts = [(0.1, -0.2), (0.05, 0.3), (-0.4, 0.01)]
    sx, sy = [], []
    for dx, dy in shifts:
        x, y = O.identity_map(gm)
        x[:n * n] += dx
        y[:n * n] += dy
        sx.append(x)
        sy.append(y)
    xy = O.compose_positions(gm, gt, cx, cy, np.concatenate(sx), np.concatenate(sy), 3)
    h = TWO_PI / 64.0
    g1 = np.arange(64) * h
    X, Y = np.meshgrid(g1, g1)
    ex = np.mod(X + sum(s[0] for s in shifts), TWO_PI).ravel()
    ey = np.mod(Y + sum(s[1] for s in shifts), TWO_PI).ravel()
    dx = np.abs(xy[0::2] - ex)
    dx = np.minimum(dx, TWO_PI - dx)
    assert dx.max() < 1e-12
    dy = np.abs(xy[1::2] - ey)
    dy = np.minimum(dy, TWO_PI - dy)
    assert dy.max() < 1e-12
    # empty stack with the identity map reproduces the initial condition on the grid
    xy0 = O.compose_positions(gm, gt, cx, cy, None, None, 0)
    w = O.sample_init(gt, xy0, 0)
    assert np.max(np.abs(w - (np.cos(X) + np.cos(Y) + 0.6 * np.cos(2 * X) + 0.2 * np.cos(3 * X)).ravel())) < 1e-13


def test_random_spectrum_initial_condition():
    n = 256
    w = O.ic_random_spectrum(n, 0).reshape(n, n)
    h = TWO_PI / float(np.float32(n))
    assert abs(0.5 * h * h * np.sum(w ** 2) - 1.0) < 1e-12     # enstrophy 1
    assert abs(w.mean()) < 1e-14                                 # zero mean
    F = np.fft.rfft2(w)
    ky = np.abs(np.fft.fftfreq(n, 1 / n))[:, None] + 0 * np.arange(n // 2 + 1)[None, :]
    kx = np.arange(n // 2 + 1)[None, :] + 0 * ky
    assert np.abs(F[(kx > 64) | (ky > 64)]).max() < 1e-12 * np.abs(F).max()   # support |kx|,|ky| <= 64
    # amplitude law |omega_hat|^2 = A^2 k E(k)/pi: ratio of two modes is fixed by E(k)
    E = lambda k: k ** 4 * np.exp(-2 * (k / 6) ** 2)
    a1, a2 = np.abs(F[0, 3]), np.abs(F[4, 0])        # k = 3 and k = 4
    assert abs((a1 / a2) ** 2 - (3 * E(3)) / (4 * E(4))) < 1e-10
    assert not np.allclose(O.ic_random_spectrum(n, 1), w.ravel())   # the seed matters


def test_double_shear_layer_is_periodic_and_odd():
    n = 256
    f = lambda x, y: O.lib().cmmo_init_vorticity(O.C.c_double(x), O.C.c_double(y), 100, None)
    assert abs(f(0.3, 0.0) - f(0.3, TWO_PI)) < 1e-12
    assert abs(f(0.0, 1.1) - f(TWO_PI, 1.1)) < 1e-12
    delta = np.pi / 15
    # at y = pi/2 the upper layer peaks: 1/delta + tail of the lower one + 0.05 cos x
    # at x = pi/2 (cos x = 0) and y = pi/2 the upper layer peaks at 1/delta; the lower layer
    # contributes -sech^2(15)/delta ~ -2e-12
    assert abs(f(np.pi / 2, np.pi / 2) - 1 / delta) < 1e-10
    # odd under y -> 2 pi - y about the layers' midpoint for the shear part: omega(x, pi - y) = -omega
    assert abs(f(0.7, np.pi / 2 + 0.3) + f(0.7 + np.pi, 3 * np.pi / 2 + 0.3)) < 1e-12


def test_particles_uniform_velocity_straight_lines():
    n = 32
    g = O.grid(n)
    a, b = 0.3, -0.2
    P = psi_linear(n, a, b)
    psi = np.concatenate([P, P, P])
    pos = np.array([1.0, 2.0, 3.0, 1.5, 2.5, 2.5])
    for scheme in (10, 20, 30, 40, 31, 41):
        out = O.particles_advect(g, psi, 3, [0.375, 0.75, -0.125], 0.1, scheme, pos)
        assert np.max(np.abs(out[0::2] - (pos[0::2] + 0.1 * a))) < 1e-12
        assert np.max(np.abs(out[1::2] - (pos[1::2] + 0.1 * b))) < 1e-12


def test_particles_rk3_third_order():
    # rotation psi = -(x^2+y^2)/2 around the domain centre is not periodic; use the steady cellular
    # flow psi = sin x sin y and compare one step against a fine RK4 reference: error ~ dt^4
    n = 256
    g = O.grid(n)
    P, h = herm_from_poly(n, lambda X, Y: np.sin(X) * np.sin(Y), lambda X, Y: np.cos(X) * np.sin(Y),
                          lambda X, Y: np.sin(X) * np.cos(Y), lambda X, Y: np.cos(X) * np.cos(Y))
    psi = np.concatenate([P, P, P])
    p0 = np.array([1.0, 0.5])

    def ref(p, T, m=2000):
        p = p.copy()
        dt = T / m
        f = lambda q: np.array([np.sin(q[0]) * np.cos(q[1]), -np.cos(q[0]) * np.sin(q[1])])
        for _ in range(m):
            k1 = f(p); k2 = f(p + dt / 2 * k1); k3 = f(p + dt / 2 * k2); k4 = f(p + dt * k3)
            p = p + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        return p

    errs = []
    for dt in (0.2, 0.1):
        out = O.particles_advect(g, psi, 3, [0.375, 0.75, -0.125], dt, 30, p0)
        errs.append(np.max(np.abs(out - ref(p0, dt))))
    assert errs[1] < errs[0] / 10        # local error of a third-order step scales like dt^4


def vp_bump_H(n, v0, hv, Lv):
    """H(v) of compute_periodification_function computed with numpy's FFT: the centred bump
    divided by its maximum, then H_hat = S_hat / (n ky^2), DC 0."""
    v = v0 + np.arange(n) * hv
    sg = 0.5 - 0.5 * np.tanh(2 * np.pi * (np.abs(v) - 0.4 * Lv) / (0.1 * Lv))
    sg -= sg.mean()
    sg /= sg.max()
    S = np.fft.rfft(sg) / n
    ky = 2 * np.pi / (hv * n) * np.arange(n // 2 + 1)
    S[1:] /= ky[1:] ** 2
    S[0] = 0
    return np.fft.irfft(S, n) * n, sg


def test_vp_psi_closed_form():
    """f(x, v) = A(x)/Lv on the vort grid (64^2), identity map, psi grid 128^2:
    I(x) = hv (f(x,v0) + sum_v f) - 1 = (1 + 1/Nv) A(x) - 1, so with
    A = 1 + eps cos(k1 x) + c cos(pi ix) (the vort grid's Nyquist mode)
    phi = (1 + 1/Nv) [eps cos(k1 x)/k1^2 + 2 c cos(kN x)/kN^2] on the psi grid: +1/k^2 (the
    reference divides by +k^2), the double-counted first row (1 + 1/Nv), and the Nyquist bin copied
    into an interior bin by the zero padding (factor 2).  psi = phi(x) - H(v) with H'' = -sigma_c."""
    nv, npsi = 64, 128
    gv, gp = O.vp_grid(nv), O.vp_grid(npsi)
    Lx, Lv = gv.x1 - gv.x0, gv.y1 - gv.y0
    hx = Lx / float(np.float32(nv))
    hv = Lv / float(np.float32(nv))
    eps, c = 0.1, 0.03
    ix = np.arange(nv)
    A = 1 + eps * np.cos(2 * np.pi / Lx * ix * hx) + c * np.cos(np.pi * ix)
    f = np.tile(A / Lv, (nv, 1))
    F = O.hermite_from_real(gv, f.ravel())
    cx, cy = O.identity_map(gv)
    psi = O.vp_psi_from_f(gv, gv, gv, gp, cx, cy, F)
    n = npsi * npsi
    hxp = Lx / float(np.float32(npsi))
    hvp = Lv / float(np.float32(npsi))
    xp = np.arange(npsi) * hxp
    k1, kN = 2 * np.pi / Lx, 2 * np.pi / Lx * (nv // 2)
    fac = 1 + 1 / nv
    phi = fac * (eps * np.cos(k1 * xp) / k1 ** 2 + 2 * c * np.cos(kN * xp) / kN ** 2)
    dphi = -fac * (eps * np.sin(k1 * xp) / k1 + 2 * c * np.sin(kN * xp) / kN)
    H, sg = vp_bump_H(npsi, gp.y0, hvp, Lv)
    expect = phi[None, :] - H[:, None]
    assert np.max(np.abs(psi[:n] - expect.ravel())) < 1e-12
    assert np.max(np.abs(psi[n:2 * n] - np.tile(dphi, npsi))) < 1e-11
    # d2H/dv2 = -sigma_c: spectral second derivative of -H gives back the centred bump
    Hk = np.fft.rfft(H)
    kv = 2 * np.pi / (hvp * npsi) * np.arange(npsi // 2 + 1)
    d2 = np.fft.irfft(-(kv ** 2) * Hk, npsi)
    assert np.max(np.abs(-d2 - sg)) < 1e-9
    # the same density on equal grids: no padding, the Nyquist mode keeps its single weight
    psi_eq = O.vp_psi_from_f(gv, gv, gv, gv, cx, cy, F)
    x = ix * hx
    phi_eq = fac * (eps * np.cos(k1 * x) / k1 ** 2 + c * np.cos(kN * x) / kN ** 2)
    H_eq, _ = vp_bump_H(nv, gv.y0, hv, Lv)
    assert np.max(np.abs(psi_eq[:nv * nv] - (phi_eq[None, :] - H_eq[:, None]).ravel())) < 1e-12


def test_vp_psi_landau_density():
    """Landau initial condition through the identity map: rho - 1 = 0.05 cos(x/2) (the Maxwellian
    over [-2.5 pi, 2.5 pi) integrates to 1 within 1e-13), so phi = 0.05/0.25 cos(x/2) and
    dpsi/dx = -0.1 sin(x/2)."""
    n = 128
    g = O.vp_grid(n)
    cx, cy = O.identity_map(g)
    xy = O.compose_positions(g, g, cx, cy, None, None, 0)
    f = O.sample_init(g, xy, 200)
    F = O.hermite_from_real(g, f)
    psi = O.vp_psi_from_f(g, g, g, g, cx, cy, F)
    h = (g.x1 - g.x0) / float(np.float32(n))
    x = np.arange(n) * h
    nn = n * n
    dpsi = psi[nn:2 * nn].reshape(n, n)
    assert np.max(np.abs(dpsi - (-0.1 * np.sin(0.5 * x))[None, :])) < 1e-11
    # a vort grid finer than the psi grid is rejected (the reference's error 1013)
    with pytest.raises(AssertionError):
        O.vp_psi_from_f(g, g, g, O.vp_grid(64), cx, cy, F)


def test_spectral_operators_closed_form():
    """laplacian(cos 2x cos 3y) = -13 cos 2x cos 3y; i_laplacian gives -1/13 of it (mean removed:
    a constant added to F disappears); grad planes are the exact derivatives."""
    n = 64
    g = O.grid(n)
    h = TWO_PI / float(np.float32(n))
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    F = np.cos(2 * X) * np.cos(3 * Y) + 0.7
    base = np.cos(2 * X) * np.cos(3 * Y)
    nn = n * n
    lap = O.spectral_from_real(g, F.ravel(), 2)
    assert np.max(np.abs(lap[:nn] - (-13 * base).ravel())) < 1e-11
    il = O.spectral_from_real(g, F.ravel(), 1)
    assert np.max(np.abs(il[:nn] - (-base / 13).ravel())) < 1e-13
    assert np.max(np.abs(il[nn:2 * nn] - (2 * np.sin(2 * X) * np.cos(3 * Y) / 13).ravel())) < 1e-13
    assert np.max(np.abs(il[2 * nn:3 * nn] - (3 * np.cos(2 * X) * np.sin(3 * Y) / 13).ravel())) < 1e-13
    he = O.spectral_from_real(g, F.ravel(), 0)
    assert np.max(np.abs(he - O.hermite_from_real(g, F.ravel()))) == 0.0


def test_forward_advection_uniform_velocity():
    """Forward map (direction +1) under a constant velocity u = (a, b): the old map is applied at the
    footpoint first, then the point moves forward, so one step from the identity gives x + dt*u in
    the interior for every scheme but RK4Mod, whose forward formula is the backward one in the
    reference (Cuda-CMM/src/numerical/cmm-timestep.cu:384-430) and gives x - dt*u."""
    n = 32
    g = O.grid(n)
    a, b = 0.3, -0.2
    P = psi_linear(n, a, b)
    cx, cy = O.identity_map(g)
    dt = 0.05
    nn = n * n
    inner = np.zeros((n, n), bool)
    inner[4:-4, 4:-4] = True
    inner = inner.ravel()
    for tin in (10, 20, 21, 30, 31, 40, 41):
        L = {10: 1, 20: 2, 21: 2, 30: 3, 31: 3, 40: 4, 41: 4}[tin]
        t = np.array([-(L - 1 - i) * dt for i in range(L)] + [dt])
        c = O.advect_coeffs(t, np.full(L + 1, dt), 0, L, tin, 1, 1e-3, direction=1)
        nx, ny = O.map_advect(g, g, cx, cy, np.concatenate([P] * L), c)
        s = -1.0 if tin == 41 else 1.0
        assert np.max(np.abs(nx[:nn][inner] - (cx[:nn][inner] + s * dt * a))) < 1e-12, tin
        assert np.max(np.abs(ny[:nn][inner] - (cy[:nn][inner] + s * dt * b))) < 1e-12, tin
        assert np.max(np.abs(nx[nn:2 * nn][inner] - 1)) < 1e-9


def _advect_steps(g, psi_of_t, steps, dt, tin, L, direction, cx, cy):
    """steps map updates with psi levels psi_of_t(t) at the Lagrange nodes t_n, t_n-1, ..."""
    for k in range(steps):
        t = np.array([(k - (L - 1 - i)) * dt for i in range(L)] + [(k + 1) * dt])
        c = O.advect_coeffs(t, np.full(L + 1, dt), 0, L, tin, 1, 1e-3, direction=direction)
        psi = np.concatenate([psi_of_t(t[L - 1 - l]) for l in range(L)])
        cx, cy = O.map_advect(g, g, cx, cy, psi, c)
    return cx, cy


def test_forward_rk3_is_third_order_and_inverts_the_backward_map():
    """Forward RK3 with a time-dependent stream function psi(t) = (1 + t) sin(x) sin(y): its one-step
    error against a fine reference falls by ~2^4 per halving of dt, which holds only when the stage
    velocities are taken at t_n, t_n + dt/2 and t_n+1 (level 0, L12 and L1 in the reference's
    order), and after a few steps the backward map undoes the forward one (cmmo_invert_error is the
    largest |x - X_b(X_f(x))| over cell centres)."""
    n = 64
    g = O.grid(n)

    def psi_of_t(t):
        return herm_from_poly(n, lambda X, Y: (1 + t) * np.sin(X) * np.sin(Y),
                              lambda X, Y: (1 + t) * np.cos(X) * np.sin(Y),
                              lambda X, Y: (1 + t) * np.sin(X) * np.cos(Y),
                              lambda X, Y: (1 + t) * np.cos(X) * np.cos(Y))[0]

    ident = O.identity_map(g)
    nn = n * n
    errs = []
    for dt in (0.2, 0.1):
        one = _advect_steps(g, psi_of_t, 1, dt, 30, 3, 1, *ident)
        fine = _advect_steps(g, psi_of_t, 16, dt / 16, 30, 3, 1, *ident)
        errs.append(np.max(np.abs(one[0][:nn] - fine[0][:nn])))
    print("forward RK3 one-step errors", errs, errs[0] / errs[1])
    assert errs[0] / errs[1] > 10
    fx, fy = _advect_steps(g, psi_of_t, 4, 0.02, 30, 3, 1, *ident)
    bx, by = _advect_steps(g, psi_of_t, 4, 0.02, 30, 3, -1, *ident)
    e = O.lib().cmmo_invert_error(g, g, g, O.ptr(bx), O.ptr(by), O.ptr(fx), O.ptr(fy))
    # brute force at a few cell centres
    h = TWO_PI / float(np.float32(n))
    worst = 0.0
    xo, yo = O.C.c_double(), O.C.c_double()
    for (i, j) in ((0, 0), (5, 17), (40, 63), (63, 1)):
        x, y = i * h + 0.5 * h, j * h + 0.5 * h
        O.lib().cmmo_diffeo_eval(g, O.ptr(fx), O.ptr(fy), O.C.c_double(x), O.C.c_double(y), O.C.byref(xo), O.C.byref(yo))
        O.lib().cmmo_diffeo_eval(g, O.ptr(bx), O.ptr(by), xo, yo, O.C.byref(xo), O.C.byref(yo))
        worst = max(worst, np.hypot(x - xo.value, y - yo.value))
    print("invertibility", e, worst)
    assert worst <= e < 1e-4


def _shears(n, a, b):
    """Hermite maps of two shears that do not commute: A(x, y) = (x + a sin y, y), B(x, y) = (x, y + b sin x)."""
    one, zero = (lambda X, Y: 1 + 0 * X), (lambda X, Y: 0 * X)
    ax, _ = herm_from_poly(n, lambda X, Y: X + a * np.sin(Y), one, lambda X, Y: a * np.cos(Y), zero)
    ay, _ = herm_from_poly(n, lambda X, Y: Y, zero, one, zero)
    bx, _ = herm_from_poly(n, lambda X, Y: X, one, zero, zero)
    by, _ = herm_from_poly(n, lambda X, Y: Y + b * np.sin(X), lambda X, Y: b * np.cos(X), one, zero)
    return (ax, ay), (bx, by)


def _torus_err(got, want):
    d = np.abs(np.asarray(got) - np.asarray(want)) % TWO_PI
    return np.max(np.minimum(d, TWO_PI - d))


def test_forward_composition_order_with_non_commuting_maps():
    """Forward composition applies the oldest stacked map first and the current map last; backward
    composition the other way round.  Two shears that do not commute tell the orders apart by
    O(a b) = 0.09, far above the bicubic interpolation error (~1e-6 at n = 128)."""
    n, a, b = 128, 0.3, 0.3
    gm, gt = O.grid(n), O.grid(48)
    (ax, ay), (bx, by) = _shears(n, a, b)
    h = TWO_PI / 48.0
    X, Y = np.meshgrid(np.arange(48) * h, np.arange(48) * h)
    X, Y = X.ravel(), Y.ravel()
    # forward: stack [A], current B -> B(A(p))
    xy = O.forward_positions(gm, gt, bx, by, ax, ay, 1)
    x1 = X + a * np.sin(Y)
    assert _torus_err(xy[0::2], x1) < 1e-5 and _torus_err(xy[1::2], Y + b * np.sin(x1)) < 1e-5
    # backward with the same inputs: A(B(p))
    xyb = O.compose_positions(gm, gt, bx, by, ax, ay, 1)
    y1 = Y + b * np.sin(X)
    assert _torus_err(xyb[0::2], X + a * np.sin(y1)) < 1e-5 and _torus_err(xyb[1::2], y1) < 1e-5
    assert _torus_err(xy[0::2], xyb[0::2]) > 0.05
    # empty stack: the current map alone
    xy0 = O.forward_positions(gm, gt, ax, ay, None, None, 0)
    assert _torus_err(xy0[0::2], x1) < 1e-5 and _torus_err(xy0[1::2], Y) < 1e-12


def test_forward_points_selection_and_bounds():
    """Forwarded particles: maps first_map.. of the stack then (optionally) the current map, each
    application only for points strictly inside the bounds (others become NaN); with nothing to
    apply the points are copied."""
    n, a, b = 128, 0.3, 0.3
    gm = O.grid(n)
    (ax, ay), (bx, by) = _shears(n, a, b)
    cx, cy = O.identity_map(gm)
    # stack [identity, A], current B
    sx, sy = np.concatenate([cx, ax]), np.concatenate([cy, ay])
    rng = np.random.default_rng(3)
    p = rng.uniform(0.1, TWO_PI - 0.1, size=(200, 2)).ravel()
    px, py = p[0::2], p[1::2]
    full = (0.0, TWO_PI, 0.0, TWO_PI)
    both = O.forward_points(gm, full, bx, by, sx, sy, 2, 0, True, p)
    x1 = px + a * np.sin(py)
    assert _torus_err(both[0::2], x1) < 1e-5 and _torus_err(both[1::2], py + b * np.sin(np.mod(x1, TWO_PI))) < 1e-5
    # the same as the forward positions of those points on a grid when all maps apply
    only_cur = O.forward_points(gm, full, bx, by, sx, sy, 2, 2, True, p)
    assert _torus_err(only_cur[0::2], px) < 1e-12 and _torus_err(only_cur[1::2], py + b * np.sin(px)) < 1e-5
    no_cur = O.forward_points(gm, full, bx, by, sx, sy, 2, 1, False, p)
    assert _torus_err(no_cur[0::2], x1) < 1e-5 and _torus_err(no_cur[1::2], py) < 1e-12
    copy = O.forward_points(gm, full, bx, by, sx, sy, 2, 2, False, p)
    assert np.array_equal(copy, p)
    # bounds: a point on the boundary (x = 0) or outside a window is NaN after the first application
    q = np.array([0.0, 1.0, 1.0, 1.0, 3.0, 3.0])
    out = O.forward_points(gm, (0.5, 2.0, 0.5, 2.0), bx, by, sx, sy, 2, 2, True, q)
    assert np.isnan(out[0]) and np.isnan(out[1]) and np.isnan(out[4]) and np.isnan(out[5])
    assert abs(out[2] - 1.0) < 1e-12 and abs(out[3] - (1.0 + b * np.sin(1.0))) < 1e-5


def test_ensemble_distribution_functions():
    """The ensemble's Landau (203) and two-stream (204) distributions: values by their closed forms,
    mass 4 pi on x in [0, 4 pi), v in [-8, 8) (both v-profiles integrate to 1 up to e^-32 tails),
    and the density's first x-mode equal to the amplitude."""
    n = 256
    b = (0.0, 2 * O.TWO_PI, -8.0, 8.0)
    g = O.grid(n, n, b)
    hx, hv = 2 * O.TWO_PI / n, 16.0 / n
    x = (np.arange(n) * hx)[None, :].repeat(n, 0).ravel()
    v = (-8.0 + np.arange(n) * hv)[:, None].repeat(n, 1).ravel()
    xy = np.empty(2 * n * n)
    xy[0::2], xy[1::2] = x, v
    for ic, prof in ((203, lambda v: np.exp(-v * v / 2)), (204, lambda v: v * v * np.exp(-v * v / 2))):
        for a in (0.005, 0.03):
            f = O.sample_init(g, xy, ic, [a])
            ref = (1 + a * np.cos(0.5 * x)) * prof(v) / np.sqrt(2 * np.pi)
            assert np.max(np.abs(f - ref)) <= 1e-15 * np.max(np.abs(ref)) + 1e-300
            mass = hx * hv * f.sum()
            assert abs(mass - 4 * np.pi) < 1e-9, (ic, a, mass)
            rho = hv * f.reshape(n, n).sum(axis=0)
            mode = np.fft.rfft(rho)[1] / (n / 2)
            assert abs(abs(mode) - a) < 1e-9, (ic, a, mode)
