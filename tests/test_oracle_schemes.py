"""Pins of the oracle's time integrators and footpoint stencils by what the mathematics fixes.

Time schemes (Cuda-CMM/src/numerical/cmm-timestep.cu:28-470, particles
Cuda-CMM/src/numerical/cmm-particles.cu:119-276):
- a spatially uniform velocity that is a polynomial in time, u = (p(t), q(t)), is extrapolated
  exactly by Lagrange weights over L levels when deg p < L, so every scheme must return the
  quadrature it is exact for: the footpoint is x - int_{t_n}^{t_{n+1}} u dt to rounding.  A stage
  taken at the wrong time (L1 in place of L12), a wrong final weight or a wrong level index moves
  it by O(dt^2) and fails;
- on a non-uniform, time-dependent flow whose Hermite form is exact (a bicubic polynomial stream
  function times a linear factor in time) the local error of a step of order p falls by 2^(p+1)
  when dt halves; the exact footpoint comes from a fine RK4 integration written here.  A stage
  evaluated at a wrong position degrades the order.

Footpoint stencils (Cuda-CMM/src/simulation/cmm-simulation-host.cu:77-90 with
Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:226-289): with psi = 0 the footpoints are the
stencil points, and the new map at a node is the stencil applied to the old map.  For a
polynomial old map the averages over the four diagonal points at distance k*eps are moments of
(x + s k eps)^a, so the 2nd/4th/6th-order stencils reproduce polynomials of total degree 1/3/5
(and every bicubic for the 6th order) exactly, and leave closed-form residuals beyond that:
x^2 -> x^2 + eps^2 (2nd order), x^2 y^2 -> x^2 y^2 - 4 eps^4 (4th order).
"""
import numpy as np
import pytest

import oracle_lib as O

TWO_PI = O.TWO_PI
N = 32
G = O.grid(N)
H = TWO_PI / float(np.float32(N))
XS = np.arange(N) * H
C0 = np.pi          # polynomials are centred on the middle of the domain, far from the seam


def lagrange_weights(nodes, target):
    """Lagrange basis polynomials over `nodes` evaluated at `target` (textbook formula)."""
    w = []
    for i, ti in enumerate(nodes):
        v = 1.0
        for j, tj in enumerate(nodes):
            if j != i:
                v *= (target - tj) / (ti - tj)
        w.append(v)
    return np.array(w)


def psi_uniform(pv, qv):
    """Hermite planes of psi = p*y - q*x: velocity (psi_y, -psi_x) = (p, q) everywhere in the
    interior (the value plane is not periodic, so only cells away from the seam are exact)."""
    Xg, Yg = np.meshgrid(XS, XS)
    return np.concatenate([(pv * Yg - qv * Xg).ravel(), np.full(N * N, -qv), np.full(N * N, pv), np.zeros(N * N)])


def psi_levels(p_poly, q_poly, times):
    """levels newest first: slot l holds psi at times[-1-l]"""
    return np.concatenate([psi_uniform(p_poly(t), q_poly(t)) for t in times[::-1]])


# backward schemes: (scheme, Lagrange order, polynomial degree in time it must integrate exactly,
# whether it reads level 1 as t_n - dt)
BACKWARD = [(21, 2, 1, False), (20, 2, 1, True), (30, 3, 2, False), (31, 3, 2, True), (40, 4, 3, False),
            (41, 4, 3, False)]


def time_setup(L, even, dt):
    if even:
        levels = [1.0 - (L - 1 - i) * dt for i in range(L)]
    else:
        gaps = [0.13, 0.08, 0.11][:L - 1]
        levels = [1.0 - sum(gaps[i:]) for i in range(L - 1)] + [1.0]
    return np.array(levels), levels[-1] + dt


@pytest.mark.parametrize("tin,L,deg,even", BACKWARD)
def test_backward_schemes_time_polynomial_uniform_flow(tin, L, deg, even):
    dt = 0.07
    levels, t_next = time_setup(L, even, dt)
    rng = np.random.default_rng(tin)
    cp, cq = rng.uniform(-1, 1, deg + 1), rng.uniform(-1, 1, deg + 1)
    p = np.polynomial.Polynomial(cp)
    q = np.polynomial.Polynomial(cq)
    psi = psi_levels(p, q, levels)
    t = np.append(levels, t_next)
    c = O.advect_coeffs(t, np.append(np.diff(t, prepend=t[0] - dt), []), 0, L, tin, 1, 1e-3)
    x0, y0 = C0 + 0.31, C0 - 0.27
    xf = O.footpoint(G, psi, c, x0, y0)
    dx = p.integ()(t_next) - p.integ()(levels[-1])
    dy = q.integ()(t_next) - q.integ()(levels[-1])
    assert abs(xf[0] - (x0 - dx)) < 2e-14 and abs(xf[1] - (y0 - dy)) < 2e-14, (xf, x0 - dx, y0 - dy)


def test_backward_euler_closed_form():
    # explicit Euler from t_{n+1}: x - dt * u(t_{n+1}), with u(t_{n+1}) extrapolated exactly from
    # two levels of a linear p; with one level it is x - dt * u(t_n)
    dt = 0.07
    p = np.polynomial.Polynomial([0.3, -0.8])
    q = np.polynomial.Polynomial([-0.2, 0.5])
    for L in (1, 2):
        levels, t_next = time_setup(L, False, dt)
        psi = psi_levels(p, q, levels)
        t = np.append(levels, t_next)
        c = O.advect_coeffs(t, np.full(len(t), dt), 0, L, 10, 1, 1e-3)
        xf = O.footpoint(G, psi, c, C0, C0)
        tu = t_next if L == 2 else levels[-1]
        assert abs(xf[0] - (C0 - dt * p(tu))) < 1e-14 and abs(xf[1] - (C0 - dt * q(tu))) < 1e-14


def test_time_polynomial_check_is_sharp():
    # the same closed form fails for the schemes of lower exactness one degree up: Heun (trapezoid)
    # on a quadratic and RK3Mod on a cubic, with levels enough to extrapolate them exactly
    dt = 0.07
    for tin, L, deg in ((21, 3, 2), (31, 4, 3)):
        levels, t_next = time_setup(L, True, dt)
        p = np.polynomial.Polynomial([0.1] * deg + [1.0])
        q = np.polynomial.Polynomial([0.0])
        psi = psi_levels(p, q, levels)
        t = np.append(levels, t_next)
        c = O.advect_coeffs(t, np.full(len(t), dt), 0, L, tin, 1, 1e-3)
        xf = O.footpoint(G, psi, c, C0, C0)
        exact = C0 - (p.integ()(t_next) - p.integ()(levels[-1]))
        assert abs(xf[0] - exact) > 1e-6


# a bicubic stream function (cubic in each variable), exact in the Hermite basis, times a linear
# factor in time
def P(X, Y):
    return 0.5 * X ** 2 * Y - 0.3 * Y ** 3 + 0.4 * X * Y + 0.2 * X ** 3 - 0.1 * X ** 2 * Y ** 2 + 0.6 * Y


def P_x(X, Y):
    return X * Y + 0.4 * Y + 0.6 * X ** 2 - 0.2 * X * Y ** 2


def P_y(X, Y):
    return 0.5 * X ** 2 - 0.9 * Y ** 2 + 0.4 * X - 0.2 * X ** 2 * Y + 0.6


def P_xy(X, Y):
    return X + 0.4 - 0.4 * X * Y


def s(t):
    return 1.0 + 0.4 * t


def psi_bicubic(scale):
    Xg, Yg = np.meshgrid(XS - C0, XS - C0)
    return scale * np.concatenate([f(Xg, Yg).ravel() for f in (P, P_x, P_y, P_xy)])


def flow(z, t, steady=False):
    X, Y = z[0] - C0, z[1] - C0
    return (1.0 if steady else s(t)) * np.array([P_y(X, Y), -P_x(X, Y)])


def exact_step(z, t0, t1, m=4000, steady=False):
    """RK4 with m substeps from t0 to t1 (t1 < t0 integrates backward)"""
    h = (t1 - t0) / m
    t = t0
    z = np.array(z, dtype=np.float64)
    for _ in range(m):
        k1 = flow(z, t, steady)
        k2 = flow(z + h / 2 * k1, t + h / 2, steady)
        k3 = flow(z + h / 2 * k2, t + h / 2, steady)
        k4 = flow(z + h * k3, t + h, steady)
        z = z + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        t += h
    return z


def observed_orders(errs):
    return [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


# (scheme, Lagrange order, order of the scheme)
ORDERS = [(10, 2, 1), (21, 2, 2), (20, 2, 2), (30, 3, 3), (31, 3, 3), (40, 4, 4), (41, 4, 4)]


@pytest.mark.parametrize("tin,L,order", ORDERS)
def test_backward_schemes_local_error_order(tin, L, order):
    z0 = (C0 + 0.3, C0 - 0.2)
    errs = []
    for dt in (0.16, 0.08, 0.04, 0.02):
        levels = [1.0 - (L - 1 - i) * dt for i in range(L)]
        t_next = 1.0 + dt
        psi = np.concatenate([psi_bicubic(s(tl)) for tl in levels[::-1]])
        t = np.append(levels, t_next)
        c = O.advect_coeffs(t, np.full(len(t), dt), 0, L, tin, 1, 1e-3)
        xf = O.footpoint(G, psi, c, *z0)
        errs.append(np.max(np.abs(xf - exact_step(z0, t_next, 1.0))))
    orders = observed_orders(errs)
    print(tin, errs, orders)
    assert abs(orders[-1] - (order + 1)) < 0.3 and abs(orders[-2] - (order + 1)) < 0.4


# particles move forward from t_n to t_{n+1}; their levels are newest first from t_{n+1}
# (slot 0 = t_{n+1}, slot 1 = t_n, slot 2 = t_n - dt ...), Lp = Lagrange weights to t_n + dt/2
PARTICLES = [(20, 2, 1, False), (30, 3, 2, False), (31, 3, 2, True), (40, 4, 3, False), (41, 4, 3, False)]


def particle_times(L, even, dt):
    base, t_next = time_setup(L - 1, even, dt) if L > 1 else (np.array([]), None)
    # levels t_{n+1}, t_n and older; base holds the L-1 older ones ending at t_n = 1.0
    levels = np.append(base, 1.0 + dt)
    return levels


@pytest.mark.parametrize("scheme,L,deg,even", PARTICLES)
def test_particle_schemes_time_polynomial_uniform_flow(scheme, L, deg, even):
    dt = 0.07
    levels = particle_times(L, even, dt)     # ascending, last = t_{n+1}
    rng = np.random.default_rng(scheme)
    p = np.polynomial.Polynomial(rng.uniform(-1, 1, deg + 1))
    q = np.polynomial.Polynomial(rng.uniform(-1, 1, deg + 1))
    psi = psi_levels(p, q, levels)
    Lp = lagrange_weights(levels[::-1], 1.0 + dt / 2)
    pos = np.array([C0 + 0.2, C0 - 0.1])
    out = O.particles_advect(G, psi, L, Lp, dt, scheme, pos)
    ex = pos + np.array([p.integ()(1 + dt) - p.integ()(1.0), q.integ()(1 + dt) - q.integ()(1.0)])
    assert np.max(np.abs(out - ex)) < 2e-14, (out, ex)


def test_particle_euler_closed_form():
    dt = 0.07
    p = np.polynomial.Polynomial([0.3, -0.8])
    q = np.polynomial.Polynomial([-0.2, 0.5])
    levels = particle_times(2, False, dt)
    psi = psi_levels(p, q, levels)
    pos = np.array([C0, C0])
    out = O.particles_advect(G, psi, 2, lagrange_weights(levels[::-1], 1 + dt / 2), dt, 10, pos)
    assert np.max(np.abs(out - (pos + dt * np.array([p(1.0), q(1.0)])))) < 1e-14


@pytest.mark.parametrize("scheme,L,order", [(10, 2, 1), (20, 2, 2), (30, 3, 3), (31, 3, 3), (40, 4, 4), (41, 4, 4)])
def test_particle_schemes_local_error_order(scheme, L, order):
    z0 = np.array([C0 + 0.3, C0 - 0.2])
    errs = []
    for dt in (0.16, 0.08, 0.04, 0.02):
        levels = np.array([1.0 - (L - 2 - i) * dt for i in range(L - 1)] + [1.0 + dt])
        psi = np.concatenate([psi_bicubic(s(tl)) for tl in levels[::-1]])
        Lp = lagrange_weights(levels[::-1], 1.0 + dt / 2)
        out = O.particles_advect(G, psi, L, Lp, dt, scheme, z0)
        errs.append(np.max(np.abs(out - exact_step(z0, 1.0, 1.0 + dt))))
    orders = observed_orders(errs)
    print(scheme, errs, orders)
    assert abs(orders[-1] - (order + 1)) < 0.3 and abs(orders[-2] - (order + 1)) < 0.4


@pytest.mark.parametrize("tin,order", [(10, 1), (21, 2), (30, 3), (40, 4)])
def test_forward_schemes_steady_local_error_order(tin, order):
    # the forward map's integrators in a steady flow (all levels equal): the reference's forward
    # RK4 takes its stage times from L1/L12 in an order of its own, which a steady flow does not see
    L = 4
    z0 = np.array([C0 + 0.3, C0 - 0.2])
    psi = np.concatenate([psi_bicubic(s(0.0))] * L)
    errs = []
    for dt in (0.16, 0.08, 0.04, 0.02):
        t = np.array([1.0 - (L - 1 - i) * dt for i in range(L)] + [1.0 + dt])
        c = O.advect_coeffs(t, np.full(len(t), dt), 0, L, tin, 1, 1e-3, direction=1)
        xf = O.footpoint(G, psi, c, *z0)
        # the steady flow with factor s(0) = 1, forward over dt
        errs.append(np.max(np.abs(xf - exact_step(z0, 0.0, dt, steady=True))))
    orders = observed_orders(errs)
    print(tin, errs, orders)
    assert abs(orders[-1] - (order + 1)) < 0.3 and abs(orders[-2] - (order + 1)) < 0.4


def test_forward_rk4mod_is_the_backward_formula():
    # the reference's forward RK4Mod case runs the backward formula (cmm-timestep.cu:384-430 is
    # called for direction 1 but integrates with -dt like :432-470)
    L = 4
    dt = 0.05
    psi = np.concatenate([psi_bicubic(1.0)] * L)
    t = np.array([1.0 - (L - 1 - i) * dt for i in range(L)] + [1.0 + dt])
    cf = O.advect_coeffs(t, np.full(len(t), dt), 0, L, 41, 1, 1e-3, direction=1)
    cb = O.advect_coeffs(t, np.full(len(t), dt), 0, L, 41, 1, 1e-3, direction=-1)
    z0 = (C0 + 0.3, C0 - 0.2)
    assert np.array_equal(O.footpoint(G, psi, cf, *z0), O.footpoint(G, psi, cb, *z0))


# ---- footpoint stencils -------------------------------------------------------------------------

def map_planes(f, fx, fy, fxy):
    Xg, Yg = np.meshgrid(XS - C0, XS - C0)
    return np.concatenate([g(Xg, Yg).ravel() for g in (f, fx, fy, fxy)])


def stencil_apply(muo, eps, mx, my):
    """new map with psi = 0 (footpoints = stencil points) from the old map planes mx, my"""
    dt = 0.1
    t = np.array([-2 * dt, -dt, 0.0, dt])
    c = O.advect_coeffs(t, np.full(4, dt), 0, 3, 30, muo, eps)
    return O.map_advect(G, G, mx, my, np.zeros(3 * 4 * N * N), c)


INNER = np.zeros((N, N), bool)
INNER[6:-6, 6:-6] = True
INNER = INNER.ravel()


def plane(a, p):
    return a[p * N * N:(p + 1) * N * N][INNER]


def check_exact(muo, fx_, fy_, eps=1e-3):
    """fx_, fy_: (f, f_x, f_y, f_xy) of the old map's X and Y components"""
    mx, my = map_planes(*fx_), map_planes(*fy_)
    nx, ny = stencil_apply(muo, eps, mx, my)
    # rounding of the stencil points (~1e-16 |X|) grows by 1/eps per derivative
    for new, old in ((nx, mx), (ny, my)):
        sc = np.max(np.abs(plane(old, 0))) + 1
        for p, m in ((0, 0), (1, 1), (2, 1), (3, 2)):
            err = np.max(np.abs(plane(new, p) - plane(old, p)))
            assert err < 64 * 2.2e-16 * sc / eps ** m, (muo, p, err)


def test_stencil_2nd_order():
    # exact for 1, x, y, xy and harmonic quadratics
    check_exact(0, (lambda X, Y: 0.3 + X - 0.2 * Y + 0.4 * X * Y + 0.1 * (X ** 2 - Y ** 2),
                    lambda X, Y: 1 + 0.4 * Y + 0.2 * X, lambda X, Y: -0.2 + 0.4 * X - 0.2 * Y,
                    lambda X, Y: 0.4 + 0 * X),
                (lambda X, Y: Y - 0.7 * X + 0.25 * X * Y, lambda X, Y: -0.7 + 0.25 * Y,
                 lambda X, Y: 1 + 0.25 * X, lambda X, Y: 0.25 + 0 * X))
    # x^2 -> x^2 + eps^2 in the value, derivatives exact
    eps = 0.05
    mx = map_planes(lambda X, Y: X ** 2, lambda X, Y: 2 * X, lambda X, Y: 0 * X, lambda X, Y: 0 * X)
    my = map_planes(lambda X, Y: Y ** 2, lambda X, Y: 0 * X, lambda X, Y: 2 * Y, lambda X, Y: 0 * X)
    nx, ny = stencil_apply(0, eps, mx, my)
    assert np.max(np.abs(plane(nx, 0) - (plane(mx, 0) + eps ** 2))) < 1e-13
    assert np.max(np.abs(plane(ny, 0) - (plane(my, 0) + eps ** 2))) < 1e-13
    assert np.max(np.abs(plane(nx, 1) - plane(mx, 1))) < 1e-11 and np.max(np.abs(plane(ny, 2) - plane(my, 2))) < 1e-11
    assert np.max(np.abs(plane(nx, 2))) < 1e-11 and np.max(np.abs(plane(nx, 3))) < 1e-9


def test_stencil_4th_order():
    # exact for total degree <= 3 (asymmetric in x and y so that a transposed plane fails)
    fX = (lambda X, Y: 0.3 * X ** 3 + 0.2 * X ** 2 * Y - 0.5 * X * Y ** 2 + 0.1 * Y ** 3 + X,
          lambda X, Y: 0.9 * X ** 2 + 0.4 * X * Y - 0.5 * Y ** 2 + 1,
          lambda X, Y: 0.2 * X ** 2 - 1.0 * X * Y + 0.3 * Y ** 2,
          lambda X, Y: 0.4 * X - 1.0 * Y)
    fY = (lambda X, Y: -0.2 * X ** 3 + 0.6 * X * Y ** 2 + Y + 0.3 * X,
          lambda X, Y: -0.6 * X ** 2 + 0.6 * Y ** 2 + 0.3,
          lambda X, Y: 1.2 * X * Y + 1,
          lambda X, Y: 1.2 * Y)
    check_exact(1, fX, fY)
    # x^2 y^2 -> x^2 y^2 - 4 eps^4 in the value; the derivative planes stay exact
    eps = 0.05
    mx = map_planes(lambda X, Y: X ** 2 * Y ** 2, lambda X, Y: 2 * X * Y ** 2, lambda X, Y: 2 * X ** 2 * Y,
                    lambda X, Y: 4 * X * Y)
    nx, _ = stencil_apply(1, eps, mx, mx.copy())
    assert np.max(np.abs(plane(nx, 0) - (plane(mx, 0) - 4 * eps ** 4))) < 1e-13
    for p in (1, 2, 3):
        assert np.max(np.abs(plane(nx, p) - plane(mx, p))) < 1e-9


def test_stencil_6th_order():
    # exact for every bicubic, x^3 y^3 included
    fX = (lambda X, Y: 0.1 * X ** 3 * Y ** 3 - 0.2 * X ** 3 * Y ** 2 + 0.3 * X ** 2 * Y + X,
          lambda X, Y: 0.3 * X ** 2 * Y ** 3 - 0.6 * X ** 2 * Y ** 2 + 0.6 * X * Y + 1,
          lambda X, Y: 0.3 * X ** 3 * Y ** 2 - 0.4 * X ** 3 * Y + 0.3 * X ** 2,
          lambda X, Y: 0.9 * X ** 2 * Y ** 2 - 1.2 * X ** 2 * Y + 0.6 * X)
    fY = (lambda X, Y: -0.15 * X ** 2 * Y ** 3 + 0.25 * X * Y ** 3 + Y,
          lambda X, Y: -0.3 * X * Y ** 3 + 0.25 * Y ** 3,
          lambda X, Y: -0.45 * X ** 2 * Y ** 2 + 0.75 * X * Y ** 2 + 1,
          lambda X, Y: -0.9 * X * Y ** 2 + 0.75 * Y ** 2)
    check_exact(2, fX, fY)
    # and the 4th-order stencil is not exact on the same map (the check is sharp)
    eps = 0.05
    mx, my = map_planes(*fX), map_planes(*fY)
    nx, _ = stencil_apply(1, eps, mx, my)
    assert np.max(np.abs(plane(nx, 0) - plane(mx, 0))) > 1e-7


# ---- inertial particles -------------------------------------------------------------------------
# x' = v, v' = -(v - u(t))/tau in a uniform flow u(t) = u0 + u1 t has the closed form
#   v(t) = u(t) - u1 tau + (v0 - u0(t0) + u1 tau) e^{-(t-t0)/tau},
#   x(t) = x0 + int_{t0}^{t} v,
# so one step of a scheme of order p misses it by O(dt^(p+1)), and the stage velocities must come
# from the right time levels for that order to hold.

def inertial_exact(x0, v0, u0, u1, tau, t0, t1):
    c = v0 - (u0 + u1 * t0) + u1 * tau
    T = t1 - t0
    dx = u0 * T + u1 * (t1 ** 2 - t0 ** 2) / 2 - u1 * tau * T + c * tau * (1 - np.exp(-T / tau))
    v = u0 + u1 * t1 - u1 * tau + c * np.exp(-T / tau)
    return x0 + dx, v


@pytest.mark.parametrize("scheme,L,order", [(10, 2, 1), (20, 2, 2), (30, 3, 3), (31, 3, 3), (40, 4, 4), (41, 4, 4)])
def test_inertial_particles_local_error_order(scheme, L, order):
    tau = 0.5
    u0 = np.array([0.3, -0.2])
    u1 = np.array([0.4, 0.25])
    x0 = np.array([C0 + 0.1, C0 - 0.3])
    v0 = np.array([-0.5, 0.7])
    errs = []
    for dt in (0.2, 0.1, 0.05, 0.025):
        levels = np.array([1.0 - (L - 2 - i) * dt for i in range(L - 1)] + [1.0 + dt])   # ascending, last t_{n+1}
        psi = np.concatenate([psi_uniform(u0[0] + u1[0] * t, u0[1] + u1[1] * t) for t in levels[::-1]])
        Lp = lagrange_weights(levels[::-1], 1.0 + dt / 2)
        pos, vel = O.particles_inertial(G, psi, L, Lp, dt, scheme, tau, x0, v0)
        ex, ev = inertial_exact(x0, v0, u0, u1, tau, 1.0, 1.0 + dt)
        errs.append(max(np.max(np.abs(pos - ex)), np.max(np.abs(vel - ev))))
    orders = observed_orders(errs)
    print(scheme, errs, orders)
    assert abs(orders[-1] - (order + 1)) < 0.3 and abs(orders[-2] - (order + 1)) < 0.45


def test_inertial_particles_follow_the_fluid_as_tau_vanishes():
    # with tau -> 0 an inertial particle released at the fluid velocity follows the fluid particle:
    # the lag is u'(t) tau to first order
    dt, steps = 2e-3, 200
    L = 3
    z = np.array([C0 + 0.3, C0 - 0.2])
    lag = []
    for tau in (4e-2, 2e-2, 1e-2):
        p_in = z.copy()
        p_fl = z.copy()
        X, Y = p_in[0] - C0, p_in[1] - C0
        v = s(1.0) * np.array([P_y(X, Y), -P_x(X, Y)])
        for k in range(steps):
            t = 1.0 + k * dt
            levels = np.array([t - (L - 2 - i) * dt for i in range(L - 1)] + [t + dt])
            psi = np.concatenate([psi_bicubic(s(tl)) for tl in levels[::-1]])
            Lp = lagrange_weights(levels[::-1], t + dt / 2)
            p_in, v = O.particles_inertial(G, psi, L, Lp, dt, 30, tau, p_in, v)
            p_fl = O.particles_advect(G, psi, L, Lp, dt, 30, p_fl)
        lag.append(np.max(np.abs(p_in - p_fl)))
    print("lag", lag)
    r1, r2 = lag[0] / lag[1], lag[1] / lag[2]
    assert 1.7 < r1 < 2.3 and 1.7 < r2 < 2.3 and abs(r2 - 2) < abs(r1 - 2)


def test_particle_start_time_in_the_loop():
    # a set with init_time = 4 dt stays put for the first three steps and then moves; an inertial set
    # keeps its start velocity rule: zero velocity until the step reaching init_time sets the fluid one
    ogc, ogf = O.grid(16), O.grid(32)
    dt = 1.0 / 16
    pos = np.array([1.0, 2.0, 3.0, 1.5])
    r3 = O.euler_run(ogc, ogf, ogf, ogf, dt=dt, final_time=3 * dt, particles=pos, particle_init_time=4 * dt,
                     want=("incomp",))
    assert np.array_equal(r3["particles"], pos)
    r4 = O.euler_run(ogc, ogf, ogf, ogf, dt=dt, final_time=4 * dt, particles=pos, particle_init_time=4 * dt,
                     want=("incomp",))
    assert not np.array_equal(r4["particles"], pos)
    # inertial, init_vel = 0: the velocity stays zero (and the particle still, since v = 0 and the
    # first stage moves it by dt * v) until the start; then the set moves
    ri = O.euler_run(ogc, ogf, ogf, ogf, dt=dt, final_time=3 * dt, particles=pos, particle_init_time=4 * dt,
                     particle_tau=0.1, particle_init_vel=1, want=("incomp",))
    assert np.array_equal(ri["particles"], pos)
    assert np.all(ri["particle_vel"] != 0)          # init_vel = 1: fluid velocity after the warm-up
