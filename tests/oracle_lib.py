"""ctypes binding of the CPU oracle (oracle/build/libcmm_oracle.so) for the tests.

Test infrastructure only: the oracle is the checker, never the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "libcmm_oracle.so")


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double)]


class AdvectCoeffs(C.Structure):
    _fields_ = [("L1", C.c_double * 4), ("L12", C.c_double * 4), ("c1", C.c_double * 12),
                ("cx", C.c_double * 12), ("cy", C.c_double * 12), ("cxy", C.c_double * 12),
                ("dt", C.c_double), ("eps", C.c_double), ("time_integration_num", C.c_int),
                ("map_update_order_num", C.c_int), ("l_order", C.c_int), ("direction", C.c_int)]


class EulerParams(C.Structure):
    _fields_ = [("g_coarse", Grid), ("g_fine", Grid), ("g_psi", Grid), ("g_vort", Grid),
                ("dt", C.c_double), ("final_time", C.c_double),
                ("time_integration_num", C.c_int), ("map_update_order_num", C.c_int), ("l_order", C.c_int),
                ("lagrange_init_higher_order", C.c_int),
                ("map_epsilon", C.c_double), ("incomp_threshold", C.c_double), ("freq_cut_psi", C.c_double),
                ("molly_stencil", C.c_int), ("skip_remapping", C.c_int),
                ("ic", C.c_int), ("ic_params", C.c_double * 10),
                ("discrete", C.c_int), ("g_disc", Grid), ("disc_herm", C.POINTER(C.c_double)),
                ("stack_capacity", C.c_int), ("n_particles", C.c_int), ("particle_scheme", C.c_int),
                ("particles", C.POINTER(C.c_double)), ("max_steps", C.c_int), ("sim", C.c_int),
                ("forward_map", C.c_int), ("particle_init_time", C.c_double), ("particle_tau", C.c_double),
                ("particle_init_vel", C.c_int), ("particle_vel_given", C.c_int),
                ("particle_vel", C.POINTER(C.c_double))]


class EulerResult(C.Structure):
    _fields_ = [("steps", C.c_int), ("warmup_incomp", C.c_double * 8), ("n_warmup", C.c_int),
                ("time", C.POINTER(C.c_double)), ("incomp", C.POINTER(C.c_double)),
                ("map_counter", C.POINTER(C.c_double)),
                ("chix", C.POINTER(C.c_double)), ("chiy", C.POINTER(C.c_double)),
                ("psi", C.POINTER(C.c_double)), ("w_herm", C.POINTER(C.c_double)),
                ("vort_coarse", C.POINTER(C.c_double)),
                ("incomp_fwd", C.POINTER(C.c_double)), ("invert", C.POINTER(C.c_double)),
                ("fchix", C.POINTER(C.c_double)), ("fchiy", C.POINTER(C.c_double))]


TWO_PI = 6.283185307179586476925286766559005768394338798750211641949889184615632812572
DP = C.POINTER(C.c_double)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        L.cmmo_incomp_error.restype = C.c_double
        L.cmmo_invert_error.restype = C.c_double
        L.cmmo_hermite_eval.restype = C.c_double
        L.cmmo_diffeo_det.restype = C.c_double
        L.cmmo_lagrange_coeff.restype = C.c_double
        L.cmmo_init_vorticity.restype = C.c_double
        L.cmmo_grid_h.restype = C.c_double
        L.cmmo_grid_h.argtypes = [C.c_int, C.c_double, C.c_double]
        L.cmmo_ic_random_spectrum.argtypes = [C.c_int, C.c_uint64, DP]
        L.cmmo_hermite_eval.argtypes = [Grid, DP, C.c_double, C.c_double]
        L.cmmo_hermite_grad.argtypes = [Grid, DP, C.c_int, C.c_double, C.c_double, DP]
        L.cmmo_diffeo_eval.argtypes = [Grid, DP, DP, C.c_double, C.c_double, DP, DP]
        L.cmmo_diffeo_det.argtypes = [Grid, DP, DP, C.c_double, C.c_double]
        L.cmmo_init_vorticity.argtypes = [C.c_double, C.c_double, C.c_int, DP]
        _lib = L
    return _lib


def grid(n, ny=None, b=(0.0, TWO_PI, 0.0, TWO_PI)):
    return Grid(n, n if ny is None else ny, b[0], b[1], b[2], b[3])


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def advect_coeffs(t, dt, loop_ctr, l_order, tin, muo, eps, direction=-1):
    c = AdvectCoeffs()
    t = np.ascontiguousarray(t, dtype=np.float64)
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    rc = lib().cmmo_advect_coeffs_fill(ptr(t), ptr(dt), loop_ctr, l_order, tin, muo, C.c_double(eps),
                                       direction, C.byref(c))
    assert rc == 0
    return c


def identity_map(g):
    n = g.nx * g.ny
    cx = np.zeros(4 * n)
    cy = np.zeros(4 * n)
    lib().cmmo_map_init_identity(g, ptr(cx), ptr(cy))
    return cx, cy


def map_advect(gm, gp, cx, cy, psi, coeffs):
    n = gm.nx * gm.ny
    ncx = np.zeros(4 * n)
    ncy = np.zeros(4 * n)
    rc = lib().cmmo_map_advect(gm, gp, ptr(cx), ptr(cy), ptr(psi), C.byref(coeffs), ptr(ncx), ptr(ncy))
    assert rc == 0
    return ncx, ncy


def incomp_error(gm, gc, cx, cy):
    return lib().cmmo_incomp_error(gm, gc, ptr(cx), ptr(cy))


def compose_positions(gm, gt, cx, cy, stack_x, stack_y, n_stack):
    xy = np.zeros(2 * gt.nx * gt.ny)
    sx = stack_x if stack_x is not None else np.zeros(1)
    sy = stack_y if stack_y is not None else np.zeros(1)
    lib().cmmo_compose_positions(gm, gt, ptr(cx), ptr(cy), ptr(sx), ptr(sy), n_stack, ptr(xy))
    return xy


def forward_positions(gm, gt, cx, cy, stack_x, stack_y, n_stack):
    xy = np.zeros(2 * gt.nx * gt.ny)
    sx = stack_x if stack_x is not None else np.zeros(1)
    sy = stack_y if stack_y is not None else np.zeros(1)
    lib().cmmo_forward_positions(gm, gt, ptr(cx), ptr(cy), ptr(sx), ptr(sy), n_stack, ptr(xy))
    return xy


def forward_points(gm, bounds, cx, cy, stack_x, stack_y, n_stack, first_map, apply_current, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros_like(pos)
    b = np.ascontiguousarray(bounds, dtype=np.float64)
    sx = stack_x if stack_x is not None else np.zeros(1)
    sy = stack_y if stack_y is not None else np.zeros(1)
    lib().cmmo_forward_points(gm, ptr(b), ptr(cx), ptr(cy), ptr(sx), ptr(sy), n_stack, first_map, int(apply_current),
                              pos.size // 2, ptr(pos), ptr(out))
    return out


def sample_init(gt, xy, ic, params=None, disc=None, gdisc=None, discrete=0):
    out = np.zeros(gt.nx * gt.ny)
    prm = np.zeros(10) if params is None else np.ascontiguousarray(params, dtype=np.float64)
    d = np.zeros(1) if disc is None else disc
    gd = gdisc if gdisc is not None else gt
    lib().cmmo_sample_init(gt, ptr(xy), ic, ptr(prm), ptr(d), gd, discrete, ptr(out))
    return out


def sample_vort(gm, gt, gf, cx, cy, w, molly=0):
    out = np.zeros(gt.nx * gt.ny)
    lib().cmmo_sample_vort(gm, gt, gf, ptr(cx), ptr(cy), ptr(w), molly, ptr(out))
    return out


def rfft2(a):
    ny, nx = a.shape
    out = np.zeros((ny, nx // 2 + 1), dtype=np.complex128)
    rc = lib().cmmo_rfft2(nx, ny, ptr(np.ascontiguousarray(a, dtype=np.float64)),
                          out.ctypes.data_as(DP))
    assert rc == 0
    return out


def irfft2(F, nx):
    ny = F.shape[0]
    out = np.zeros((ny, nx))
    Fc = np.ascontiguousarray(F, dtype=np.complex128)
    rc = lib().cmmo_irfft2(nx, ny, Fc.ctypes.data_as(DP), ptr(out))
    assert rc == 0
    return out


def hermite_from_real(g, real):
    out = np.zeros(4 * g.nx * g.ny)
    rc = lib().cmmo_hermite_from_real(g, ptr(np.ascontiguousarray(real, dtype=np.float64).ravel()), ptr(out))
    assert rc == 0
    return out


def psi_from_vort(gm, gv, gf, gp, cx, cy, w, molly, fcut):
    out = np.zeros(4 * gp.nx * gp.ny)
    rc = lib().cmmo_psi_from_vort(gm, gv, gf, gp, ptr(cx), ptr(cy), ptr(w), molly, C.c_double(fcut), ptr(out))
    assert rc == 0
    return out


def particles_advect(gp, psi, l_order, Lp12, dt, scheme, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
    Lp = np.zeros(4)
    Lp[:len(Lp12)] = Lp12
    lib().cmmo_particles_advect(gp, ptr(psi), l_order, ptr(Lp), C.c_double(dt), scheme, len(pos) // 2, ptr(pos))
    return pos


def particles_inertial(gp, psi, l_order, Lp12, dt, scheme, tau, pos, vel):
    pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).copy()
    Lp = np.zeros(4)
    Lp[:len(Lp12)] = Lp12
    lib().cmmo_particles_inertial(gp, ptr(psi), l_order, ptr(Lp), C.c_double(dt), scheme, C.c_double(tau),
                                  len(pos) // 2, ptr(pos), ptr(vel))
    return pos, vel


def ic_random_spectrum(n, seed):
    out = np.zeros(n * n)
    rc = lib().cmmo_ic_random_spectrum(n, seed, ptr(out))
    assert rc == 0
    return out


def spectral_from_real(g, real, op):
    out = np.zeros(4 * g.nx * g.ny)
    rc = lib().cmmo_spectral_from_real(g, ptr(np.ascontiguousarray(real, dtype=np.float64).ravel()), op, ptr(out))
    assert rc == 0, rc
    return out


def vp_psi_from_f(gm, gv, gf, gp, cx, cy, f_herm, molly=0):
    out = np.zeros(4 * gp.nx * gp.ny)
    rc = lib().cmmo_vp_psi_from_f(gm, gv, gf, gp, ptr(cx), ptr(cy), ptr(f_herm), molly, ptr(out))
    assert rc == 0, rc
    return out


def vp_grid(n, nv=None):
    """Vlasov-Poisson grid on x in [0, 4 pi), v in [-2.5 pi, 2.5 pi) (the reference's bounds
    {0, 4, -2.5, 2.5} times pi)."""
    return grid(n, nv, (0.0, 2 * TWO_PI, -1.25 * TWO_PI, 1.25 * TWO_PI))


def euler_run(g_coarse, g_fine, g_psi, g_vort, dt, final_time, tin=30, muo=1, l_order=3, lihi=1,
              eps=1e-3, threshold=1e-4, fcut=128.0, molly=0, skip_remapping=0, ic=0, ic_params=None,
              disc=None, g_disc=None, stack_capacity=1000, particles=None, particle_scheme=30,
              max_steps=0, want=("incomp", "time", "map_counter"), sim=0, forward_map=0, particle_init_time=0.0,
              particle_tau=0.0, particle_init_vel=0, particle_vel=None):
    p = EulerParams()
    p.sim = sim
    p.forward_map = forward_map
    p.g_coarse, p.g_fine, p.g_psi, p.g_vort = g_coarse, g_fine, g_psi, g_vort
    p.dt, p.final_time = dt, final_time
    p.time_integration_num, p.map_update_order_num, p.l_order = tin, muo, l_order
    p.lagrange_init_higher_order = lihi
    p.map_epsilon, p.incomp_threshold, p.freq_cut_psi = eps, threshold, fcut
    p.molly_stencil, p.skip_remapping, p.ic = molly, skip_remapping, ic
    if ic_params is not None:
        for i, v in enumerate(ic_params):
            p.ic_params[i] = v
    keep = []
    if disc is not None:
        p.discrete = 1
        p.g_disc = g_disc
        d = np.ascontiguousarray(disc, dtype=np.float64)
        keep.append(d)
        p.disc_herm = ptr(d)
    p.stack_capacity = stack_capacity
    if particles is not None:
        pos = np.ascontiguousarray(particles, dtype=np.float64).copy()
        keep.append(pos)
        p.n_particles = len(pos) // 2
        p.particle_scheme = particle_scheme
        p.particles = ptr(pos)
        p.particle_init_time = particle_init_time
        p.particle_tau = particle_tau
        p.particle_init_vel = int(particle_init_vel)
        if particle_tau != 0:
            vel = np.zeros(len(pos)) if particle_vel is None else np.ascontiguousarray(particle_vel,
                                                                                     dtype=np.float64).copy()
            p.particle_vel_given = int(particle_vel is not None)
            keep.append(vel)
            p.particle_vel = ptr(vel)
    p.max_steps = max_steps
    iter_max = int(1.05 * np.ceil(final_time / dt)) + 1
    res = EulerResult()
    out = {}
    for name in ("time", "incomp", "map_counter", "incomp_fwd", "invert"):
        a = np.zeros(iter_max)
        out[name] = a
        setattr(res, name, ptr(a))
    nc = g_coarse.nx * g_coarse.ny
    npsi = g_psi.nx * g_psi.ny
    nf = g_fine.nx * g_fine.ny
    sizes = {"chix": 4 * nc, "chiy": 4 * nc, "psi": 4 * npsi * l_order, "w_herm": 4 * nf, "vort_coarse": nc,
             "fchix": 4 * nc, "fchiy": 4 * nc}
    for name, n in sizes.items():
        if name in want:
            a = np.zeros(n)
            out[name] = a
            setattr(res, name, ptr(a))
    rc = lib().cmmo_euler_run(C.byref(p), C.byref(res))
    assert rc == 0
    steps = res.steps
    for name in ("time", "incomp", "map_counter", "incomp_fwd", "invert"):
        out[name] = out[name][:steps]
    out["steps"] = steps
    out["warmup_incomp"] = [res.warmup_incomp[i] for i in range(res.n_warmup)]
    if particles is not None:
        out["particles"] = pos
        if particle_tau != 0:
            out["particle_vel"] = vel
    return out


def footpoint(gp, psi, coeffs, x, y):
    """One footpoint of the oracle's time integrator (direction from coeffs) at (x, y)."""
    out = np.zeros(2)
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    rc = lib().cmmo_footpoint(gp, ptr(psi), C.byref(coeffs), C.c_double(x), C.c_double(y), ptr(out))
    assert rc == 0
    return out
