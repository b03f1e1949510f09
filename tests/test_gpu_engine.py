"""End-to-end GPU runs of the Euler engine (warm-up, steps, remap with the device stack) against
the reference's own fixtures (tests/golden/fixtures.json) and the CPU oracle.

Fixture tolerances: the reference's harness compares sums of absolute differences with
eps = 1e-13 on the same GPU architecture and states 1e-10..1e-12 across architectures
(Cuda-CMM/tests/test_functions.py:39, :68-71); statistics here are held to 1e-10 relative and
E_inc to 2e-11 absolute per step.  GPU against oracle: relative L-infinity <= 1e-12 for map
values, psi and omega after one step.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import hermite_errors, rel_linf

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))
TWO_PI = O.TWO_PI


def h256():
    # (x1 - x0) / (float)N in double, Cuda-CMM/src/grid/cmm-grid2d.cu:27
    return TWO_PI / float(np.float32(256))


def run_engine(dt, final_time, skip_remapping, steps=None):
    import cmm
    e = cmm.Euler(cmm.grid(256), cmm.grid(512), cmm.grid(256), cmm.grid(512), dt, freq_cut_psi=128.0,
                  skip_remapping=skip_remapping, incomp_threshold=1e-4, stack_capacity=11)
    warm = e.init()
    infos = []
    t = 0.0
    n = int(round(final_time / dt)) if steps is None else steps
    for _ in range(n):
        info = e.step(dt)
        infos.append((info.t, info.incomp, info.stack_size))
        t = info.t
    return e, warm, infos


def test_one_step_against_fixture_and_oracle(dev):
    fx = FIX["one_step"]
    e, warm, infos = run_engine(1.0, 1.0, True)
    # warm-up errors logged with 3 significant digits
    ref_w = fx["log_warmup_incomp"]
    assert len(warm) == len(ref_w)
    for a, b in zip(warm, ref_w):
        assert abs(a - b) <= 0.006 * abs(b)
    einc = infos[0][1]
    print("E_inc", einc, fx["Error_incompressibility"][0])
    assert abs(einc - fx["Error_incompressibility"][0]) <= 1e-10 * fx["Error_incompressibility"][0]
    n = 256 * 256
    h = float(h256())
    cx = e.field("chix")
    psi = e.field("psi")
    w = e.field("w_coarse")
    stats = {
        "Debug_globals/Map_ChiX_coarse_Min": cx[:n].min(),
        "Debug_globals/Map_ChiX_coarse_Max": cx[:n].max(),
        "Debug_globals/Map_ChiX_coarse_L2": cx[:n].sum() * h * h,
        "Debug_globals/Stream_function_Psi_psi_Min": psi[:n].min(),
        "Debug_globals/Stream_function_Psi_psi_Max": psi[:n].max(),
        "Debug_globals/Velocity_UX_psi_Max": psi[n:2 * n].max(),
        "Debug_globals/Velocity_UY_psi_Max": psi[2 * n:3 * n].max(),
        "Debug_globals/Vorticity_W_coarse_Min": w.min(),
        "Mesure/Energy": 0.5 * h * h * np.sum(psi[n:3 * n] ** 2),
        "Mesure/Enstrophy": 0.5 * h * h * np.sum(w ** 2),
    }
    for k, v in stats.items():
        ref = fx[k][1] if k.startswith("Mesure") else fx[k][1]
        print(f"{k}: {v!r} fixture {ref!r}")
        assert abs(v - ref) <= 1e-10 * abs(ref)
    # against the oracle's run of the same case
    r = O.euler_run(O.grid(256), O.grid(512), O.grid(256), O.grid(512), dt=1.0, final_time=1.0, skip_remapping=1,
                    want=("chix", "chiy", "psi", "w_herm"))
    cy = e.field("chiy")
    for name, g, ref in (("X", cx, r["chix"]), ("Y", cy, r["chiy"])):
        err = rel_linf(g[:n], ref[:n])
        print("map", name, "value plane rel Linf", err)
        assert err <= 1e-12
    err = rel_linf(psi[:3 * 4 * n], r["psi"])
    print("psi rel Linf", err)
    assert err <= 1e-12
    e.close()


def test_remapping_11b(dev):
    fx = FIX["remap_11b"]
    e, warm, infos = run_engine(1.0 / 32, 2.0, False)
    einc = np.array([i[1] for i in infos])
    ref = np.array(fx["Error_incompressibility"])
    counter = np.array([i[2] for i in infos], dtype=float)
    print("max abs E_inc diff", np.abs(einc - ref).max())
    assert np.array_equal(counter, np.array(fx["Map_counter"]))
    assert np.abs(einc - ref).max() <= 2e-11
    times = np.array([i[0] for i in infos])
    assert np.array_equal(times, np.array(fx["Time_s"]))
    n = 256 * 256
    h = float(h256())
    psi = e.field("psi")
    w = e.field("w_coarse")
    energy = 0.5 * h * h * np.sum(psi[n:3 * n] ** 2)
    enstrophy = 0.5 * h * h * np.sum(w ** 2)
    print("energy", energy, fx["Mesure/Energy"][4], "enstrophy", enstrophy, fx["Mesure/Enstrophy"][4])
    assert abs(energy - fx["Mesure/Energy"][4]) <= 1e-10 * fx["Mesure/Energy"][4]
    assert abs(enstrophy - fx["Mesure/Enstrophy"][4]) <= 1e-10 * fx["Mesure/Enstrophy"][4]
    e.close()


def test_vlasov_poisson_against_oracle(dev):
    """The 1D-1V Vlasov-Poisson loop (nonlinear Landau damping, map 64^2, fine = psi = vort 128^2
    on [0, 4 pi) x [-2.5 pi, 2.5 pi), RK3, dt 1/16, T 1, remap at E_inc > 1e-4, which here remaps
    every step) on the GPU against the oracle's run: the same remap steps, E_inc to 1e-10 absolute,
    psi to 1e-10 relative L-inf; diagnostics against the closed-form initial values."""
    import cmm
    gc, gf = cmm.vp_grid(64), cmm.vp_grid(128)
    dt, T = 1.0 / 16, 1.0
    e = cmm.Euler(gc, gf, gf, gf, dt, initial_condition="nonlin_landau_damping", incomp_threshold=1e-4,
                  stack_capacity=32, simulation_type="cmm_vlasov_poisson_1d")
    warm = e.init()
    d0 = e.vp_diagnostics()
    infos = [e.step(dt) for _ in range(int(round(T / dt)))]
    einc = np.array([i.incomp for i in infos])
    counter = np.array([i.stack_size for i in infos], dtype=float)
    psi = e.field("psi")
    e.close()
    r = O.euler_run(O.vp_grid(64), O.vp_grid(128), O.vp_grid(128), O.vp_grid(128), dt=dt, final_time=T, ic=202,
                    sim=1, threshold=1e-4, stack_capacity=32, want=("incomp", "time", "map_counter", "psi"))
    print("warm-up", warm, list(r["warmup_incomp"]))
    print("E_inc", einc, r["incomp"])
    print("maps", counter, r["map_counter"])
    err = rel_linf(psi, r["psi"])
    d_einc = np.abs(einc - r["incomp"])
    d_warm = np.abs(np.array(warm) - np.array(r["warmup_incomp"]))
    print("psi rel Linf", err, "max |dE_inc|", d_einc.max(), "warm-up", d_warm.max())
    assert counter.max() >= 1, "the case should remap at least once"
    assert np.array_equal(counter, r["map_counter"])
    # E_inc = max |1 - det grad X| is a cancellation: its rounding is absolute, a few 1e-11 here
    # (the warm-up steps are 1/16 long on a phase-space map with large v-derivatives, and the remap
    # every step re-samples f through the stack, so it accumulates)
    assert d_einc.max() <= 1e-10 and d_warm.max() <= 1e-10
    assert err <= 1e-10
    # initial diagnostics: mass 4 pi (Maxwellian over [-2.5 pi, 2.5 pi) times (1 + 0.5 cos(x/2))),
    # kinetic energy 2 pi, potential energy 1/2 int (0.5 sin(x/2)/0.5 ... ) = 1/2 * 4 pi/2 * 1 = pi,
    # momentum 0, first field mode |E_1| = 0.5
    Etot, Ekin, Epot, mass, mom = d0[:5]
    print("diagnostics at t=0", d0)
    assert abs(mass - 4 * np.pi) < 1e-9 and abs(Ekin - 2 * np.pi) < 1e-9
    assert abs(Epot - np.pi) < 1e-9 and abs(Etot - (Ekin + Epot)) < 1e-12 and abs(mom) < 1e-9
    assert abs(d0[5] - 0.5) < 1e-9 and max(d0[6:9]) < 1e-9


def test_particles_in_the_loop(dev):
    """Fluid particles advanced by the device loop (on its second stream, overlapping the next
    step) against the oracle's run: 4_nodes, map 64^2, fine = psi = vort 128^2, dt 1/16, 16 steps
    with a remap threshold that remaps, 500 particles, RK3 and RK4; positions to 1e-12 relative."""
    import cmm
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, TWO_PI, 1000)
    for scheme, tin, L in ((30, "RK3", 3), (40, "RK4", 4)):
        e = cmm.Euler(cmm.grid(64), cmm.grid(128), cmm.grid(128), cmm.grid(128), 1.0 / 16, time_integration=tin,
                      incomp_threshold=1e-4, stack_capacity=64)
        s = e.add_particles(pos, scheme=scheme)
        e.init()
        infos = [e.step(1.0 / 16) for _ in range(16)]
        got = e.particles(s)
        e.close()
        r = O.euler_run(O.grid(64), O.grid(128), O.grid(128), O.grid(128), dt=1.0 / 16, final_time=1.0, tin=scheme,
                        l_order=L, threshold=1e-4, stack_capacity=64, particles=pos, particle_scheme=scheme,
                        want=("incomp", "time", "map_counter", "particles"))
        maps = [i.stack_size for i in infos]
        print(scheme, "maps", maps[-1], "particles rel Linf", rel_linf(got, r["particles"]))
        assert maps == list(r["map_counter"]) and maps[-1] >= 1
        assert rel_linf(got, r["particles"]) <= 1e-12


def test_zoom_and_restart_against_oracle(dev):
    """A run that remaps (map 64^2, fine 128^2, threshold 1e-5), then: a zoom window (non-square,
    off-centre, cmm_euler_zoom) against the oracle's composition through the same map and stack,
    its initial conditions (own, passive scalar ring, Landau distribution) and psi there; and a
    second engine restarted from that state (cmm_euler_restart) against the oracle's translate and
    psi solve."""
    import cmm
    gc, gf = cmm.grid(64), cmm.grid(128)
    e = cmm.Euler(gc, gf, gf, gf, 1.0 / 32, incomp_threshold=1e-5, stack_capacity=16)
    e.init()
    for _ in range(24):
        info = e.step(1.0 / 32)
    nstack = info.stack_size
    assert nstack >= 2, "the run should have remapped"
    cx, cy = e.field("chix"), e.field("chiy")
    maps = [e.stack_map(m) for m in range(nstack)]
    psi = e.field("psi")
    npsi = 4 * 128 * 128
    gz = cmm.grid(48, 40, (1.0, 3.5, 2.0, 4.25))
    ics = [-1, 302, 200]
    z = e.zoom(gz, ics)
    ogc, ogf = O.grid(64), O.grid(128)
    ogz = O.grid(48, 40, (1.0, 3.5, 2.0, 4.25))
    sx = np.concatenate([m[0] for m in maps])
    sy = np.concatenate([m[1] for m in maps])
    xy = O.compose_positions(ogc, ogz, cx, cy, sx, sy, nstack)
    # positions on the torus (a point within rounding of the wrap can land on either side)
    d = np.concatenate([z["x"] - xy[0::2], z["y"] - xy[1::2]])
    d = np.mod(d + np.pi, 2 * np.pi) - np.pi
    print("zoom positions", np.abs(d).max())
    assert np.abs(d).max() <= 1e-12 * TWO_PI
    for k, ic in enumerate(ics):
        ref = O.sample_init(ogz, xy, 0 if ic < 0 else ic)
        print("zoom ic", ic, rel_linf(z["ic"][k], ref))
        assert rel_linf(z["ic"][k], ref) <= 1e-12
    L = O.lib()
    # the zoom grid's spacing is (x1 - x0)/(float)N as for every grid
    p0 = np.ascontiguousarray(psi[:npsi])
    hx, hy = L.cmmo_grid_h(48, 1.0, 3.5), L.cmmo_grid_h(40, 2.0, 4.25)
    ref_psi = np.array([L.cmmo_hermite_eval(ogf, O.ptr(p0), 1.0 + i * hx, 2.0 + j * hy)
                        for j in range(40) for i in range(48)])
    print("zoom psi", rel_linf(z["psi"], ref_psi))
    assert rel_linf(z["psi"], ref_psi) <= 1e-12
    # restart a second engine from this state
    r = cmm.Euler(gc, gf, gf, gf, 1.0 / 32, incomp_threshold=1e-5, stack_capacity=16)
    r.restart(0.75, maps, cx, cy, [psi[npsi:2 * npsi], psi[2 * npsi:3 * npsi]])
    w_ref = O.hermite_from_real(ogf, O.sample_init(ogf, O.compose_positions(ogc, ogf, *O.identity_map(ogc), sx, sy,
                                                                            nstack), 0))
    w = r.field("w_fine")
    print("restart omega", rel_linf(w[:128 * 128], w_ref[:128 * 128]))
    assert rel_linf(w[:128 * 128], w_ref[:128 * 128]) <= 1e-12
    psi_ref = O.psi_from_vort(ogc, ogf, ogf, ogf, cx, cy, w_ref, 0, 128.0)
    rp = r.field("psi")
    print("restart psi", rel_linf(rp[:128 * 128], psi_ref[:128 * 128]))
    assert rel_linf(rp[:128 * 128], psi_ref[:128 * 128]) <= 1e-12
    assert np.array_equal(rp[npsi:], psi[npsi:])
    info = r.step(1.0 / 32)
    assert abs(info.t - (0.75 + 1.0 / 32)) < 1e-15
    e.close()
    r.close()


def test_forward_map_in_the_loop(dev):
    """forward_map = 1 (map 64^2, fine 128^2, threshold 1e-5 so that the run remaps): E_inc of both
    maps and the invertibility error per step and the map counter against the oracle's loop, the
    final forward map, its composition through the forward stack and forwarded points; then a
    remap requested below the threshold (cmm_euler_request_remap) pushes both maps and resets the
    forward map to the identity."""
    import cmm
    gc, gf = cmm.grid(64), cmm.grid(128)
    dt, steps = 1.0 / 32, 24
    e = cmm.Euler(gc, gf, gf, gf, dt, incomp_threshold=1e-5, stack_capacity=16, forward_map=True)
    e.init()
    assert np.array_equal(e.field("chix_f"), O.identity_map(O.grid(64))[0])
    inc, inc_f, inv, ctr = [], [], [], []
    for _ in range(steps):
        info = e.step(dt)
        inc.append(info.incomp)
        inc_f.append(info.incomp_forward)
        inv.append(info.invertibility)
        ctr.append(info.stack_size)
    ogc, ogf = O.grid(64), O.grid(128)
    r = O.euler_run(ogc, ogf, ogf, ogf, dt=dt, final_time=steps * dt, threshold=1e-5, stack_capacity=16, forward_map=1,
                    want=("incomp", "time", "map_counter", "fchix", "fchiy"))
    print("E_inc", inc[:3], r["incomp"][:3])
    print("E_inc_f", inc_f[:3], r["incomp_fwd"][:3])
    print("E_inv", inv[:3], r["invert"][:3])
    assert np.array_equal(np.array(ctr, dtype=float), r["map_counter"]) and ctr[-1] >= 2
    assert np.all(np.abs(np.array(inc) - r["incomp"]) <= 1e-10 * np.abs(r["incomp"]) + 2e-11)
    assert np.all(np.abs(np.array(inc_f) - r["incomp_fwd"]) <= 1e-10 * np.abs(r["incomp_fwd"]) + 2e-11)
    # |x - X_b(X_f(x))| is a cancellation between positions of size 2 pi: its rounding level is
    # absolute, ~1e-13 after 24 steps of accumulated map rounding
    assert np.all(np.abs(np.array(inv) - r["invert"]) <= 1e-10 * np.abs(r["invert"]) + 2e-12)
    fx, fy = e.field("chix_f"), e.field("chiy_f")
    for name, g, ref in (("X_f", fx, r["fchix"]), ("Y_f", fy, r["fchiy"])):
        errs = hermite_errors(g, ref, 64 * 64, 1.0 / 1e-3)
        print("forward map", name, errs)
        assert max(errs) <= 1e-10
    n_st = ctr[-1]
    maps = [e.forward_stack_map(m) for m in range(n_st)]
    sx = np.concatenate([m[0] for m in maps])
    sy = np.concatenate([m[1] for m in maps])
    gt = cmm.grid(96)
    x, y = e.forward_positions(gt)
    xy = O.forward_positions(ogc, O.grid(96), fx, fy, sx, sy, n_st)
    d = np.mod(np.concatenate([x - xy[0::2], y - xy[1::2]]) + np.pi, 2 * np.pi) - np.pi
    print("forward positions", np.abs(d).max())
    assert np.abs(d).max() <= 1e-12 * TWO_PI
    pts = np.random.default_rng(5).uniform(0.01, TWO_PI - 0.01, size=(300, 2)).ravel()
    for first, cur in ((0, True), (1, True), (1, False)):
        got = e.forward_points(gc, pts, first, cur)
        ref = O.forward_points(ogc, (0.0, TWO_PI, 0.0, TWO_PI), fx, fy, sx, sy, n_st, first, cur, pts)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        dd = np.mod(got - ref + np.pi, 2 * np.pi) - np.pi
        assert np.nanmax(np.abs(dd)) <= 1e-12 * TWO_PI
    # a requested remap below the threshold
    e.request_remap()
    info = e.step(dt)
    assert info.remapped == 1 and info.stack_size == n_st + 1
    assert np.array_equal(e.field("chix_f"), O.identity_map(ogc)[0])
    assert not np.array_equal(e.forward_stack_map(n_st)[0], O.identity_map(ogc)[0])
    e.close()


@pytest.mark.parametrize("tin,L,schemes", [("RK3", 3, (30, 30)), ("RK4", 4, (41, 40))])
def test_particle_sets_start_time_and_inertia(dev, tin, L, schemes):
    """Particle sets with the reference's per-set settings in the device loop (4_nodes, map 64^2,
    fine = psi = vort 128^2, dt 1/16, 16 steps, remapping at 1e-4) against the oracle's run of each
    set: a fluid set starting at t = 0.5, an inertial set (tau 0.1, fluid start velocity) from t = 0
    and an inertial set (tau 0.05, zero start velocity) starting at t = 0.25; positions and
    velocities within 1e-12 relative, and the fluid set's velocity output (d psi/dx, d psi/dy at the
    positions) against the oracle's psi."""
    import cmm
    rng = np.random.default_rng(11)
    dt = 1.0 / 16
    sets = [dict(scheme=schemes[0], init_time=0.5, tau=0.0, init_vel=False),
            dict(scheme=schemes[1], init_time=0.0, tau=0.1, init_vel=True),
            dict(scheme=schemes[1], init_time=0.25, tau=0.05, init_vel=False)]
    pos = [rng.uniform(0, TWO_PI, 400) for _ in sets]
    e = cmm.Euler(cmm.grid(64), cmm.grid(128), cmm.grid(128), cmm.grid(128), dt, time_integration=tin,
                  incomp_threshold=1e-4, stack_capacity=64)
    ids = [e.add_particles(p, **kw) for p, kw in zip(pos, sets)]
    e.init()
    mid = None
    for k in range(16):
        e.step(dt)
        if k == 5:
            mid = [e.particles(i) for i in ids]
    got = [(e.particles(i), e.particle_velocities(i)) for i in ids]
    psi = e.field("psi")
    e.close()
    # the fluid set has not moved before its start (t = 6/16 < 0.5)
    assert np.array_equal(mid[0], pos[0])
    og = O.grid(128)
    for k, (kw, p0) in enumerate(zip(sets, pos)):
        r = O.euler_run(O.grid(64), og, og, og, dt=dt, final_time=1.0, tin={3: 30, 4: 40}[L], l_order=L,
                        threshold=1e-4, stack_capacity=64, particles=p0,
                        particle_scheme=kw["scheme"], particle_init_time=kw["init_time"], particle_tau=kw["tau"],
                        particle_init_vel=int(kw["init_vel"]), want=("incomp", "map_counter", "psi"))
        ep = rel_linf(got[k][0], r["particles"])
        print(tin, "set", k, kw, "positions", ep)
        assert ep <= 1e-12
        if kw["tau"] != 0:
            ev = rel_linf(got[k][1], r["particle_vel"])
            print("   velocities", ev)
            assert ev <= 1e-12
        else:
            # (d psi/dx, d psi/dy) of the newest psi at the positions
            p0r = np.ascontiguousarray(r["psi"][:4 * 128 * 128])
            ref = np.zeros_like(got[k][1])
            u = np.zeros(2)
            xy = r["particles"]
            for j in range(len(xy) // 2):
                O.lib().cmmo_hermite_grad(og, O.ptr(p0r), 1, O.C.c_double(xy[2 * j]), O.C.c_double(xy[2 * j + 1]),
                                          O.ptr(u))
                ref[2 * j], ref[2 * j + 1] = -u[1], u[0]
            ok = np.isfinite(got[k][1])
            print("   velocity output", rel_linf(got[k][1][ok], ref[ok]), "outside", (~ok).sum())
            assert rel_linf(got[k][1][ok], ref[ok]) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_check_in_the_sampler_pass(dev, precision):
    """The step's incompressibility check inside the psi sampler's pass (cmm_euler_set_check_fusion)
    against the check kernel of its own, on the double shear layer (map 128^2, fine = vort 512^2,
    psi 256^2, dt 1/64) with a threshold that remaps every few steps: the same remap steps, the same
    E_inc and the same psi.  The check points and the terms of the sums are the same; the compiler
    may contract products into fused multiply-adds differently in the two kernels, and the sums
    cancel about 6/h of magnitude to h, so E_inc is held to 1e-12 absolute (measured 7e-15).  psi
    is bit for bit the same."""
    import cmm
    runs = []
    for fused in (False, True):
        e = cmm.Euler(cmm.grid(128), cmm.grid(512), cmm.grid(256), cmm.grid(512), 1.0 / 64, freq_cut_psi=128.0,
                      initial_condition="double_shear_layer", incomp_threshold=8e-6, stack_capacity=32,
                      precision=precision)
        e.set_check_fusion(fused)
        e.init()
        infos = [e.step(1.0 / 64) for _ in range(32)]
        runs.append(([x.incomp for x in infos], [x.remapped for x in infos], e.field("psi")))
        e.close()
    (e0, r0, p0), (e1, r1, p1) = runs
    print(precision, "remaps", sum(r0), "E_inc max", max(e0), "max diff", max(abs(a - b) for a, b in zip(e0, e1)))
    assert sum(r0) >= 2 and r0 == r1
    assert all(abs(a - b) <= 1e-12 for a, b in zip(e0, e1))
    assert np.array_equal(p0, p1)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_psi_precombined_by_the_transform(dev, precision):
    """The next step's Lagrange-combined psi fields written by the psi solve's last transform
    (cmm_euler_set_psi_precombine) against the combination kernel of the advection: the double
    shear layer (map 128^2, fine = vort 512^2, psi 256^2, dt 1/64) with remaps (the remap's psi
    solve writes them again), a shortened step (whose weights differ, so the advection combines
    anew) and the steps after it.  The sums are the same, so the runs agree bit for bit: map, psi
    levels and every E_inc."""
    import cmm
    dt = 1.0 / 64
    plan = [dt] * 12 + [dt / 2] + [dt] * 12
    runs = []
    for pre in (False, True):
        e = cmm.Euler(cmm.grid(128), cmm.grid(512), cmm.grid(256), cmm.grid(512), dt, freq_cut_psi=128.0,
                      initial_condition="double_shear_layer", incomp_threshold=8e-6, stack_capacity=32,
                      precision=precision)
        e.set_psi_precombine(pre)
        e.init()
        infos = [e.step(d) for d in plan]
        runs.append(([x.incomp for x in infos], [x.remapped for x in infos], e.field("psi"), e.field("chix"),
                     e.field("chiy")))
        e.close()
    (e0, r0, *f0), (e1, r1, *f1) = runs
    print(precision, "remaps", sum(r0))
    assert sum(r0) >= 1 and r0 == r1
    assert e0 == e1
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)
