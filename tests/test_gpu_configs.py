"""The workloads the benchmark and its companion configurations are quoted on, run on the GPU and
compared with the oracle on identical inputs.

- Initial conditions new to this engine (no reference counterpart): the double shear layer sampled
  on the GPU, and the random-spectrum field (seed 0) built by the engine's own spectrum upload and
  transforms, against the oracle's.
- The 1024^2 map / 2048^2 psi / 4096^2 fine random-spectrum case (seed 0, dt 1/256, RK3, 4th-order
  stencil, eps 1e-3, freq_cut_psi 256, 4096 RK3 particles seeded by SplitMix64 seed 0): discrete
  initial condition, Lagrange warm-up and one step; map, every psi level, omega on the coarse grid
  and particle positions within 1e-12 relative L-inf (derivative planes in their own scale).
- The double shear layer at 128^2 map / 256^2 psi / 512^2 fine, dt 1/64, to T = 1 with remapping at
  E_inc > 1e-4: the same remap steps, E_inc within 2e-11 absolute per step, omega, psi and the map at
  T within 1e-10 relative L-inf (64 steps and a remap accumulate rounding beyond one step's 1e-12).
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev_array, hermite_errors, host, rel_linf

pytestmark = pytest.mark.gpu

TWO_PI = O.TWO_PI


def h_of(n):
    return TWO_PI / float(np.float32(n))


def test_double_shear_layer_sampled_on_gpu(dev):
    import cmm
    import torch
    n = 512
    g, og = cmm.grid(n), O.grid(n)
    # positions of the identity map on the 512^2 grid: the initial condition at the grid points
    cx, cy = O.identity_map(O.grid(128))
    out = torch.zeros(n * n, dtype=torch.float64, device=dev)
    cmm.sample_init(cmm.grid(128), g, dev_array(cx), dev_array(cy), None, 0, cmm.IC["double_shear_layer"], out)
    got = host(out)
    xy = O.compose_positions(O.grid(128), og, cx, cy, None, None, 0)
    ref = O.sample_init(og, xy, 100)
    print("double shear layer rel Linf", rel_linf(got, ref))
    assert rel_linf(got, ref) <= 1e-14


def c2_engine(particles):
    import cmm
    g = cmm.grid
    e = cmm.Euler(g(1024), g(4096), g(2048), g(4096), 1.0 / 256, freq_cut_psi=256.0, map_epsilon=1e-3,
                  incomp_threshold=1e-4, initial_condition="random_spectrum", seed=0, stack_capacity=4)
    s = e.add_particles(particles, 30)
    return e, s


def test_random_spectrum_case_one_step(dev):
    import cmm
    pos = cmm.uniform_particles(4096, 0)
    e, s = c2_engine(pos)
    # the engine's discrete initial condition against the oracle's field and Hermite form
    nd = 4096 * 4096
    disc = e.field("disc")
    w_ref = O.ic_random_spectrum(4096, 0)
    H_ref = O.hermite_from_real(O.grid(4096), w_ref)
    errs = hermite_errors(disc, H_ref, nd, np.pi / h_of(4096))
    print("random-spectrum initial condition (per plane)", errs)
    assert max(errs) <= 1e-12
    warm = e.init()
    info = e.step(1.0 / 256)
    cx, cy, psi, w = e.field("chix"), e.field("chiy"), e.field("psi"), e.field("w_coarse")
    part = e.particles(s)
    e.close()
    r = O.euler_run(O.grid(1024), O.grid(4096), O.grid(2048), O.grid(4096), dt=1.0 / 256, final_time=1.0 / 256,
                    fcut=256.0, disc=H_ref, g_disc=O.grid(4096), stack_capacity=4, particles=pos,
                    particle_scheme=30, want=("chix", "chiy", "psi", "vort_coarse", "incomp", "time", "map_counter"))
    assert r["steps"] == 1 and info.stack_size == 0
    print("warm-up E_inc", warm, r["warmup_incomp"])
    d_warm = np.abs(np.array(warm) - np.array(r["warmup_incomp"]))
    assert len(warm) == len(r["warmup_incomp"]) and d_warm.max() <= 2e-11
    print("E_inc", info.incomp, r["incomp"][0])
    assert abs(info.incomp - r["incomp"][0]) <= 2e-11
    nc, npsi = 1024 * 1024, 2048 * 2048
    for name, got, ref in (("X", cx, r["chix"]), ("Y", cy, r["chiy"])):
        errs = hermite_errors(got, ref, nc, 1.0 / 1e-3)
        print("map", name, errs)
        assert max(errs) <= 1e-12
    for lev in range(3):
        sl = slice(4 * npsi * lev, 4 * npsi * (lev + 1))
        errs = hermite_errors(psi[sl], r["psi"][sl], npsi, np.pi / h_of(2048))
        print("psi level", lev, errs)
        assert max(errs) <= 1e-12
    ew = rel_linf(w, r["vort_coarse"])
    ep = rel_linf(part, r["particles"])
    print("omega on the coarse grid", ew, "particles", ep)
    assert ew <= 1e-12 and ep <= 1e-12


def test_double_shear_layer_case_to_t1(dev):
    import cmm
    g = cmm.grid
    dt = 1.0 / 64
    e = cmm.Euler(g(128), g(512), g(256), g(512), dt, freq_cut_psi=128.0, initial_condition="double_shear_layer",
                  incomp_threshold=1e-4, stack_capacity=64)
    warm = e.init()
    infos = [e.step(dt) for _ in range(64)]
    cx, cy, psi, w = e.field("chix"), e.field("chiy"), e.field("psi"), e.field("w_coarse")
    e.close()
    r = O.euler_run(O.grid(128), O.grid(512), O.grid(256), O.grid(512), dt=dt, final_time=1.0, ic=100,
                    threshold=1e-4, fcut=128.0, stack_capacity=64,
                    want=("incomp", "time", "map_counter", "psi", "vort_coarse", "chix", "chiy"))
    einc = np.array([i.incomp for i in infos])
    counter = np.array([i.stack_size for i in infos], dtype=float)
    times = np.array([i.t for i in infos])
    print("remaps", counter.max(), "max |dE_inc|", np.abs(einc - r["incomp"]).max())
    assert np.array_equal(times, r["time"])
    assert np.array_equal(counter, r["map_counter"]) and counter.max() >= 1
    assert np.abs(np.array(warm) - np.array(r["warmup_incomp"])).max() <= 2e-11
    assert np.abs(einc - r["incomp"]).max() <= 2e-11
    n, npsi = 128 * 128, 256 * 256
    ew = rel_linf(w, r["vort_coarse"])
    print("omega", ew)
    assert ew <= 1e-10
    errs = hermite_errors(psi[:4 * npsi], r["psi"][:4 * npsi], npsi, np.pi / h_of(256))
    print("psi", errs)
    assert max(errs) <= 1e-10
    for name, got, ref in (("X", cx, r["chix"]), ("Y", cy, r["chiy"])):
        errs = hermite_errors(got, ref, n, 1.0 / 1e-3)
        print("map", name, errs)
        assert max(errs) <= 1e-10
