"""Ensembles of Vlasov-Poisson instances of mixed grid sizes stepped together (cmm_batch_*):
with the queued step work of all instances captured into one graph per step phase, or with every
kind of launch of all instances in one grouped, persistent launch (FP64), every instance must
follow exactly the steps it takes alone through cmm_euler_step (same block bodies, same order,
same coefficients), including remaps, in both precision forms.  The autotuner's choice must leave the check's value unchanged up to the
rounding of det near 1 (1e-13 absolute)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 16


def make(cmm, n, ic, amp, precision):
    g = cmm.ensemble_grid
    e = cmm.Euler(g(n), g(4 * n), g(2 * n), g(2 * n), DT, initial_condition=ic, ic_params=[amp],
                  simulation_type="cmm_vlasov_poisson_1d", incomp_threshold=1e-4, stack_capacity=32,
                  freq_cut_psi=1e9, precision=precision)
    e.init()
    return e


SPECS = [(32, "landau_amp", 0.05), (64, "two_stream_v2", 0.02), (32, "two_stream_v2", 0.04), (64, "landau_amp", 0.01)]


@pytest.mark.parametrize("precision,grouping", [("fp64", False), ("split", False), ("fp64", True)])
def test_batch_graph_matches_single_engines(dev, precision, grouping):
    import cmm
    steps = 60
    alone = [make(cmm, *s, precision) for s in SPECS]
    ref_infos = [[e.step(DT) for _ in range(steps)] for e in alone]
    batched = [make(cmm, *s, precision) for s in SPECS]
    b = cmm.Batch(batched, use_graphs=True, grouping=grouping)
    infos = [b.step(DT) for _ in range(steps)]
    st = b.stats()
    print("batch stats", st)
    remaps = [sum(i.remapped for i in r) for r in ref_infos]
    print("remaps per instance", remaps)
    if grouping:
        assert st["grouped_steps"] >= steps - 1
        assert st["grouped_steps"] + st["graph_launches"] + st["eager_steps"] == steps
    else:
        assert st["graph_launches"] == steps - 1 and st["eager_steps"] == 1 and st["grouped_steps"] == 0
    for k, (ea, eb) in enumerate(zip(alone, batched)):
        einc_a = np.array([r.incomp for r in ref_infos[k]])
        einc_b = np.array([infos[s][k].incomp for s in range(steps)])
        assert np.array_equal(einc_a, einc_b), (k, np.abs(einc_a - einc_b).max())
        assert [r.stack_size for r in ref_infos[k]] == [infos[s][k].stack_size for s in range(steps)]
        for f in ("chix", "chiy", "psi"):
            assert np.array_equal(ea.field(f), eb.field(f)), (k, f)
    assert sum(remaps) > 0
    b.close()
    for e in alone + batched:
        e.close()


def test_autotune_keeps_the_check(dev):
    import cmm
    e = make(cmm, 64, "landau_amp", 0.05, "fp64")
    for _ in range(5):
        e.step(DT)
    g = cmm.ensemble_grid
    before = e.field("chix"), e.field("chiy")
    import torch
    cx = torch.from_numpy(before[0]).cuda()
    cy = torch.from_numpy(before[1]).cuda()
    ref = cmm.incompressibility_check(g(256), g(64), cx, cy)
    t = cmm.autotune(e)
    print("autotune", t)
    assert t["tile"] in (1, 2, 4) and t["max_blocks"] > 0
    info = e.step(DT)
    assert info.incomp > 0
    # the check on the same map with the tuned launch shape gives the same value up to the
    # rounding of det near 1 (the tile variants group the same FP64 terms with other fusions)
    e2 = make(cmm, 64, "landau_amp", 0.05, "fp64")
    for _ in range(5):
        e2.step(DT)
    a = e2.step(DT).incomp
    print("E_inc default tile", a, "tuned", info.incomp)
    assert abs(a - info.incomp) <= 1e-13, (a, info.incomp)
    assert ref > 0
    e.close()
    e2.close()


@pytest.mark.parametrize("ic,code", [("landau_amp", 203), ("two_stream_v2", 204)])
def test_ensemble_conditions_against_oracle(dev, ic, code):
    """The ensemble's distribution functions (x in [0, 4 pi), v in [-8, 8), amplitude 0.04) through
    the FP64 Vlasov-Poisson loop (map 64^2, fine 256^2, psi and vort 128^2, RK3, dt 1/16, T 1, remap
    at 1e-4) against the oracle's run: the same remap steps, E_inc within 1e-10 absolute, psi within
    1e-10 relative L-inf (the tolerances of the reference's Vlasov-Poisson case)."""
    import cmm
    import oracle_lib as O
    from gpu_util import rel_linf
    g = cmm.ensemble_grid
    e = cmm.Euler(g(64), g(256), g(128), g(128), DT, initial_condition=ic, ic_params=[0.04],
                  simulation_type="cmm_vlasov_poisson_1d", incomp_threshold=1e-4, stack_capacity=32)
    e.init()
    infos = [e.step(DT) for _ in range(16)]
    einc = np.array([i.incomp for i in infos])
    counter = np.array([i.stack_size for i in infos], dtype=float)
    psi = e.field("psi")
    e.close()
    b = (0.0, 2 * O.TWO_PI, -8.0, 8.0)
    og = lambda n: O.grid(n, n, b)
    r = O.euler_run(og(64), og(256), og(128), og(128), dt=DT, final_time=1.0, ic=code, ic_params=[0.04], sim=1,
                    threshold=1e-4, stack_capacity=32, want=("incomp", "map_counter", "psi"))
    err = rel_linf(psi, r["psi"])
    print(ic, "remaps", counter.max(), "max |dE_inc|", np.abs(einc - r["incomp"]).max(), "psi rel Linf", err)
    assert np.array_equal(counter, r["map_counter"])
    assert np.abs(einc - r["incomp"]).max() <= 1e-10
    assert err <= 1e-10


def run_pair(cmm, specs, dt, dts, precision="fp64", grid_of=None, grouping=False, mode=None):
    """the same engines stepped alone (cmm_euler_step) and together in a batch (graphs, or grouped
    launches), with the step sequence dts; returns both sets of step infos and engines"""
    mk = grid_of or (lambda n: (n, 4 * n, 2 * n, 2 * n))

    def make_e(n, ic, amp):
        g = cmm.ensemble_grid
        nc, nf, np_, nv = mk(n)
        e = cmm.Euler(g(nc), g(nf), g(np_), g(nv), dt, initial_condition=ic, ic_params=[amp],
                      simulation_type="cmm_vlasov_poisson_1d", incomp_threshold=1e-4, stack_capacity=32,
                      freq_cut_psi=1e9, precision=precision)
        e.init()
        return e

    alone = [make_e(*sp) for sp in specs]
    ref = [[e.step(d) for d in dts] for e in alone]
    batched = [make_e(*sp) for sp in specs]
    b = cmm.Batch(batched, use_graphs=True, grouping=grouping)
    if mode:
        b.set_mode(mode)
    got = [b.step(d) for d in dts]
    return alone, ref, batched, got, b


def test_grouped_batch_with_set_and_mixed_check_tiles(dev):
    """Engines with one tile edge set for their checks (different block caps) step in grouped
    launches bit for bit as alone; engines whose tile edges differ (as separately autotuned engines
    can) or whose time schemes differ step in the other forms, decided before any engine records a
    step, again bit for bit as alone."""
    import cmm
    dt = 1.0 / 16
    steps = 24
    g = cmm.ensemble_grid
    tiles = [(2, 256), (2, 1024), (2, 64), (2, 8192)]

    def make_t(n, ic, amp, tile, cap, scheme="RK3"):
        e = cmm.Euler(g(n), g(4 * n), g(2 * n), g(2 * n), dt, initial_condition=ic, ic_params=[amp],
                      simulation_type="cmm_vlasov_poisson_1d", incomp_threshold=1e-4, stack_capacity=32,
                      freq_cut_psi=1e9, time_integration=scheme)
        e.init()
        e.set_check_tile(tile, cap)
        return e

    alone = [make_t(*sp, *t) for sp, t in zip(SPECS, tiles)]
    ref = [[e.step(dt) for _ in range(steps)] for e in alone]
    batched = [make_t(*sp, *t) for sp, t in zip(SPECS, tiles)]
    b = cmm.Batch(batched, use_graphs=True, grouping=True)
    got = [b.step(dt) for _ in range(steps)]
    print("batch stats", b.stats())
    assert b.stats()["grouped_steps"] >= steps - 1
    assert_same(alone, ref, batched, got, steps)
    b.close()
    # unlike engines: no grouped step, the same results as alone
    unlike = [[(SPECS[0], 4, "RK3"), (SPECS[1], 2, "RK3"), (SPECS[2], 1, "RK3")],
              [(SPECS[0], 4, "RK3"), (SPECS[1], 4, "RK4")]]
    others = []
    for cfg in unlike:
        mixed = [make_t(*sp, t, 256, scheme=sc) for sp, t, sc in cfg]
        solo = [make_t(*sp, t, 256, scheme=sc) for sp, t, sc in cfg]
        b2 = cmm.Batch(mixed, use_graphs=True, grouping=True)
        got2 = [b2.step(dt) for _ in range(8)]
        ref2 = [[e.step(dt) for _ in range(8)] for e in solo]
        assert b2.stats()["grouped_steps"] == 0
        assert_same(solo, ref2, mixed, got2, 8)
        b2.close()
        others += mixed + solo
    for e in alone + batched + others:
        e.close()


def assert_same(alone, ref, batched, got, steps):
    for k, (ea, eb) in enumerate(zip(alone, batched)):
        einc_a = np.array([r.incomp for r in ref[k]])
        einc_b = np.array([got[s][k].incomp for s in range(steps)])
        assert np.array_equal(einc_a, einc_b), (k, np.abs(einc_a - einc_b).max())
        assert [r.stack_size for r in ref[k]] == [got[s][k].stack_size for s in range(steps)]
        for f in ("chix", "chiy", "psi"):
            assert np.array_equal(ea.field(f), eb.field(f)), (k, f)


@pytest.mark.parametrize("grouping", [False, True])
def test_batch_graph_after_a_shorter_step_and_at_a_non_dyadic_step(dev, grouping):
    """Graph replays take their Lagrange weights from the time history of the step they replay:
    after a shortened step (eager) the nominal steps replay graphs captured under an even history,
    and at dt = 1/20 the weights change in their last bits from step to step.  Grouped steps take
    the coefficients of each step (the shortened one too).  Every instance must still equal its
    single-engine run bit for bit."""
    import cmm
    specs = SPECS[:2]
    for dt, dts in ((DT, [DT] * 14 + [DT / 2] + [DT] * 14), (1.0 / 20, [1.0 / 20] * 40)):
        alone, ref, batched, got, b = run_pair(cmm, specs, dt, dts, grouping=grouping)
        st = b.stats()
        print("dt", dt, "batch stats", st)
        assert (st["grouped_steps"] if grouping else st["graph_launches"]) >= len(dts) - 3
        assert_same(alone, ref, batched, got, len(dts))
        b.close()
        for e in alone + batched:
            e.close()


@pytest.mark.parametrize("grouping", [False, True, "grouped_lanes", "grouped_persistent"])
def test_batch_mixed_grid_sizes_of_the_ensemble_workload(dev, grouping):
    """One instance each of the ensemble workload's grid sizes (maps 256^2, 512^2, 1024^2; fine 4x,
    psi and vort 2x; x in [0, 4 pi), v in [-8, 8); dt 1/20; Landau and two-stream with seeded
    amplitudes), 20 steps: the graph batch and the grouped batch equal the instances run alone bit
    for bit, and the 256^2 instance follows the oracle's run (remap steps, E_inc within 1e-10, psi
    within 1e-10)."""
    import cmm
    import oracle_lib as O
    from gpu_util import rel_linf
    amps = np.random.default_rng(0).uniform(0.005, 0.05, 3)
    specs = [(256, "landau_amp", amps[0]), (512, "two_stream_v2", amps[1]), (1024, "landau_amp", amps[2])]
    dt = 1.0 / 20
    steps = 20
    alone, ref, batched, got, b = run_pair(cmm, specs, dt, [dt] * steps, grouping=bool(grouping),
                                           mode=grouping if isinstance(grouping, str) else None)
    print("batch stats", b.stats(), "remaps", [sum(i.remapped for i in r) for r in ref])
    if grouping:
        assert b.stats()["grouped_steps"] >= steps - 1
    assert_same(alone, ref, batched, got, steps)
    psi = alone[0].field("psi")
    bb = (0.0, 2 * O.TWO_PI, -8.0, 8.0)
    og = lambda n: O.grid(n, n, bb)
    r = O.euler_run(og(256), og(1024), og(512), og(512), dt=dt, final_time=steps * dt, ic=203, ic_params=[amps[0]],
                    sim=1, threshold=1e-4, stack_capacity=32, fcut=1e9, want=("incomp", "map_counter", "psi"))
    einc = np.array([i.incomp for i in ref[0]])
    counter = np.array([i.stack_size for i in ref[0]], dtype=float)
    print("256^2 instance: remaps", counter.max(), "max |dE_inc|", np.abs(einc - r["incomp"]).max(),
          "psi", rel_linf(psi, r["psi"]))
    assert r["steps"] == steps
    assert np.array_equal(counter, r["map_counter"])
    assert np.abs(einc - r["incomp"]).max() <= 1e-10
    assert rel_linf(psi, r["psi"]) <= 1e-10
    b.close()
    for e in alone + batched:
        e.close()


@pytest.mark.parametrize("precision,grouping", [("fp64", False), ("split", False), ("fp64", True)])
def test_engines_alternate_between_single_steps_and_the_batch(dev, precision, grouping):
    """Batched engines also step alone between graph steps: a single-engine step runs the check and
    its E_inc copy on the engine's third stream (forked and joined by events), a graph step captures
    or replays the engine streams' work from the batch's master stream, and neither may leave state
    that the other reads wrongly (the check's events re-recorded outside a capture, the Lagrange
    weights taken by value again, graphs keyed on the step phase that the single steps advanced).
    Every instance must equal its run through single-engine steps alone, bit for bit."""
    import cmm
    specs = SPECS[:2]
    plan = ["b"] * 10 + ["s"] * 5 + ["b"] * 15 + ["s"] * 5 + ["b"] * 5
    alone = [make(cmm, *s, precision) for s in specs]
    ref = [[e.step(DT) for _ in plan] for e in alone]
    mixed = [make(cmm, *s, precision) for s in specs]
    b = cmm.Batch(mixed, use_graphs=True, grouping=grouping)
    got = [b.step(DT) if how == "b" else [e.step(DT) for e in mixed] for how in plan]
    st = b.stats()
    print("batch stats", st, "remaps", [sum(i.remapped for i in r) for r in ref])
    assert (st["grouped_steps"] if grouping else st["graph_launches"]) >= 20
    assert_same(alone, ref, mixed, got, len(plan))
    b.close()
    for e in alone + mixed:
        e.close()


def test_batch_autotune_keeps_the_results(dev):
    """cmm_batch_autotune runs real steps in each launch form (grouped, grouped persistent, graphs,
    eager) and keeps the fastest; the steps it takes and the ones after it must equal the engines'
    own steps bit for bit."""
    import cmm
    specs = SPECS[:3]
    alone = [make(cmm, *s, "fp64") for s in specs]
    batched = [make(cmm, *s, "fp64") for s in specs]
    b = cmm.Batch(batched, use_graphs=True)
    t = b.autotune(DT, 2)
    steps_done = batched[0].state()[0]
    print("autotune", t, "steps taken", steps_done, "stats", b.stats())
    assert t["mode"] in ("eager", "graphs", "grouped", "grouped_persistent") and t["ms_per_step"] > 0
    assert steps_done >= 1 + 2 + 2 + 12 + 2 + 2
    got = [b.step(DT) for _ in range(10)]
    ref = [[e.step(DT) for _ in range(steps_done + 10)][-10:] for e in alone]
    for k, (ea, eb) in enumerate(zip(alone, batched)):
        assert [r.incomp for r in ref[k]] == [got[s][k].incomp for s in range(10)]
        for f in ("chix", "chiy", "psi"):
            assert np.array_equal(ea.field(f), eb.field(f)), (k, f)
    b.close()
    for e in alone + batched:
        e.close()


@pytest.mark.parametrize("ic", ["landau_amp", "two_stream_v2"])
def test_split_vlasov_poisson_against_fp64(dev, ic):
    """A 256^2 instance of the ensemble workload (fine 1024^2, psi and vort 512^2, dt 1/20) to T = 2
    in both forms: the split form (its advection with the separable velocity from the 1D tables,
    as the FP64 form's) remaps on the same steps and its distribution and psi stay within the
    chosen form's 1e-5 relative L2 of the FP64 run's."""
    import cmm
    dt = 1.0 / 20
    g = cmm.ensemble_grid
    out = {}
    for prec in ("fp64", "split"):
        e = cmm.Euler(g(256), g(1024), g(512), g(512), dt, initial_condition=ic, ic_params=[0.05],
                      simulation_type="cmm_vlasov_poisson_1d", incomp_threshold=1e-4, stack_capacity=64,
                      precision=prec)
        e.init()
        infos = [e.step(dt) for _ in range(40)]
        out[prec] = ([i.stack_size for i in infos], e.field("w_coarse"), e.field("psi")[:4 * 512 * 512])
        e.close()
    rl2 = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    ef, ep = rl2(out["split"][1], out["fp64"][1]), rl2(out["split"][2], out["fp64"][2])
    print(f"{ic}: remaps fp64 {out['fp64'][0][-1]} split {out['split'][0][-1]}; f rel L2 {ef:.3e}, psi rel L2 {ep:.3e}")
    assert out["split"][0] == out["fp64"][0]
    assert ef <= 1e-5 and ep <= 1e-5
