#!/usr/bin/env python3
"""CMM steps/s on one MI355X for the 1024^2 map / 2048^2 psi / 4096^2 fine Euler workload.

A step is one pass of the map-advection path: backward RK3 map advection, incompressibility check
on the fine grid, remap when E_inc > 1e-4 (composition of the device-resident stack and new fine
vorticity Hermite), psi solve, and RK3 advection of 4096 particles.  The run starts from a seeded
random-spectrum vorticity (E(k) ~ k^4 exp(-2 (k/6)^2), enstrophy 1), dt = 1/256; W untimed
steps then K timed steps (default W + K = 512 steps = T 2).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cmm-turbulence_amd"))

METRIC = "CMM steps/s @1024² map/4096² fine: plain FP64 vs chosen form at equal L2 error"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

C2 = dict(n_coarse=1024, n_psi=2048, n_fine=4096, n_vort=4096, dt=1.0 / 256, final_time=2.0, threshold=1e-4,
          eps=1e-3, freq_cut_psi=256.0, stack_capacity=256, n_particles=4096, seed=0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic(slot, c, kx_max, n_stack=0):
    """Algorithmic HBM bytes and FP64 flops of one launch (each input read once, each output
    written once; flop counts from the operation counts of the device code)."""
    nc, npsi, nf, nv = c["n_coarse"] ** 2, c["n_psi"] ** 2, c["n_fine"] ** 2, c["n_vort"] ** 2
    L = 3
    if slot == "advect":
        # map in (8 doubles/node) + psi L levels x 4 planes + map out
        return 64 * nc + 32 * L * npsi + 64 * nc, 9.2e3 * nc
    if slot == "incomp":
        return 64 * nc, 320.0 * nf
    if slot == "psi_row_r2c":
        # map + fine omega Hermite read, transposed half-spectrum columns kx <= kx_max written
        return 64 * nc + 32 * nf + 16 * (kx_max + 1) * c["n_vort"], 270.0 * nv + 2.5 * nv * 24
    if slot == "psi_col":
        return 16 * (kx_max + 1) * (c["n_vort"] + 2 * c["n_psi"]), 5.0 * (kx_max + 1) * 36 * c["n_psi"]
    if slot == "psi_row_c2r":
        return 2 * 16 * (kx_max + 1) * c["n_psi"] + 32 * npsi, 2 * 5.0 * npsi * 11
    if slot == "particles":
        return 16 * c["n_particles"], 0.0
    if slot == "remap":
        return (64 * nc * (n_stack + 1) + 32 * nf), 170.0 * nf * (n_stack + 1)
    return 0, 0.0


def uniform_particles(n, seed):
    import cmm
    return cmm.uniform_particles(n, seed)


def cpu_baseline(c, threads):
    """The oracle (C++/OpenMP, FP64) timed on one step's worth of the hot-path operators at the same
    sizes (advect 1024^2, check on 4096^2, psi solve 4096^2 -> 2048^2, 4096 particles) from a
    random-spectrum state and the identity map; returns steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import numpy as np
    import oracle_lib as O
    gc, gp, gf = O.grid(c["n_coarse"]), O.grid(c["n_psi"]), O.grid(c["n_fine"])
    w = O.ic_random_spectrum(c["n_fine"], c["seed"])
    W = O.hermite_from_real(gf, w)
    cx, cy = O.identity_map(gc)
    psi1 = O.psi_from_vort(gc, gf, gf, gp, cx, cy, W, 0, c["freq_cut_psi"])
    psi = np.concatenate([psi1, psi1, psi1])
    pos = uniform_particles(c["n_particles"], c["seed"])
    dt = c["dt"]
    t = np.array([-2 * dt, -dt, 0.0, dt])
    coeff = O.advect_coeffs(t, np.full(4, dt), 0, 3, 30, 1, c["eps"])
    t0 = time.perf_counter()
    nx, ny = O.map_advect(gc, gp, cx, cy, psi, coeff)
    t1 = time.perf_counter()
    O.incomp_error(gc, gf, nx, ny)
    t2 = time.perf_counter()
    O.psi_from_vort(gc, gf, gf, gp, nx, ny, W, 0, c["freq_cut_psi"])
    t3 = time.perf_counter()
    O.particles_advect(gp, psi, 3, [0.375, 0.75, -0.125], dt, 30, pos)
    t4 = time.perf_counter()
    log(f"cpu baseline ({threads} threads): advect {t1 - t0:.2f}s incomp {t2 - t1:.2f}s psi {t3 - t2:.2f}s "
        f"particles {t4 - t3:.3f}s")
    return 1.0 / (t4 - t0), t4 - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=504)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device copy, the reference kernel for counter scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs one engine instance per GPU and has no multi-GPU "
              f"path (requested --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import torch
    import cmm
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    c = dict(C2)
    log("device:", cmm.device_info())
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    g = cmm.grid
    t_setup = time.perf_counter()
    eng = cmm.Euler(g(c["n_coarse"]), g(c["n_fine"]), g(c["n_psi"]), g(c["n_vort"]), c["dt"],
                    incomp_threshold=c["threshold"], map_epsilon=c["eps"], freq_cut_psi=c["freq_cut_psi"],
                    initial_condition="random_spectrum", seed=c["seed"], stack_capacity=c["stack_capacity"])
    eng.add_particles(uniform_particles(c["n_particles"], c["seed"]), scheme=30)
    warm = eng.init()
    t_init = time.perf_counter() - t_setup
    log(f"init + warm-up {t_init:.2f}s, warm-up E_inc {warm}")
    dt = c["dt"]
    t = 0.0
    for i in range(args.warmup):
        info = eng.step(dt)
    torch.cuda.synchronize()
    eng.profile(True)
    remaps = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        info = eng.step(dt)
        remaps += info.remapped
        if (i + 1) % 128 == 0:
            log(f"step {args.warmup + i + 1}: t={info.t:.4f} E_inc={info.incomp:.3e} maps={info.stack_size}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = eng.profile_read()
    steps, t_now, n_stack = eng.state()
    elapsed = t1 - t0
    sps = args.steps / elapsed

    # dominant kernel and its roofline
    kx_max = int(min(c["n_psi"] // 2, c["n_vort"] // 2, c["freq_cut_psi"]))
    per = {}
    for name, (ms, cnt) in prof.items():
        if cnt:
            b, f = algorithmic(name, c, kx_max, n_stack)
            per[name] = dict(ms_total=ms, launches=cnt, ms_avg=ms / cnt, bytes=b, flops=f)
    for name, d in sorted(per.items(), key=lambda kv: -kv[1]["ms_total"]):
        gbs = d["bytes"] / (d["ms_avg"] * 1e-3) / 1e9
        tfs = d["flops"] / (d["ms_avg"] * 1e-3) / 1e12
        log(f"  {name:12s} {d['ms_total']:9.2f} ms total  {d['launches']:5d} launches  {d['ms_avg']:.4f} ms avg"
            f"  {gbs:8.1f} GB/s  {tfs:6.2f} TF/s")
    dom = max((k for k in per if k != "remap"), key=lambda k: per[k]["ms_total"])
    d = per[dom]
    gbs = d["bytes"] / (d["ms_avg"] * 1e-3) / 1e9
    tfs = d["flops"] / (d["ms_avg"] * 1e-3) / 1e12
    if tfs / FP64_PEAK_TFLOPS > gbs / HBM_PEAK_GBS:
        roof = {"bound": "mfma", "achieved": round(tfs, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tfs / FP64_PEAK_TFLOPS, 4)}
    else:
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    roof.update({"kernel": dom, "traffic": None, "ms_avg": round(d["ms_avg"], 5),
                 "algorithmic_bytes": d["bytes"], "algorithmic_flops": d["flops"]})

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        v, secs = cpu_baseline(c, threads)
        cpu = {"value": round(v, 5), "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": f"one step's worth of the hot-path operators at full size (advect 1024^2, check 4096^2, "
                         f"psi 4096^2->2048^2, 4096 particles), {secs:.1f} s of CPU work"}

    out = {
        "metric": METRIC, "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Euler 2D plain FP64: map 1024^2, psi 2048^2, fine/vort 4096^2, random-spectrum "
                               "IC seed 0, dt 1/256, RK3, 4th-order stencil, eps 1e-3, remap 1e-4, "
                               "freq_cut_psi 256, 4096 RK3 particles",
                   "form": "plain FP64", "t_end": round(t_now, 6), "remaps_timed": remaps, "maps": n_stack,
                   "init_s": round(t_init, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_ms_avg": {k: round(v["ms_avg"], 5) for k, v in per.items()},
    }
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
