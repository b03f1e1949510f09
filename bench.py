#!/usr/bin/env python3
"""CMM steps/s on one MI355X for the 1024^2 map / 2048^2 psi / 4096^2 fine Euler workload.

A step is one pass of the map-advection path: backward RK3 map advection, incompressibility check
on the fine grid, remap when E_inc > 1e-4 (composition of the device-resident stack and new fine
vorticity Hermite), psi solve, and RK3 advection of 4096 particles.  The run starts from a seeded
random-spectrum vorticity (E(k) ~ k^4 exp(-2 (k/6)^2), enstrophy 1), dt = 1/256; W untimed
steps then K timed steps (default W + K = 512 steps = T 2).  Prints one JSON line.

The JSON line also carries: the split-precision form at the same step and both forms at twice and
at four times the step run to the same end time (speed-up in simulated time and omega error against
plain FP64 at the nominal step); map-stack composition at depth; the double-shear-layer workload (128^2 / 256^2 /
512^2) stepped singly and from captured graphs; a 100-step window of the 32-instance Vlasov-Poisson
ensemble, graph batch against one by one; and the CPU baseline (the C++/OpenMP oracle's whole steps
of the same workload and one timed remap, with the thread count, nproc and CPU model).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-ensemble] [--copy-calibration]
  python bench.py summarize-profiles KT_DIR FETCH_DIR WRITE_DIR OUT_DIR "kt cmd" "fetch cmd" "write cmd"

roofline.traffic is the dominant kernel's HBM bytes per launch from the committed counter summary
(FETCH_SIZE + WRITE_SIZE, scaled by the 1 GiB copy calibration), null when that file is absent.
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
        # map in (8 doubles/node, interleaved) + the two Lagrange-weighted psi fields and psi level 0
        # (4 planes each) + map out, plane-major and interleaved for the next step
        return 64 * nc + 32 * 3 * npsi + 2 * 64 * nc, 9.2e3 * nc
    if slot == "psi_combine":
        # L psi levels read, C1, C12 and the level-0 copy written (4 planes each)
        return 32 * (L + 3) * npsi, 2 * 2 * L * 4 * npsi
    if slot == "incomp":
        # map read once; separable Hermite sums over 4x4 tiles of check points, about 85 flop each
        return 64 * nc, 85.0 * nf
    if slot == "psi_row_r2c":
        # sampling: map + fine omega Hermite read, omega(X(x)) written; row transform: that field read,
        # transposed half-spectrum columns kx <= kx_max written
        return 64 * nc + 32 * nf + 16 * nv + 16 * (kx_max + 1) * c["n_vort"], 270.0 * nv + 2.5 * nv * 24
    if slot == "psi_row_r2c_fused":
        # the same with the incompressibility check in the sampler's pass (fine grid = vorticity
        # grid): the check's flops at the cell centres, the map read once for both
        b, f = algorithmic("psi_row_r2c", c, kx_max, n_stack)
        return b, f + 85.0 * nf
    if slot in ("psi_row_c2r_pre", "psi_row_c2r_pre_split"):
        # the transform also writes the next step's combined psi fields: the older L - 1 levels
        # read, C1, C12 and C0 written (32-byte FP64 records, 16-byte packed records in the split
        # form), the sums' flops
        b, f = algorithmic("psi_row_c2r", c, kx_max, n_stack)
        rec = 16 if slot.endswith("_split") else 32
        return b + 32 * (L - 1) * npsi + 3 * rec * npsi, f + 2 * 2 * L * 4 * npsi
    if slot == "psi_col":
        return 16 * (kx_max + 1) * (c["n_vort"] + 2 * c["n_psi"]), 5.0 * (kx_max + 1) * 36 * c["n_psi"]
    if slot == "psi_row_c2r":
        return 2 * 16 * (kx_max + 1) * c["n_psi"] + 32 * npsi, 2 * 5.0 * npsi * 11
    if slot == "particles":
        return 16 * c["n_particles"], 0.0
    # split-precision form: 48-byte map records, 16-byte records of the combined psi fields (the
    # value's low part coded in the cross derivative's low byte), 16-byte records of the fine
    # vorticity
    if slot == "advect_split":
        return 48 * nc + 16 * 3 * npsi + 48 * nc, 9.2e3 * nc
    if slot == "psi_combine_split":
        return 32 * L * npsi + 16 * 3 * npsi, 2 * 2 * L * 4 * npsi
    if slot == "incomp_split":
        return 48 * nc, 85.0 * nf
    if slot == "psi_row_r2c_split":
        # the sampled vorticity is written and read back as FP32
        return 48 * nc + 16 * nf + 8 * nv + 16 * (kx_max + 1) * c["n_vort"], 270.0 * nv + 2.5 * nv * 24
    if slot == "remap":
        return (64 * nc * (n_stack + 1) + 32 * nf), 170.0 * nf * (n_stack + 1)
    return 0, 0.0


def uniform_particles(n, seed):
    import cmm
    return cmm.uniform_particles(n, seed)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def stack_maps_host(nc, n_maps):
    """n_maps smooth shears of the identity on an nc^2 grid (derivative planes exact), as the
    composition leg stacks them: (stack_x, stack_y) plane-major, oldest first"""
    import numpy as np
    h = (2 * np.pi) / float(np.float32(nc))
    xs = np.arange(nc) * h
    X, Y = np.meshgrid(xs, xs)
    X, Y = X.ravel(), Y.ravel()
    sx, sy = [], []
    for m in range(n_maps):
        a, ph = 0.01 * (1 + m % 3), 0.37 * m
        sx.append(np.concatenate([X + a * np.sin(Y + ph), np.ones_like(X), a * np.cos(Y + ph), np.zeros_like(X)]))
        sy.append(np.concatenate([Y + a * np.sin(X - ph), a * np.cos(X - ph), np.ones_like(X), np.zeros_like(X)]))
    return np.concatenate(sx), np.concatenate(sy)


def cpu_baseline(c, threads, steps=16, remap_maps=32):
    """The oracle (C++/OpenMP, FP64, the same algorithm) on the host cores, whole steps of the same
    workload: one run through the Lagrange warm-up only and one through the warm-up and `steps`
    steps (advection, check on the 4096^2 grid, remap when due, psi solve, 4096 RK3 particles);
    the difference is the steady stepping.  Then one remap timed on its own: the 4096^2 fine grid
    composed through the current map and remap_maps stacked 1024^2 maps, the initial condition
    sampled there and its Hermite form built (translate_initial_condition_through_map_stack)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import numpy as np
    import oracle_lib as O
    gc, gp, gf = O.grid(c["n_coarse"]), O.grid(c["n_psi"]), O.grid(c["n_fine"])
    t0 = time.perf_counter()
    H = O.hermite_from_real(gf, O.ic_random_spectrum(c["n_fine"], c["seed"]))
    pos = uniform_particles(c["n_particles"], c["seed"])
    kw = dict(dt=c["dt"], final_time=c["final_time"], fcut=c["freq_cut_psi"], eps=c["eps"], threshold=c["threshold"],
              disc=H, g_disc=gf, stack_capacity=c["stack_capacity"], particles=pos, particle_scheme=30,
              want=("incomp",))
    t1 = time.perf_counter()
    O.euler_run(gc, gf, gp, gf, max_steps=-1, **kw)
    t2 = time.perf_counter()
    r = O.euler_run(gc, gf, gp, gf, max_steps=steps, **kw)
    t3 = time.perf_counter()
    assert r["steps"] == steps
    sps = steps / ((t3 - t2) - (t2 - t1))
    # one remap at depth remap_maps + 1
    sx, sy = stack_maps_host(c["n_coarse"], remap_maps)
    cx, cy = O.identity_map(gc)
    t4 = time.perf_counter()
    xy = O.compose_positions(gc, gf, cx, cy, sx, sy, remap_maps)
    w = O.sample_init(gf, xy, 101, disc=H, gdisc=gf, discrete=1)
    O.hermite_from_real(gf, w)
    t5 = time.perf_counter()
    log(f"cpu baseline ({threads} threads): initial field {t1 - t0:.2f}s, warm-up {t2 - t1:.2f}s, warm-up + {steps} "
        f"steps {t3 - t2:.2f}s -> {sps:.3f} steps/s; remap through {remap_maps + 1} maps {t5 - t4:.2f}s")
    return sps, (t3 - t2) - (t2 - t1), t5 - t4, (t1 - t0) + (t3 - t1) + (t5 - t4)


# slot model (algorithmic()) -> kernel-name patterns of the launches the slot times (the row-transform
# slot of the psi solve is the sampling of omega(X(x)) followed by the row transform of the sampled
# field; "_fused": the sampler with the check in its pass; "_pre": the C2R transform that also writes
# the next step's combined psi fields)
SLOT_KERNELS = {"advect": ("k_map_advect_pair<30, true>",), "psi_combine": ("k_psi_combine<3>",),
                "incomp": ("k_incomp_tensor<",),
                "psi_row_r2c": ("k_sample_field<cmm::SampleVortT<true>, double, 4, false>", "k_row_r2c<cmm::SampleReal>"),
                "psi_row_r2c_fused": ("k_sample_field<cmm::SampleVortT<true>, double, 4, true>",
                                      "k_row_r2c<cmm::SampleReal>"),
                "psi_col": ("k_col(",), "psi_row_c2r": ("k_row_c2r4(",), "psi_row_c2r_pre": ("k_row_c2r4_pre<0, 3>",),
                "particles": ("k_particles(",),
                # the split-precision form's own kernels (its column stage is the FP64 one)
                "advect_split": ("k_map_advect32_pair<30, true>",), "psi_combine_split": ("k_psi_combine32<3, true>",),
                "incomp_split": ("k_incomp_tensor32<",),
                "psi_row_r2c_split": ("k_sample_field<cmm::SampleVort32, float, 4, false>", "k_row_r2c<cmm::SampleRealF>"),
                "psi_row_c2r_pre_split": ("k_row_c2r4_pre<1, 3>",)}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cmm_euler_c2_hbm_traffic.json")
FETCH_SCALE_DEFAULT, WRITE_SCALE_DEFAULT = 2.0, 1.0


def _read_csv(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(d, suffix):
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith(suffix):
                return os.path.join(root, f)
    raise FileNotFoundError(f"no *{suffix} under {d}")


def summarize_profiles(kt_dir, fetch_dir, write_dir, out_dir, commands):
    """Turns rocprofv3 output (a --kernel-trace --stats run and two --pmc runs, FETCH_SIZE and
    WRITE_SIZE) into the text summaries kept with the sources: per-kernel durations, and HBM bytes
    per dispatch scaled by what the same counters read for the 1 GiB copy (rocclr copyBuffer)."""
    import collections
    stats = _read_csv(_find(kt_dir, "kernel_stats.csv"))
    lines = ["# rocprofv3 kernel statistics, 1024^2 map / 2048^2 psi / 4096^2 fine Euler workload (MI355X)", "",
             "command: " + commands["kt"], "",
             f"{'kernel':78s} {'calls':>6s} {'total_ms':>10s} {'avg_us':>10s} {'pct':>6s}"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"])):
        lines.append(f"{r['Name'][:78]:78s} {int(r['Calls']):6d} {float(r['TotalDurationNs']) / 1e6:10.3f} "
                     f"{float(r['AverageNs']) / 1e3:10.2f} {float(r['Percentage']):6.2f}")
    # per slot, from the kernel trace: each kernel's launches at its most frequent grid size (the
    # time loop's), their average duration, summed over the slot's kernels (the bench's slot times,
    # taken with events, also hold the gaps between the kernels)
    dur = collections.defaultdict(list)
    for r in _read_csv(_find(kt_dir, "kernel_trace.csv")):
        key = (r["Kernel_Name"], r["Grid_Size_X"], r["Grid_Size_Y"])
        dur[key].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    best = {}
    for (k, _, _), v in dur.items():
        if k not in best or len(v) > len(best[k]):
            best[k] = v
    lines += ["", "slots from the kernel trace (each kernel's launches at its most frequent grid size; average us, "
              "the slot's kernels summed)", f"{'slot':22s} {'kernels':64s} {'launches':>8s} {'avg_us':>10s}"]
    for slot, pats in SLOT_KERNELS.items():
        parts = []
        for pat in pats:
            ks = [k for k in best if pat in k]
            if not ks:
                break
            v = best[ks[0]]
            parts.append((ks[0].split("(")[0].replace("void ", "").replace("cmm::", ""), len(v), sum(v) / len(v)))
        else:
            lines.append(f"{slot:22s} {' + '.join(p[0] for p in parts)[:64]:64s} {min(p[1] for p in parts):8d} "
                         f"{sum(p[2] for p in parts):10.2f}")
    with open(os.path.join(out_dir, "cmm_euler_c2_kernel_stats.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")

    def per_kernel(d, counter):
        # (kernel, grid size) -> values; the time loop's launches are the most frequent grid size
        acc = collections.defaultdict(list)
        for r in _read_csv(_find(d, "counter_collection.csv")):
            if r["Counter_Name"] == counter:
                acc[(r["Kernel_Name"], r["Grid_Size"])].append(float(r["Counter_Value"]))
        best = {}
        for (k, g), vs in acc.items():
            if k not in best or len(vs) > len(best[k]):
                best[k] = vs
        copies = [v for (k, g), vs in acc.items() if "copyBuffer" in k for v in vs]
        return best, copies
    (fetch, cf), (write, cw) = per_kernel(fetch_dir, "FETCH_SIZE"), per_kernel(write_dir, "WRITE_SIZE")
    copy_kib = float(1 << 20)  # the calibration copy moves 1 GiB each way
    fs = copy_kib / max(cf) if cf else FETCH_SCALE_DEFAULT
    ws = copy_kib / max(cw) if cw else WRITE_SCALE_DEFAULT
    c = dict(C2)
    kx_max = int(min(c["n_psi"] // 2, c["n_vort"] // 2, c["freq_cut_psi"]))
    lines = ["# HBM traffic per dispatch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch),",
             "# scaled by what the counters read for a 1 GiB device copy, 1024^2 / 2048^2 / 4096^2 Euler workload",
             "", "fetch command: " + commands["fetch"], "write command: " + commands["write"], "",
             f"copy calibration: FETCH_SIZE {max(cf) if cf else float('nan'):.1f} KiB, WRITE_SIZE "
             f"{max(cw) if cw else float('nan'):.1f} KiB for 1048576 KiB -> scale fetch x{fs:.3f}, write x{ws:.3f}", "",
             f"{'slot':22s} {'kernel':40s} {'n':>4s} {'fetch_MiB':>10s} {'write_MiB':>10s} {'alg_MiB':>10s} "
             f"{'ratio':>6s}"]
    out = {"fetch_scale": fs, "write_scale": ws, "slots": {}}
    for slot, pats in SLOT_KERNELS.items():
        # per-dispatch bytes of each kernel of the slot, summed over its kernels
        fb = wb = 0.0
        names, n = [], None
        for pat in pats:
            ks = [k for k in fetch if pat in k]
            if not ks:
                break
            fv = [v for k in ks for v in fetch[k]]
            wv = [v for k in write if pat in k for v in write[k]]
            fb += 1024.0 * fs * sum(fv) / len(fv)
            wb += 1024.0 * ws * sum(wv) / len(wv) if wv else float("nan")
            names.append(ks[0].split("(")[0].replace("void cmm::", ""))
            n = len(fv) if n is None else min(n, len(fv))
        else:
            alg = algorithmic(slot, c, kx_max)[0]
            out["slots"][slot] = {"fetch_bytes": fb, "write_bytes": wb, "algorithmic_bytes": alg, "dispatches": n,
                                  "kernels": names}
            lines.append(f"{slot:22s} {' + '.join(names)[:40]:40s} {n:4d} {fb / 2**20:10.1f} {wb / 2**20:10.1f} "
                         f"{alg / 2**20:10.1f} {(fb + wb) / alg:6.2f}")
    with open(os.path.join(out_dir, "cmm_euler_c2_hbm_traffic.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(out_dir, "cmm_euler_c2_hbm_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))


def composition(cmm, torch, c, n_maps=32, reps=5):
    """Map-stack composition at full size: the 4096^2 fine grid through the current 1024^2 map and
    n_maps stacked maps (each a smooth shear of the identity, derivatives exact), positions written
    (cmm_compose_positions).  Algorithmic bytes: every map read once (64 B per coarse node) and
    16 B of positions per fine point; the kernel's time from events on its stream."""
    import numpy as np
    nc, nf = c["n_coarse"], c["n_fine"]
    N = nc * nc
    gm, gt = cmm.grid(nc), cmm.grid(nf)
    dev = torch.device("cuda")
    h = (2 * np.pi) / float(np.float32(nc))
    xs = torch.arange(nc, dtype=torch.float64, device=dev) * h
    X = xs[None, :].expand(nc, nc).reshape(-1)
    Y = xs[:, None].expand(nc, nc).reshape(-1)
    stack = torch.empty(n_maps * 8 * N, dtype=torch.float64, device=dev)
    for m in range(n_maps):
        a, ph = 0.01 * (1 + m % 3), 0.37 * m
        sl = stack[m * 8 * N:(m + 1) * 8 * N].view(8, N)
        sl[0] = X + a * torch.sin(Y + ph)
        sl[1] = 1.0
        sl[2] = a * torch.cos(Y + ph)
        sl[3] = 0.0
        sl[4] = Y + a * torch.sin(X - ph)
        sl[5] = a * torch.cos(X - ph)
        sl[6] = 1.0
        sl[7] = 0.0
    cx = torch.empty(4 * N, dtype=torch.float64, device=dev)
    cy = torch.empty_like(cx)
    cmm.map_init_identity(gm, cx, cy)
    xy = torch.empty(2 * nf * nf, dtype=torch.float64, device=dev)
    cmm.apply_map_stack(gt, gm, cx, cy, stack, n_maps, xy)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        cmm.apply_map_stack(gt, gm, cx, cy, stack, n_maps, xy)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    b = (n_maps + 1) * 64 * N + 16 * nf * nf
    gbs = b / (ms * 1e-3) / 1e9
    # the same stack as FP64 node records (64 B per node, the engine's stack): bit-identical positions
    rec = torch.empty(n_maps * 8 * N, dtype=torch.float64, device=dev)
    cmm.stack_pack(gm, stack, n_maps, "fp64", rec)
    xyr = torch.empty_like(xy)
    cmm.apply_map_stack_packed(gt, gm, cx, cy, rec, n_maps, "fp64", xyr)
    ev0.record()
    for _ in range(reps):
        cmm.apply_map_stack_packed(gt, gm, cx, cy, rec, n_maps, "fp64", xyr)
    ev1.record()
    torch.cuda.synchronize()
    ms_r = ev0.elapsed_time(ev1) / reps
    same = bool(torch.equal(xyr, xy))
    # split records (FP64 values, FP32 derivatives, 48 B per stacked node)
    packed = torch.empty(n_maps * 6 * N, dtype=torch.float64, device=dev)
    cmm.stack_pack(gm, stack, n_maps, "split", packed)
    xys = torch.empty_like(xy)
    cmm.apply_map_stack_packed(gt, gm, cx, cy, packed, n_maps, "split", xys)
    ev0.record()
    for _ in range(reps):
        cmm.apply_map_stack_packed(gt, gm, cx, cy, packed, n_maps, "split", xys)
    ev1.record()
    torch.cuda.synchronize()
    ms_s = ev0.elapsed_time(ev1) / reps
    d = (xys - xy).abs()
    dmax = float(torch.minimum(d, 2 * np.pi - d).max())
    b_s = 64 * N + n_maps * 48 * N + 16 * nf * nf
    gbs_s = b_s / (ms_s * 1e-3) / 1e9
    gbs_r = b / (ms_r * 1e-3) / 1e9
    # the reference's dataflow per layer: positions read and written (32 B per fine point) and the
    # map read once; it is not this kernel's traffic (positions stay in registers)
    ref_b = (n_maps + 1) * (32 * nf * nf + 64 * N)
    del stack, xy, packed, xys, rec, xyr
    return {"maps": n_maps + 1, "algorithmic_bytes": b,
            "plane_major": {"kernel": "k_compose_multi<0, 2>", "ms_avg": round(ms, 4), "achieved_GBs": round(gbs, 1),
                            "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                            "point_layers_per_s": round((n_maps + 1) * nf * nf / (ms * 1e-3) / 1e9, 3)},
            "records_fp64": {"kernel": "k_compose_multi<1, 2>", "ms_avg": round(ms_r, 4), "speedup": round(ms / ms_r, 3),
                             "achieved_GBs": round(gbs_r, 1), "frac_hbm": round(gbs_r / HBM_PEAK_GBS, 4),
                             "point_layers_per_s": round((n_maps + 1) * nf * nf / (ms_r * 1e-3) / 1e9, 3),
                             "reference_dataflow_GBs": round(ref_b / (ms_r * 1e-3) / 1e9, 1),
                             "positions_equal_plane_major": same},
            "records_split": {"kernel": "k_compose_multi<2, 2>", "ms_avg": round(ms_s, 4), "speedup": round(ms / ms_s, 3),
                              "algorithmic_bytes": b_s, "achieved_GBs": round(gbs_s, 1),
                              "frac_hbm": round(gbs_s / HBM_PEAK_GBS, 4),
                              "point_layers_per_s": round((n_maps + 1) * nf * nf / (ms_s * 1e-3) / 1e9, 3),
                              "reference_dataflow_GBs": round(ref_b / (ms_s * 1e-3) / 1e9, 1),
                              "max_position_difference": dmax}}


def roofline(c, prof, n_stack, suffix="", fused=False, pre=False):
    """Per-slot rates from the engine's event profile and the roofline object of the dominant
    kernel: algorithmic bytes and flops per launch over its average duration; traffic from the
    committed counter summary.  suffix selects the algorithmic model of a form's slots; fused: the
    check ran in the psi sampler's pass (its own slot then closes at once and is left out); pre:
    the psi solve's last transform wrote the next step's combined fields (the combination's slot
    then closes at once and is left out)."""
    kx_max = int(min(c["n_psi"] // 2, c["n_vort"] // 2, c["freq_cut_psi"]))
    per = {}
    for name, (ms, cnt) in prof.items():
        if (fused and name == "incomp") or (pre and name == "psi_combine"):
            continue
        if cnt:
            model = name + suffix if algorithmic(name + suffix, c, kx_max, n_stack)[0] else name
            if fused and name == "psi_row_r2c":
                model = "psi_row_r2c_fused"
            if pre and name == "psi_row_c2r":
                model = "psi_row_c2r_pre" + suffix
            b, f = algorithmic(model, c, kx_max, n_stack)
            per[name] = dict(ms_total=ms, launches=cnt, ms_avg=ms / cnt, bytes=b, flops=f, model=model)
    for name, d in sorted(per.items(), key=lambda kv: -kv[1]["ms_total"]):
        gbs = d["bytes"] / (d["ms_avg"] * 1e-3) / 1e9
        tfs = d["flops"] / (d["ms_avg"] * 1e-3) / 1e12
        log(f"  {name + suffix:18s} {d['ms_total']:9.2f} ms total  {d['launches']:5d} launches  {d['ms_avg']:.4f} ms avg"
            f"  {gbs:8.1f} GB/s  {tfs:6.2f} TF/s")
    dom = max((k for k in per if k != "remap"), key=lambda k: per[k]["ms_total"])
    d = per[dom]
    gbs = d["bytes"] / (d["ms_avg"] * 1e-3) / 1e9
    tfs = d["flops"] / (d["ms_avg"] * 1e-3) / 1e12
    if tfs / FP64_PEAK_TFLOPS > gbs / HBM_PEAK_GBS:
        roof = {"bound": "fp64", "achieved": round(tfs, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tfs / FP64_PEAK_TFLOPS, 4)}
    else:
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        tr = json.load(open(TRAFFIC_FILE))["slots"].get(d["model"])
        if tr:
            traffic = int(tr["fetch_bytes"] + tr["write_bytes"])
    # the slot's kernels (its events also hold the launches and gaps between them: the check's
    # clearing launch and E_inc copy in the fused sampler's slot)
    roof.update({"kernel": d["model"], "kernels": [k.split("(")[0] for k in SLOT_KERNELS.get(d["model"], ())],
                 "traffic": traffic, "ms_avg": round(d["ms_avg"], 5), "algorithmic_bytes": d["bytes"],
                 "algorithmic_flops": d["flops"]})
    return roof, per


def run_form(cmm, torch, c, precision, warmup, steps, dt=None, t_end=None, profile_every=8):
    """One engine run of the workload: W untimed steps, then K steps timed between two device
    synchronisations; returns the rate, the kernel profile and the final fields (coarse vorticity,
    newest psi level, particles) for the error comparison.  With t_end the timed steps instead run
    until t_end, the last one shortened to land on it.  The kernel profile comes from per-kernel
    events on every profile_every-th timed step (each event adds a packet to the stream, about
    1 % of the step at one step in eight; 0: no events)."""
    import numpy as np
    g = cmm.grid
    dt = c["dt"] if dt is None else dt
    t_setup = time.perf_counter()
    eng = cmm.Euler(g(c["n_coarse"]), g(c["n_fine"]), g(c["n_psi"]), g(c["n_vort"]), dt,
                    incomp_threshold=c["threshold"], map_epsilon=c["eps"], freq_cut_psi=c["freq_cut_psi"],
                    initial_condition="random_spectrum", seed=c["seed"], stack_capacity=c["stack_capacity"],
                    precision=precision)
    eng.add_particles(uniform_particles(c["n_particles"], c["seed"]), scheme=30)
    warm = eng.init()
    fused = eng.check_fusion()
    pre = eng.psi_precombine()
    t_init = time.perf_counter() - t_setup
    log(f"[{precision}] init + warm-up {t_init:.2f}s, warm-up E_inc {warm}")
    einc = []
    for i in range(warmup):
        einc.append(eng.step(dt).incomp)
    torch.cuda.synchronize()
    if profile_every:
        eng.profile(True, every=profile_every)
    remaps = 0
    _, t_start, _ = eng.state()
    t = t_start
    i = 0
    t0 = time.perf_counter()
    while (i < steps) if t_end is None else (t_end - t > 1e-5 * dt):
        info = eng.step(dt if t_end is None else min(dt, t_end - t))
        t = info.t
        remaps += info.remapped
        einc.append(info.incomp)
        i += 1
        if i % 128 == 0:
            log(f"[{precision}] step {warmup + i}: t={info.t:.4f} E_inc={info.incomp:.3e} maps={info.stack_size}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = eng.profile_read() if profile_every else {}
    _, t_now, n_stack = eng.state()
    out = {"sps": i / (t1 - t0), "elapsed": t1 - t0, "steps": i, "sim_time": t_now - t_start, "fused": fused, "pre": pre,
           "sim_per_s": (t_now - t_start) / (t1 - t0), "prof": prof, "maps": n_stack, "t_end": t_now,
           "remaps": remaps, "init_s": t_init, "einc": np.array(einc), "w": eng.field("w_coarse"),
           "psi": eng.field("psi")[:4 * c["n_psi"] ** 2], "part": eng.particles(0)}
    eng.close()
    return out


def run_c1(cmm, torch, cpu_threads=None):
    """The double-shear-layer workload (map 128^2, psi 256^2, fine 512^2, dt 1/64, RK3, remap at
    1e-4) to T = 1: steps/s of single-engine steps (cmm_euler_step, one host wait per step) and of
    the same engine stepped from captured graphs (a batch of one engine); the oracle's steady steps
    on the host as its CPU baseline."""
    g = cmm.grid
    dt, n = 1.0 / 64, 64

    def engine():
        e = cmm.Euler(g(128), g(512), g(256), g(512), dt, freq_cut_psi=128.0, initial_condition="double_shear_layer",
                      incomp_threshold=1e-4, stack_capacity=64)
        e.init()
        return e
    out = {"workload": "double shear layer, map 128^2, psi 256^2, fine 512^2, dt 1/64, RK3, remap 1e-4, T 1"}
    for mode in ("eager", "graph"):
        e = engine()
        b = cmm.Batch([e], use_graphs=True, grouping=False) if mode == "graph" else None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        remaps = 0
        for _ in range(n):
            info = b.step(dt)[0] if b else e.step(dt)
            remaps += info.remapped
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[mode] = {"steps_per_s": round(n / el, 1), "ms_per_step": round(1e3 * el / n, 4), "remaps": remaps,
                     "t_end": info.t}
        if b:
            out[mode]["graph"] = b.stats()
            b.close()
        e.close()
    out["graph_vs_eager"] = round(out["graph"]["steps_per_s"] / out["eager"]["steps_per_s"], 3)
    if cpu_threads:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        kw = dict(dt=dt, final_time=1.0, ic=100, threshold=1e-4, fcut=128.0, stack_capacity=64, want=("incomp",))
        og = O.grid
        t0 = time.perf_counter()
        O.euler_run(og(128), og(512), og(256), og(512), max_steps=-1, **kw)
        t1 = time.perf_counter()
        O.euler_run(og(128), og(512), og(256), og(512), **kw)
        t2 = time.perf_counter()
        out["cpu_oracle_steps_per_s"] = round(n / ((t2 - t1) - (t1 - t0)), 2)
        out["cpu_threads"] = cpu_threads
    log(f"config 1: {out}")
    return out


def ensemble_leg(cmm, torch, n_inst=32, steps=100, warmup=40, precision="fp64"):
    """The 32-instance Vlasov-Poisson ensemble for a bounded window of steps after `warmup`
    untimed steps (the graph batch captures its graphs in the first 12 steps; the autotuned batch
    chooses its launch form by timing them in the first 38): the autotuned batch, the grouped
    batch (every kind of launch of all instances in one launch; FP64 only) and the captured-graph
    and eager batches against the instances stepped one by one (cmm_euler_step each), each mode on
    freshly built, identical engines over the same steps; saturated instances (stack full, remap
    skipped) are counted."""
    dt = 1.0 / 20
    specs = ensemble_specs(n_inst, 0)
    g = cmm.ensemble_grid
    out = {"workload": f"{n_inst} Vlasov-Poisson instances, maps 256^2/512^2/1024^2 (fine 4x, psi and vort 2x), "
                       f"Landau / two-stream, amplitudes U[0.005, 0.05] seed 0, dt 1/20, {precision}, steps "
                       f"{warmup + 1}..{warmup + steps} after {warmup} untimed steps"}
    modes = ["autotuned_batch", "grouped_batch", "graph_batch", "eager_batch", "one_by_one"]
    if precision != "fp64":
        modes.remove("grouped_batch")
    # an untimed pass over the same steps first, so that one-time set-up (transform tables of the
    # remap's sizes, first launches) falls on no mode's window
    for mode in ["warm"] + modes:
        es = []
        for nn, ic, amp in specs:
            e = cmm.Euler(g(nn), g(4 * nn), g(2 * nn), g(2 * nn), dt, initial_condition=ic, ic_params=[amp],
                          simulation_type="cmm_vlasov_poisson_1d", incomp_threshold=1e-4, stack_capacity=64,
                          precision=precision)
            e.init()
            es.append(e)
        b = (cmm.Batch(es, use_graphs=mode not in ("eager_batch", "warm"), grouping=mode == "grouped_batch")
             if mode != "one_by_one" else None)
        tuned = b.autotune(dt, 5) if mode == "autotuned_batch" else None
        while es[0].state()[0] < warmup:
            if b:
                b.step(dt)
            else:
                for e in es:
                    e.step(dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        remaps = saturated = 0
        for _ in range(steps):
            infos = b.step(dt) if b else [e.step(dt) for e in es]
            remaps += sum(x.remapped for x in infos)
            saturated += sum(x.saturated for x in infos)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[mode] = {"ensemble_steps_per_s": round(steps / el, 2), "instance_steps_per_s": round(steps * n_inst / el, 1),
                     "remaps": remaps, "saturated_steps": saturated}
        if tuned:
            out[mode]["autotune"] = tuned
        if b:
            out[mode]["graph"] = b.stats()
            b.close()
        for e in es:
            e.close()
        torch.cuda.synchronize()
    out.pop("warm")
    for mode in modes[:-1]:
        out[f"{mode}_vs_one_by_one"] = round(out[mode]["ensemble_steps_per_s"] /
                                             out["one_by_one"]["ensemble_steps_per_s"], 3)
    log(f"ensemble: {out}")
    return out


def ensemble_specs(n_inst, seed):
    """The mixed ensemble: map grids cycling 256^2, 512^2, 1024^2 (fine 4x, psi and vort 2x),
    Landau damping and two-stream distributions alternating, amplitudes uniform in [0.005, 0.05]
    from a seeded generator."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = (256, 512, 1024)
    return [(sizes[i % 3], "landau_amp" if i % 2 == 0 else "two_stream_v2", float(rng.uniform(0.005, 0.05)))
            for i in range(n_inst)]


def run_ensemble(args):
    """Batched 1D-1V Vlasov-Poisson ensemble to T, in the FP64 and split forms: the batch's launch
    form (grouped launches, persistent grouped launches, captured graphs, eager queues) chosen by
    the autotuner in the run's first steps, saturated instances counted; the rate of every form
    against stepping the instances one by one over a bounded window of steps on freshly built
    engines (ensemble_leg), and the split form's errors at T."""
    import numpy as np
    import torch
    import cmm
    dt, t_end = 1.0 / 20, args.t_end
    steps = int(round(t_end / dt))
    specs = ensemble_specs(args.instances, 0)
    g = cmm.ensemble_grid

    def build(prec):
        es = []
        for n, ic, amp in specs:
            e = cmm.Euler(g(n), g(4 * n), g(2 * n), g(2 * n), dt, initial_condition=ic, ic_params=[amp],
                          simulation_type="cmm_vlasov_poisson_1d", incomp_threshold=1e-4, stack_capacity=64,
                          precision=prec)
            e.init()
            es.append(e)
        return es

    out = {"workload": f"{args.instances} Vlasov-Poisson instances, maps 256^2/512^2/1024^2 (fine 4x, psi and vort 2x), "
                       f"x in [0, 4 pi), v in [-8, 8), Landau / two-stream alternating, amplitudes U[0.005, 0.05] "
                       f"seed 0, dt 1/20, T {t_end}", "steps": steps, "forms": {}}
    fields = {}
    for prec in ("fp64", "split"):
        es = build(prec)
        tunes = [cmm.autotune(e) for e in es]
        b = cmm.Batch(es, use_graphs=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # the launch form chosen by timing in the run's first steps (part of the timed run)
        form = b.autotune(dt, 5)
        remaps = saturated = 0
        for i in range(es[0].state()[0], steps):
            infos = b.step(dt)
            remaps += sum(x.remapped for x in infos)
            saturated += sum(x.saturated for x in infos)
            if (i + 1) % 200 == 0:
                log(f"[ensemble {prec}] step {i + 1}/{steps}, remaps so far {remaps}")
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        st = b.stats()
        b.close()
        fields[prec] = [(e.field("w_coarse"), e.field("psi")[:4 * (2 * n) ** 2]) for e, (n, _, _) in zip(es, specs)]
        out["forms"][prec] = {"ensemble_steps_per_s": round(steps / el, 3), "instance_steps_per_s": round(
            steps * len(es) / el, 1), "remaps_after_autotune": remaps, "saturated_steps": saturated, "graph": st,
            "autotune": form, "check_tiles": sorted({(t["tile"], t["max_blocks"]) for t in tunes})}
        log(f"[ensemble {prec}] {steps} steps in {el:.2f}s: {out['forms'][prec]}")
        for e in es:
            e.close()
        # every launch form and the instances one by one, each on fresh engines over one window
        out["forms"][prec]["windows"] = ensemble_leg(cmm, torch, n_inst=args.instances, precision=prec)
    rl2 = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    errs = [(rl2(fs[0], f6[0]), rl2(fs[1], f6[1])) for fs, f6 in zip(fields["split"], fields["fp64"])]
    out["split_vs_fp64_at_t_end"] = {"f_coarse_rel_l2_max": max(e[0] for e in errs),
                                     "psi_rel_l2_max": max(e[1] for e in errs)}
    print(json.dumps(out), flush=True)


def run_learned(args):
    """The learned psi extrapolation (map args.learned_map^2, psi args.learned_psi^2, fine and
    vorticity args.learned_fine^2; by default the double-shear-layer grids 128^2 / 256^2 / 512^2)
    with seeded random-spectrum vorticity: psi trajectories recorded every args.learned_dt for seeds
    0..n_train-1 (training) and 64.. (held out), AdamW training through the HIP kernels, the
    extrapolation errors with and without the network on the held-out seeds (extrapolations across
    two recording steps), and the held-out runs at dt = 2 learned_dt with and without the
    correction (fallback judge at 1e-6) against runs at learned_dt / 4: relative L2 error of the
    coarse vorticity at T = args.learned_run_t."""
    import numpy as np
    import torch
    import cmm
    sys.path.insert(0, os.path.join(ROOT, "cmm-turbulence_amd"))
    import learned
    dt_rec, n_psi, n_map, n_fine = args.learned_dt, args.learned_psi, args.learned_map, args.learned_fine
    fcut = args.learned_fcut if args.learned_fcut else n_psi / 2
    steps = int(round(args.learned_t / dt_rec))
    t0 = time.perf_counter()
    traj = lambda s: learned.trajectory(s, n_map=n_map, n_psi=n_psi, n_fine=n_fine, dt=dt_rec, steps=steps, fcut=fcut)
    trajs = [traj(s) for s in range(args.train_seeds)]
    held = [traj(s) for s in range(64, 64 + args.held_seeds)]
    t_data = time.perf_counter() - t0
    net = learned.Net(seed=0)
    learned.fit_scales(net, trajs)
    t0 = time.perf_counter()
    hist = learned.train(net, trajs, iters=args.train_iters, batch=args.train_batch, lr=1e-3, log=log)
    torch.cuda.synchronize()
    t_train = (time.perf_counter() - t0) / args.train_iters
    del trajs
    p1, l1, p12, l12 = learned.evaluate(net, held)
    del held
    g = cmm.grid
    rl2 = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    runs = []
    dt = 2 * dt_rec
    n = int(round(args.learned_run_t / dt))
    for seed in range(64, 64 + args.held_seeds):
        def engine(d):
            e = cmm.Euler(g(n_map), g(n_fine), g(n_psi), g(n_fine), d, initial_condition="random_spectrum", seed=seed,
                          freq_cut_psi=fcut, stack_capacity=16)
            e.init()
            return e
        ref = engine(dt / 8)
        for _ in range(8 * n):
            ref.step(dt / 8)
        w_ref = ref.field("w_coarse")
        ref.close()
        plain = engine(dt)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(n):
            plain.step(dt)
        torch.cuda.synchronize()
        tp = time.perf_counter() - ta
        w_plain = plain.field("w_coarse")
        plain.close()
        e = engine(dt)
        cor = learned.Corrector(net, e, g(n_psi), tol=1e-6)
        ta = time.perf_counter()
        for _ in range(n):
            cor.step(dt)
        torch.cuda.synchronize()
        tl = time.perf_counter() - ta
        w_l = e.field("w_coarse")
        e.close()
        na = len(cor.dev_learned)
        runs.append({"seed": seed, "omega_rel_l2_plain": rl2(w_plain, w_ref), "omega_rel_l2_learned": rl2(w_l, w_ref),
                     "fallback_step": cor.fallback_step, "active_steps": na,
                     "mean_dev_plain_while_active": float(np.mean(cor.dev_plain[:na])) if na else None,
                     "mean_dev_learned": float(np.mean(cor.dev_learned)) if na else None,
                     "ms_per_step_plain": 1e3 * tp / n, "ms_per_step_learned": 1e3 * tl / n})
        log(f"[learned] seed {seed}: {runs[-1]}")
    out = {"workload": f"learned psi extrapolation: map {n_map}^2, psi {n_psi}^2, fine {n_fine}^2, psi cut-off {fcut}, "
                       f"random-spectrum seeds; training seeds 0..{args.train_seeds - 1}, held out "
                       f"64..{63 + args.held_seeds}; psi recorded every {dt_rec} to T {args.learned_t}, extrapolated "
                       f"across {dt}; runs at dt {dt} against {dt / 8}, T {args.learned_run_t}",
           "network": "4 layers, 3x3, 32 channels, BF16 weights/activations, FP32 MFMA accumulation, AdamW with "
                      "warm-up, cosine decay and clipping",
           "data_s": round(t_data, 2), "train_ms_per_iteration": round(1e3 * t_train, 3), "train_loss": hist,
           "extrapolation_rel_l2_held_out": {"to_t_plus_dt": {"plain": p1, "learned": l1, "gain": p1 / l1},
                                             "to_t_plus_half_dt": {"plain": p12, "learned": l12, "gain": p12 / l12}},
           "runs": runs,
           "omega_rel_l2_mean": {"plain": float(np.mean([r["omega_rel_l2_plain"] for r in runs])),
                                 "learned": float(np.mean([r["omega_rel_l2_learned"] for r in runs]))}}
    print(json.dumps(out), flush=True)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "summarize-profiles":
        kt, fe, wr, out = sys.argv[2:6]
        cmds = dict(zip(("kt", "fetch", "write"), sys.argv[6:9]))
        summarize_profiles(kt, fe, wr, out, cmds)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=504)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision form")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the short ensemble leg")
    ap.add_argument("--no-config1", action="store_true", help="skip the double-shear-layer leg")
    ap.add_argument("--ensemble", action="store_true",
                    help="run the batched Vlasov-Poisson ensemble instead (its own JSON line)")
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--learned", action="store_true",
                    help="run the learned psi-extrapolation protocol instead (its own JSON line)")
    ap.add_argument("--train-seeds", type=int, default=64)
    ap.add_argument("--held-seeds", type=int, default=8)
    ap.add_argument("--train-iters", type=int, default=3000)
    ap.add_argument("--train-batch", type=int, default=8)
    ap.add_argument("--learned-dt", type=float, default=1.0 / 32, help="psi recording step of the learned leg")
    ap.add_argument("--learned-psi", type=int, default=256, help="psi grid of the learned leg")
    ap.add_argument("--learned-t", type=float, default=2.0, help="length of the recorded trajectories")
    ap.add_argument("--learned-map", type=int, default=128, help="map grid of the learned leg")
    ap.add_argument("--learned-fine", type=int, default=512, help="fine (and vorticity) grid of the learned leg")
    ap.add_argument("--learned-fcut", type=float, default=None, help="psi cut-off of the learned leg (psi grid / 2)")
    ap.add_argument("--learned-run-t", type=float, default=1.0, help="length of the learned leg's in-loop runs")
    ap.add_argument("--t-end", type=float, default=40.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device copy, the reference kernel for counter scaling")
    args = ap.parse_args()
    if args.ensemble:
        run_ensemble(args)
        return
    if args.learned:
        run_learned(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs one engine instance per GPU and has no multi-GPU "
              f"path (requested --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
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
    plain = run_form(cmm, torch, c, "fp64", args.warmup, args.steps)
    sps, elapsed, prof, n_stack, t_now, remaps, t_init = (plain[k] for k in (
        "sps", "elapsed", "prof", "maps", "t_end", "remaps", "init_s"))

    roof, per = roofline(c, prof, n_stack, fused=plain["fused"], pre=plain["pre"])

    comp = composition(cmm, torch, c)
    log(f"composition: {comp}")

    # the chosen form: split precision at the same step, with its errors against the FP64 run
    chosen = None
    if not args.no_split:
        sp = run_form(cmm, torch, c, "split", args.warmup, args.steps)
        rl2 = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        rli = lambda a, b: float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        dpart = np.abs(sp["part"] - plain["part"])
        dpart = np.minimum(dpart, 2 * np.pi - dpart)
        chosen = {"form": "split precision (map = identity + FP32 displacement, FP32 combined psi records "
                          "keeping the values' low parts, FP32 fine omega records, FP64 arithmetic, FP64 psi levels / FFT / "
                          "Poisson divide)",
                  "value": round(sp["sps"], 3), "ms_per_step": round(1e3 * sp["elapsed"] / args.steps, 4),
                  "speedup": round(sp["sps"] / sps, 3),
                  "errors_vs_fp64_at_t_end": {
                      "omega_coarse_rel_l2": rl2(sp["w"], plain["w"]), "omega_coarse_rel_linf": rli(sp["w"], plain["w"]),
                      "psi_rel_l2": rl2(sp["psi"], plain["psi"]), "psi_rel_linf": rli(sp["psi"], plain["psi"]),
                      "particles_max_abs": float(dpart.max()),
                      "e_inc_max_abs_diff": float(np.max(np.abs(sp["einc"] - plain["einc"])))},
                  "remaps_timed": sp["remaps"], "maps": sp["maps"],
                  "roofline": roofline(c, sp["prof"], sp["maps"], "_split", fused=sp["fused"], pre=sp["pre"])[0],
                  "kernels_ms_avg": {k: round(ms / cnt, 5) for k, (ms, cnt) in sp["prof"].items() if cnt}}
        log(f"split form: {sp['sps']:.1f} steps/s ({chosen['speedup']:.2f}x), errors {chosen['errors_vs_fp64_at_t_end']}")
        # both forms at twice the step (plain Lagrange extrapolation of psi) to the same end time as
        # the nominal run; the error of each against the FP64 run at the nominal step, and the rate
        # in simulated time against it
        big = {}
        for prec in ("fp64", "split"):
            r = run_form(cmm, torch, c, prec, args.warmup // 2, 0, dt=2 * c["dt"], t_end=t_now, profile_every=0)
            assert abs(r["t_end"] - t_now) <= 1e-12, (r["t_end"], t_now)
            big[prec] = {"dt": 2 * c["dt"], "steps_per_s": round(r["sps"], 3), "t_end": r["t_end"],
                         "speedup_in_simulated_time": round(r["sim_per_s"] / plain["sim_per_s"], 3),
                         "omega_coarse_rel_l2": rl2(r["w"], plain["w"]), "psi_rel_l2": rl2(r["psi"], plain["psi"])}
            log(f"[{prec} at 2 dt] {big[prec]}")
        chosen["at_twice_the_step"] = big
        # and at four times the step: how far the step can grow within the 1e-5 bound
        big4 = {}
        for prec in ("fp64", "split"):
            r = run_form(cmm, torch, c, prec, max(1, args.warmup // 4), 0, dt=4 * c["dt"], t_end=t_now, profile_every=0)
            assert abs(r["t_end"] - t_now) <= 1e-12, (r["t_end"], t_now)
            big4[prec] = {"dt": 4 * c["dt"], "steps_per_s": round(r["sps"], 3), "t_end": r["t_end"],
                          "speedup_in_simulated_time": round(r["sim_per_s"] / plain["sim_per_s"], 3),
                          "omega_coarse_rel_l2": rl2(r["w"], plain["w"]), "psi_rel_l2": rl2(r["psi"], plain["psi"])}
            log(f"[{prec} at 4 dt] {big4[prec]}")
        chosen["at_four_times_the_step"] = big4
        # the three readings at the common end time: each against plain FP64 at the nominal step
        chosen["readings"] = {
            "fp64_at_2dt_vs_fp64_at_dt": {"speedup_in_simulated_time": big["fp64"]["speedup_in_simulated_time"],
                                          "omega_coarse_rel_l2": big["fp64"]["omega_coarse_rel_l2"]},
            "split_at_2dt_vs_fp64_at_dt": {"speedup_in_simulated_time": big["split"]["speedup_in_simulated_time"],
                                           "omega_coarse_rel_l2": big["split"]["omega_coarse_rel_l2"]},
            "split_vs_fp64_at_the_same_dt": {"speedup": round(sp["sps"] / sps, 3),
                                             "omega_coarse_rel_l2": rl2(sp["w"], plain["w"])},
            "split_at_2dt_vs_fp64_at_2dt": {"speedup": round(big["split"]["steps_per_s"] / big["fp64"]["steps_per_s"], 3)},
            "fp64_at_4dt_vs_fp64_at_dt": {"speedup_in_simulated_time": big4["fp64"]["speedup_in_simulated_time"],
                                          "omega_coarse_rel_l2": big4["fp64"]["omega_coarse_rel_l2"]},
            "split_at_4dt_vs_fp64_at_dt": {"speedup_in_simulated_time": big4["split"]["speedup_in_simulated_time"],
                                           "omega_coarse_rel_l2": big4["split"]["omega_coarse_rel_l2"]}}

    cpu = None
    if not args.no_cpu_baseline:
        # the host's share of threads for one GPU (OpenMP threads as set for the job, at most 16)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        threads = min(threads, 16)
        v, secs, remap_s, total = cpu_baseline(c, threads)
        cpu = {"value": round(v, 4), "unit": "steps/s", "cores": threads, "kind": "port",
               "nproc": os.cpu_count(), "cpu_model": cpu_model(), "remap_s": round(remap_s, 3),
               "sample": f"16 whole steps of the same workload (advection, check on 4096^2, psi solve, 4096 RK3 "
                         f"particles; no remap falls due) after the Lagrange warm-up, by the C++/OpenMP oracle on "
                         f"{threads} threads ({secs:.1f} s), and one remap through 33 maps on the 4096^2 grid "
                         f"timed alone; {total:.1f} s of CPU work in all"}

    c1 = None if args.no_config1 else run_c1(cmm, torch, cpu["cores"] if cpu else None)
    ens = None if args.no_ensemble else ensemble_leg(cmm, torch)

    out = {
        "metric": METRIC, "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Euler 2D plain FP64: map 1024^2, psi 2048^2, fine/vort 4096^2, random-spectrum "
                               "IC seed 0, dt 1/256, RK3, 4th-order stencil, eps 1e-3, remap 1e-4, "
                               "freq_cut_psi 256, 4096 RK3 particles",
                   "form": "plain FP64", "t_end": round(t_now, 6), "remaps_timed": remaps, "maps": n_stack,
                   "check_in_psi_sampler": plain["fused"], "psi_precombined": plain["pre"],
                   "init_s": round(t_init, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_ms_avg": {k: round(v["ms_avg"], 5) for k, v in per.items()},
        "chosen_form": chosen,
        "composition": comp,
        "config1": c1,
        "ensemble": ens,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
