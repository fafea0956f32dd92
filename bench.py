#!/usr/bin/env python
"""Throughput of the GenPK power-spectrum path on one MI355X.

Metric: Gparticles/s of one forward pass -- CIC deposit + 3-D r2c FFT +
CIC-deconvolved binned P(k) -- at a 1024^3 grid, on the clustered
workload: 1024^3 particles in a 1000 Mpc/h box, half uniform and half in
10^5 Gaussian clumps (power-law member counts, radii 0.5-4 cells), seeded,
generated on the GPU (synthetic data).  Positions are resident in HBM
before the timed region; one step = clear + deposit + fused r2c/bin.

One JSON line is printed with the value, the roofline of the dominant
kernel (HIP events recorded on the plan's stream around each stage during
the timed steps), the whole-pipeline HBM fraction, the per-bin deviation of
the timed precision from the FP64 path on the same particles, and the CPU
oracle (the reference algorithm restated in C/OpenMP) timed on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="f32", choices=["f32", "f64"])
    ap.add_argument("--n-log2", type=int, default=30, help="log2 of the particle count (1024^3 = 2^30)")
    ap.add_argument("--dims", type=int, default=1024)
    ap.add_argument("--box", type=float, default=1000.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 comparison pass")
    ap.add_argument("--cpu-sample-log2", type=int, default=24)
    ap.add_argument("--no-backward", action="store_true", help="skip the forward+backward measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (64 uneven realisations) measurement")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--backward-steps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run plain copy kernels of known size (references for the HBM byte counters)")
    return ap.parse_args()


def stage_bytes(name, n, c, p_pad, geo, rec):
    """Algorithmic HBM bytes of one launch of a stage: what the stage must
    read and write once for its function.  geo: deposit tiling of the plan
    (tiles K, halo'd cells per tile, primary cells per tile)."""
    K, cells, prim = geo
    digit_bytes = sort_digit_bytes(K)
    if name == "sort_upsweep2" and digit_bytes:
        return n                                         # read the low digits (one byte each)
    if name.startswith("sort_upsweep"):
        return rec * n                                   # read the records
    if name == "sort_downsweep1" and digit_bytes:
        return 2 * rec * n + n                           # read + write the records, write the low digits
    if name.startswith("sort_downsweep"):
        return 2 * rec * n                               # read + write the records
    if name == "deposit_tiles":
        return rec * n + c * K * cells                   # read records, write the halo'd tiles
    if name in ("fold_y", "fold_x"):
        rows = 17 if name == "fold_y" else 33            # halo rows of 32 values + a z-halo entry each
        return 4 * c * K * rows * 33                     # read + zero the sources, read + write the faces
    if name == "fold_fft_z":
        return c * K * prim + c * p_pad                  # read the tile cells (z halos folded on load), write the grid
    if name in ("fft_z", "fft_y", "fft_x", "ifft_z", "ifft_y", "ifft_x", "adjoint_x"):
        return 2 * c * p_pad
    if name == "xpass_bin":
        return c * p_pad
    if name == "adjoint_gather":
        return (rec + 4) * n + c * K * cells + 12 * n    # records with index, G tiles, gradient write
    return 0


def sort_digit_bytes(K):
    """Whether the tile sort of K tiles runs two passes with digits of at
    most 8 bits, its first pass then writing each record's low digit as a
    byte (the key of ceil(log2 K) bits splits into a high digit and a low
    digit of half the bits, rounded up)."""
    kbits = max(int(K - 1).bit_length(), 0)
    if kbits <= 9:
        return False
    lo = (kbits + 1) // 2
    return lo <= 8 and kbits - lo <= 8


def tile_geometry(d):
    """Deposit tiling of a dims^3 plan: edges min(16|32|32, largest power of two <= d)."""
    def p2(cap):
        t = 1
        while t * 2 <= min(cap, d):
            t *= 2
        return t
    tx, ty, tz = p2(16), p2(32), p2(32)
    K = (-(-d // tx)) * (-(-d // ty)) * (-(-d // tz))
    return K, (tx + 1) * (ty + 1) * (tz + 1), tx * ty * tz


def fwd_bwd(genpk_amd, torch, pos, box, d, nbins, steps, prec, warm=1, reps=1):
    """Forward + backward of the band-power chi^2 loss (target: the forward
    P times 1.05) with the recompute policy and a masses tensor (all ones)
    that receives gradients, against the forward on the same inputs."""
    from genpk_amd.autograd import chi2_loss, power_spectrum
    n = pos.shape[0]
    plan = genpk_amd.Plan(d, prec)
    masses = torch.ones(n, dtype=torch.float32, device=pos.device)
    plan.clear()
    plan.deposit(pos, box, masses)
    P0, cnt, _ = plan.power(nbins, mass=float(n))
    target = P0 * 1.05

    def fwd():
        plan.clear()
        plan.deposit(pos, box, masses)
        return plan.power(nbins, mass=float(n))[0]

    def fb(policy):
        tp = pos.detach().requires_grad_(True)
        tm = masses.detach().requires_grad_(True)
        P = power_spectrum(tp, box, d, nbins, masses=tm, total_mass=float(n), plan=plan, recompute=policy)
        return torch.autograd.grad(chi2_loss(P, target, cnt), (tp, tm))

    def timed(f, *a):
        # the median over `reps` groups of `steps` calls (a one-off stall of
        # the device or the host does not move it)
        for _ in range(warm):
            f(*a)
        torch.cuda.synchronize()
        groups = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(steps):
                f(*a)
            torch.cuda.synchronize()
            groups.append((time.perf_counter() - t0) * 1e3 / steps)
        return sorted(groups)[len(groups) // 2]

    t_f = timed(fwd)
    out = {"loss": "sum_b (P_b - T_b)^2 / sigma_b^2, sigma_b = T_b sqrt(2/count_b), T = 1.05 P",
           "grads": "positions and masses", "steps": steps, "reps": reps, "ms_fwd_same_inputs": round(t_f, 3)}
    for policy in ("always", "auto"):
        t = timed(fb, policy)
        out[f"ms_fwd_bwd_{policy}"] = round(t, 3)
        out[f"ratio_{policy}"] = round(t / t_f, 3)
    out["policy_of_ratio"] = "always (deposit + transform recomputed in the backward)"
    out["ratio"] = out["ratio_always"]
    del plan
    torch.cuda.empty_cache()
    return out


def batched(genpk_amd, torch, prec, count, reps=5):
    """Uneven batch: `count` realisations, realisation r with seed 100 + r,
    n_side log-uniform in [64, 256] (n_side^3 particles), dims uniform over
    {128, 256, 384, 512}, 0.3-cell Gaussian lattice in a 100 Mpc/h box,
    nrbins = dims.  Plain: one plan per dims, each realisation run alone
    (clear + deposit + fused power, then synchronize), times summed.
    Batched: the autotuned hipGraph over all realisations, one launch."""
    from genpk_amd import batch as gb
    from genpk_amd import synth
    box = 100.0
    jobs = []
    for ns, d, sd in synth.uneven_realisations(count):
        pos = synth.gaussian_lattice_torch(ns, box, d, 0.3, seed=sd)
        jobs.append(dict(pos=pos, boxsize=box, dims=d, nrbins=d, total_mass=float(pos.shape[0]), precision=prec))
    n_tot = sum(j["pos"].shape[0] for j in jobs)
    plans = {d: genpk_amd.Plan(d, prec) for d in sorted({j["dims"] for j in jobs})}

    def plain():
        outs = []
        for j in jobs:
            p = plans[j["dims"]]
            p.clear()
            p.deposit(j["pos"], box)
            outs.append(p.power(j["nrbins"], mass=j["total_mass"])[0])
            torch.cuda.synchronize()
        return outs

    plain()
    t_plain = []
    for _ in range(3):
        t0 = time.perf_counter()
        ref = plain()
        t_plain.append(time.perf_counter() - t0)
    t_plain = sorted(t_plain)[1]
    del plans
    torch.cuda.empty_cache()
    best, rep = gb.tune(jobs, use_cache=False)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        best.launch()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_b = sorted(ts)[reps // 2]
    dev = max(float(((best.results(i)[0] - ref[i]).abs() / ref[i].abs().clamp(min=1e-300)).max())
              for i in range(len(jobs)))
    best.close()
    torch.cuda.empty_cache()
    return {"realisations": count, "particles": n_tot, "dims": sorted({j["dims"] for j in jobs}),
            "ms_plain_sum": round(t_plain * 1e3, 3), "ms_batched": round(t_b * 1e3, 3),
            "speedup": round(t_plain / t_b, 3), "Gparticles_per_s_batched": round(n_tot / t_b / 1e9, 3),
            "autotune": rep, "max_rel_dev_from_plain": dev}


def cpu_baseline(pos_dev, box, dims, nbins, n_total, sample_log2):
    import numpy as np
    import oracle
    n = pos_dev.shape[0]
    stride = max(1, n >> sample_log2)
    sample = pos_dev[::stride].cpu().numpy()
    ns = sample.shape[0]
    grid = oracle.new_field(dims)
    t0 = time.perf_counter()
    oracle.fieldize(box, dims, grid, sample, None, 1.0, 1)
    t_dep = time.perf_counter() - t0
    t0 = time.perf_counter()
    F = oracle.fft_r2c_3d(grid)
    t_fft = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.powerspectrum(dims, F, F, nbins, float(ns), float(ns))
    t_bin = time.perf_counter() - t0
    t_total = t_dep * (n_total / ns) + t_fft + t_bin
    return {
        "value": n_total / t_total / 1e9,
        "unit": "Gparticles/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": (f"every {stride}th of the {n_total} particles ({ns}) deposited on the {dims}^3 FP64 grid "
                   f"({t_dep:.2f} s, scaled x{n_total / ns:.0f}); FP64 r2c ({t_fft:.2f} s) and binning "
                   f"({t_bin:.2f} s) at the full {dims}^3"),
        "seconds_measured": round(t_dep + t_fft + t_bin, 2),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU slab decomposition "
              "is designed but not built.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import genpk_amd
    from genpk_amd import synth

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    n = 1 << args.n_log2
    d = args.dims
    nbins = d
    at_256 = None
    if not args.no_backward:
        # the fwd+bwd ratio at its own target size, measured before anything
        # large is allocated: 2^24 particles (the same clustered recipe, 64x
        # fewer clumps) on a 256^3 grid in a 250 Mpc/h box; steps of a few
        # ms, so the median of 5 groups of 10 after 3 warm-up calls
        n3, d3, box3 = 1 << 24, 256, 250.0
        pos3 = synth.clumped_torch(n3, box3, d3, n_clumps=100000 >> 6, seed=5, device=dev)
        at_256 = fwd_bwd(genpk_amd, torch, pos3, box3, d3, d3, 10, args.prec, warm=3, reps=5)
        at_256["workload"] = f"clustered {n3} particles, d={d3}, box {box3}, nrbins {d3}"
        del pos3
        torch.cuda.empty_cache()
    t_gen = time.perf_counter()
    pos = synth.clumped_torch(n, args.box, d, n_clumps=100000, seed=4, device=dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen

    plan = genpk_amd.Plan(d, args.prec)

    def step():
        plan.clear()
        plan.deposit(pos, args.box, mass=1.0)
        return plan.power(nbins, mass=float(n))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.timing_read()  # drop warm-up marks
    plan.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pw, cnt, keffs = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan.timing(False)
    stages = plan.timing_read()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_step * 1e-3) / 1e9

    c = 4 if args.prec == "f32" else 8
    p_pad = d * d * 2 * (d // 2 + 1)
    geo = tile_geometry(d)
    rec = 12
    per_stage = {}
    for name, (tot, launches) in stages.items():
        avg = tot / max(launches, 1)
        b = stage_bytes(name, n, c, p_pad, geo, rec)
        per_stage[name] = {"ms": round(avg, 4), "launches_per_step": launches / args.steps,
                           "alg_bytes": b, "GBps": round(b / (avg * 1e-3) / 1e9, 1) if avg > 0 and b else None}
    dom = max(per_stage, key=lambda k: per_stage[k]["ms"] * per_stage[k]["launches_per_step"])
    ds = per_stage[dom]
    achieved = ds["GBps"] or 0.0
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(dom, {}).get(args.prec)
        except Exception:
            traffic = None
    b_fwd = 12 * n + 4 * c * p_pad
    pipeline_gbs = b_fwd / (ms_step * 1e-3) / 1e9

    # forward + backward before the FP64 check and the calibration free
    # their large buffers (timed work right after large frees runs slower)
    adj = None
    if not args.no_backward:
        adj = fwd_bwd(genpk_amd, torch, pos, args.box, d, nbins, args.backward_steps, args.prec)
        adj["at_256"] = at_256

    check = None
    if not args.no_check:
        pw_t = pw.double().cpu().numpy()
        ref = genpk_amd.Plan(d, "f64")
        ref.deposit(pos, args.box, mass=1.0)
        pr, cr, kr = ref.power(nbins, mass=float(n))
        pr = pr.cpu().numpy()
        occ = cr.cpu().numpy() > 0
        rel = np.abs(pw_t[occ] / pr[occ] - 1)
        check = {"vs": "f64 path on the same particles", "max_rel": float(rel.max()),
                 "median_rel": float(np.median(rel)), "bins": int(occ.sum()),
                 "counts_equal": bool(np.array_equal(cnt.cpu().numpy(), cr.cpu().numpy()))}
        del ref
        torch.cuda.empty_cache()

    if args.copy_calibration:
        # reference kernels for the HBM byte counters, with the access width of
        # this library's kernels: k_fold (export_density) copies the d^3 cells
        # of the tile store into a padded grid with 4- or 8-byte accesses,
        # reading and writing c*d^3 bytes; a torch elementwise add over 2^30
        # floats (16-byte accesses) reads and writes 4 GiB
        plan.clear()
        plan.deposit(pos, args.box, mass=1.0)
        out = plan.export_density()
        a = torch.empty(1 << 30, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b, out


    bat = None
    if not args.no_batched:
        bat = batched(genpk_amd, torch, args.prec, args.batch_count)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pos, args.box, d, nbins, n, args.cpu_sample_log2)

    out = {
        "metric": "Gparticles/s CIC+FFT+P(k) @1024^3",
        "value": round(value, 3),
        "unit": "Gparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.prec,
        "data": "synthetic (seeded clustered particles generated on the GPU)",
        "config": {"workload": f"clustered {n} particles (half uniform, half in 1e5 power-law clumps), "
                               f"dims={d}, box={args.box} Mpc/h, forward CIC+r2c+P(k), nrbins={nbins}",
                   "n_particles": n, "dims": d, "nrbins": nbins, "precision": args.prec},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel_ms": ds["ms"], "alg_bytes_per_launch": ds["alg_bytes"]},
        "pipeline": {"alg_bytes": b_fwd, "GBps": round(pipeline_gbs, 1),
                     "frac_of_hbm_peak": round(pipeline_gbs / HBM_PEAK_GBS, 4)},
        "stages": per_stage,
        "accuracy": check,
        "fwd_bwd": adj,
        "batched": bat,
        "cpu_baseline": cpu,
        "input_gen_s": round(t_gen, 2),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
