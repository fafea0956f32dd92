#!/usr/bin/env python
"""Throughput of the GenPK power-spectrum path on one MI355X.

Metric: Gparticles/s of one forward pass -- CIC deposit + 3-D r2c FFT +
CIC-deconvolved binned P(k) -- at a 1024^3 grid, on the clustered
workload: 1024^3 particles in a 1000 Mpc/h box, half uniform and half in
10^5 Gaussian clumps (power-law member counts, radii 0.5-4 cells), seeded,
generated on the GPU (synthetic data).  Positions are resident in HBM
before the timed region; one step = clear + deposit + fused r2c/bin.

One JSON line is printed.  ``value`` is the FP32-grid plan (the workload's
float32 grid); ``f64`` times the same step with FP64 grid and FFT (the
reference's precision).  Both carry the HBM roofline of the whole forward
(the compulsory bytes 12N + 4cP over the step time) with the measured HBM
traffic per forward, the per-kernel stage times (HIP events on the plan's
stream), and the per-bin error against the FP64 CPU oracle (the reference
algorithm restated in C/OpenMP) run in full on the same particles, which
is also the CPU baseline.  Further objects: the two-species 512^3 cross
workload, forward + backward of the band-power loss, and the batched form.
Progress lines go to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
T0 = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="f32", choices=["f32", "f64"], help="precision of the headline value")
    ap.add_argument("--n-log2", type=int, default=30, help="log2 of the particle count (1024^3 = 2^30)")
    ap.add_argument("--dims", type=int, default=1024)
    ap.add_argument("--box", type=float, default=1000.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle (baseline and per-bin accuracy)")
    ap.add_argument("--cpu-sample-log2", type=int, default=None,
                    help="time the CPU oracle's deposit on a sample of 2^k particles (scaled) instead of all")
    ap.add_argument("--no-other-prec", action="store_true", help="skip the second precision's line")
    ap.add_argument("--no-backward", action="store_true", help="skip the forward+backward measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (64 uneven realisations) measurement")
    ap.add_argument("--no-cross", action="store_true", help="skip the two-species 512^3 cross measurement")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--backward-steps", type=int, default=3)
    ap.add_argument("--per-bin-table", default=None, help="write the per-bin error table to this file")
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


def measured_traffic(prec):
    """HBM bytes per launch of each forward kernel at the headline workload,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs."""
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        t = json.load(open(tfile))
    except (OSError, ValueError):
        return {}
    return {k: v[prec] for k, v in t.items() if isinstance(v, dict) and prec in v}


def time_forward(genpk_amd, torch, pos, box, d, nbins, prec, steps, warmup):
    """Wall time per step of `steps` forward steps (clear + deposit + fused
    power) bracketed by synchronisations, the per-stage HIP-event times and
    the last step's (power, count, keffs)."""
    n = pos.shape[0]
    plan = genpk_amd.Plan(d, prec)

    def step():
        plan.clear()
        plan.deposit(pos, box, mass=1.0)
        return plan.power(nbins, mass=float(n))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    plan.timing_read()  # drop warm-up marks
    plan.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pw, cnt, keffs = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan.timing(False)
    stages = plan.timing_read()
    out = [x.cpu().numpy() for x in (pw, cnt, keffs)]
    del plan
    torch.cuda.empty_cache()
    return (t1 - t0) * 1e3 / steps, stages, out


def forward_report(ms_step, stages, n, d, prec, steps):
    """Roofline of the whole forward plus per-stage figures."""
    c = 4 if prec == "f32" else 8
    p_pad = d * d * 2 * (d // 2 + 1)
    geo = tile_geometry(d)
    per_stage = {}
    for name, (tot, launches) in stages.items():
        avg = tot / max(launches, 1)
        b = stage_bytes(name, n, c, p_pad, geo, 12)
        per_stage[name] = {"ms": round(avg, 4), "launches_per_step": launches / steps,
                           "alg_bytes": b, "GBps": round(b / (avg * 1e-3) / 1e9, 1) if avg > 0 and b else None}
    dom = max(per_stage, key=lambda k: per_stage[k]["ms"] * per_stage[k]["launches_per_step"])
    ds = per_stage[dom]
    b_fwd = 12 * n + 4 * c * p_pad
    achieved = b_fwd / (ms_step * 1e-3) / 1e9
    traffic = measured_traffic(prec) if (n == 1 << 30 and d == 1024) else {}
    fwd_traffic = sum(traffic.get(k, 0) for k in per_stage) if traffic and all(k in traffic for k in per_stage
                                                                              if per_stage[k]["alg_bytes"]) else None
    roofline = {"bound": "hbm", "scope": "whole forward step (every kernel of clear + deposit + fused r2c/bin)",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes": b_fwd,
                "alg_bytes_formula": "12 N + 4 c P_pad (positions read once; grid written, FFT read + written, "
                                     "bin read once)",
                "traffic": fwd_traffic,
                "traffic_source": "sum over the forward's kernels of rocprofv3 FETCH_SIZE / WRITE_SIZE per launch, "
                                  "profiles/traffic.json" if fwd_traffic else None,
                "dominant_kernel": {"kernel": dom, "ms": ds["ms"], "alg_bytes_per_launch": ds["alg_bytes"],
                                    "achieved_GBps": ds["GBps"],
                                    "frac": round((ds["GBps"] or 0.0) / HBM_PEAK_GBS, 4),
                                    "traffic": traffic.get(dom)}}
    return roofline, per_stage


def fwd_bwd(genpk_amd, torch, pos, box, d, nbins, steps, prec, target, warm=1, reps=1):
    """Forward + backward of the band-power chi^2 loss against `target` with
    a masses tensor (all ones) that receives gradients, recompute policies
    "always" (the backward recomputes deposit + transform) and "auto";
    against two forwards: the headline step (constant mass, no autograd) and
    the same-inputs forward (the masses tensor)."""
    from genpk_amd.autograd import chi2_loss, power_spectrum
    n = pos.shape[0]
    plan = genpk_amd.Plan(d, prec)
    masses = torch.ones(n, dtype=torch.float32, device=pos.device)
    plan.clear()
    plan.deposit(pos, box, masses)
    _, cnt, _ = plan.power(nbins, mass=float(n))

    def fwd_headline():
        plan.clear()
        plan.deposit(pos, box, mass=1.0)
        return plan.power(nbins, mass=float(n))[0]

    def fwd_masses():
        plan.clear()
        plan.deposit(pos, box, masses)
        return plan.power(nbins, mass=float(n))[0]

    def fb(policy):
        tp = pos.detach().requires_grad_(True)
        tm = masses.detach().requires_grad_(True)
        P = power_spectrum(tp, box, d, nbins, masses=tm, total_mass=float(n), plan=plan, recompute=policy)
        return torch.autograd.grad(chi2_loss(P, target, cnt), (tp, tm))

    def timed(f, *a):
        # the median over `reps` groups of `steps` calls
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

    t_h = timed(fwd_headline)
    t_m = timed(fwd_masses)
    out = {"loss": "sum_b (P_b - T_b)^2 / sigma_b^2, sigma_b = T_b sqrt(2/count_b), T = P of a second seed",
           "grads": "positions and masses", "steps": steps, "reps": reps,
           "ms_fwd_headline": round(t_h, 3), "ms_fwd_same_inputs": round(t_m, 3)}
    plan.timing_read()
    for policy in ("always", "auto"):
        if policy == "always":
            plan.timing(True)
        t = timed(fb, policy)
        if policy == "always":
            plan.timing(False)
            st = plan.timing_read()
            calls = warm + steps * reps
            out["stages_always_ms"] = {k: round(v[0] / calls, 3) for k, v in st.items()}
        out[f"ms_fwd_bwd_{policy}"] = round(t, 3)
        out[f"ratio_{policy}"] = round(t / t_h, 3)
        out[f"ratio_{policy}_same_inputs"] = round(t / t_m, 3)
    out["policy_of_ratio"] = ("always: the backward recomputes the density (tile deposit from the tile order the "
                              "forward kept) and the transform; the field is not stored")
    out["ratio"] = out["ratio_always"]
    out["ratio_denominator"] = "ms_fwd_headline (constant mass, the value's step)"
    del plan
    torch.cuda.empty_cache()
    return out


def batched(genpk_amd, torch, prec, count, reps=5):
    """Uneven batch: `count` realisations, realisation r with seed 100 + r,
    n_side log-uniform in [64, 256] (n_side^3 particles), dims uniform over
    {128, 256, 384, 512}, 0.3-cell Gaussian lattice in a 100 Mpc/h box,
    nrbins = dims.  Plain: one plan per dims, each realisation run alone
    (clear + deposit + fused power, then synchronize), times summed; with
    the default plan options and with the options the autotuner picked.
    Batched: the autotuned hipGraph over all realisations, one launch."""
    from genpk_amd import batch as gb
    from genpk_amd import synth
    box = 100.0
    jobs = []
    for ns, d, sd in synth.uneven_realisations(count):
        pos = synth.gaussian_lattice_torch(ns, box, d, 0.3, seed=sd)
        jobs.append(dict(pos=pos, boxsize=box, dims=d, nrbins=d, total_mass=float(pos.shape[0]), precision=prec))
    n_tot = sum(j["pos"].shape[0] for j in jobs)

    def plain_time(options):
        plans = {d: genpk_amd.Plan(d, prec) for d in sorted({j["dims"] for j in jobs})}
        if options:
            for d, p in plans.items():
                o = options[f"{d}/{prec}"]
                p.set_options(o["tiles"], o["fft_variant"])

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
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            outs = plain()
            ts.append(time.perf_counter() - t0)
        del plans
        torch.cuda.empty_cache()
        return sorted(ts)[1], outs

    t_plain, _ = plain_time(None)
    best, rep = gb.tune(jobs, use_cache=False, log=log)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        best.launch()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_b = sorted(ts)[reps // 2]
    t_plain_tuned, ref = plain_time(rep["per_dims"])
    dev = max(float(((best.results(i)[0] - ref[i]).abs() / ref[i].abs().clamp(min=1e-300)).max())
              for i in range(len(jobs)))
    best.close()
    torch.cuda.empty_cache()
    return {"realisations": count, "particles": n_tot, "dims": sorted({j["dims"] for j in jobs}),
            "ms_plain_sum": round(t_plain * 1e3, 3), "ms_plain_sum_tuned_options": round(t_plain_tuned * 1e3, 3),
            "ms_batched": round(t_b * 1e3, 3), "speedup": round(t_plain / t_b, 3),
            "speedup_vs_plain_tuned_options": round(t_plain_tuned / t_b, 3),
            "Gparticles_per_s_batched": round(n_tot / t_b / 1e9, 3),
            "autotune": rep, "max_rel_dev_from_plain_same_options": dev}


def cross_512(genpk_amd, torch, prec, steps=10, warmup=2):
    """Two species on a 512^3 grid: 2 x 512^3 Zel'dovich particles (k^-2
    displacement spectrum, 1-cell rms, seed 1) in a 250 Mpc/h box, species A
    mass 1, species B mass 0.2 on the lattice offset by half a cell; one step
    = clear + deposit of both + one fused pass binning P_AA, P_BB and P_AB."""
    from genpk_amd import synth
    n_side, box, d = 512, 250.0, 512
    psi = synth.zeldovich_displacement_torch(n_side, 1.0, box, d, seed=1)
    A = synth.zeldovich_torch(n_side, box, d, psi=psi)
    B = synth.zeldovich_torch(n_side, box, d, offset_cells=0.5, psi=psi)
    del psi
    torch.cuda.empty_cache()
    n = A.shape[0]
    pa, pb = genpk_amd.Plan(d, prec), genpk_amd.Plan(d, prec)

    def step():
        pa.clear()
        pb.clear()
        pa.deposit(A, box, mass=1.0)
        pb.deposit(B, box, mass=0.2)
        return pa.power(d, other=pb, mass=1.0 * n, other_mass=0.2 * n)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    c = 4 if prec == "f32" else 8
    p_pad = d * d * 2 * (d // 2 + 1)
    b_alg = 2 * (12 * n + 3 * c * p_pad) + 2 * c * p_pad
    del pa, pb, A, B
    torch.cuda.empty_cache()
    return {"workload": "2 x 512^3 Zel'dovich particles (k^-2, 1-cell rms, seed 1), box 250 Mpc/h, d=512; "
                        "species A m=1, species B m=0.2 offset half a cell; auto A, auto B, cross A x B",
            "precision": prec, "particles": 2 * n, "ms_per_step": round(ms, 3),
            "Gparticles_per_s": round(2 * n / ms / 1e6, 3), "alg_bytes": b_alg,
            "alg_bytes_formula": "2 (12 N + 3 c P_pad) + 2 c P_pad (one fused bin pass reads both spectra)",
            "frac_of_hbm_peak": round(b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_oracle(pos_host, box, dims, nbins, sample_log2=None):
    """The FP64 CPU oracle (fieldize -> r2c -> powerspectrum, the reference
    algorithm restated in C/OpenMP) on the host: on every particle, or with
    sample_log2 the deposit on 2^k of them (then scaled).  Returns the
    baseline record and (power, count, keffs) of the full run (or None)."""
    import numpy as np
    import oracle
    n = pos_host.shape[0]
    sample = pos_host
    if sample_log2 is not None and n > (1 << sample_log2):
        sample = pos_host[:: n >> sample_log2]
    ns = sample.shape[0]
    grid = oracle.new_field(dims)
    t0 = time.perf_counter()
    oracle.fieldize(box, dims, grid, sample, None, 1.0, 1)
    t_dep = time.perf_counter() - t0
    log(f"cpu oracle deposit of {ns} particles: {t_dep:.1f} s")
    t0 = time.perf_counter()
    F = oracle.fft_r2c_3d(grid)
    t_fft = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = oracle.powerspectrum(dims, F, F, nbins, float(n), float(n))
    t_bin = time.perf_counter() - t0
    full = ns == n
    t_total = t_dep * (n / ns) + t_fft + t_bin
    if full:
        how = (f"all {n} particles of the workload: fieldize {t_dep:.2f} s, FP64 r2c {t_fft:.2f} s, "
               f"powerspectrum {t_bin:.2f} s (one run)")
    else:
        how = (f"every {n // ns}th of the {n} particles ({ns}) deposited ({t_dep:.2f} s, scaled x{n / ns:.0f}); "
               f"FP64 r2c ({t_fft:.2f} s) and binning ({t_bin:.2f} s) at the full {dims}^3")
    rec = {"value": n / t_total / 1e9, "unit": "Gparticles/s", "cores": oracle.num_threads(), "kind": "port",
           "sample": how, "seconds": round(t_dep + t_fft + t_bin, 2), "extrapolated": not full,
           "nproc": os.cpu_count(), "cpu_model": cpu_model()}
    return rec, (res if full else None)


def accuracy(got, ref, label):
    """Per-bin comparison of a GPU result (power, count, keffs) with the
    oracle's: counts bit-exact, keffs and P relative errors per occupied bin."""
    import numpy as np
    P, c, k = got
    po, co, ko = ref
    occ = co > 0
    rel = np.abs(P[occ] / po[occ] - 1)
    krel = np.abs(k[occ] / ko[occ] - 1)
    bins = np.nonzero(occ)[0]
    return {"vs": f"FP64 CPU oracle on the same particles ({label})", "bins": int(occ.sum()),
            "counts_equal": bool(np.array_equal(c, co)), "keffs_max_rel": float(krel.max()),
            "max_rel": float(rel.max()), "worst_bin": int(bins[rel.argmax()]), "median_rel": float(np.median(rel)),
            "p99_rel": float(np.quantile(rel, 0.99)), "bins_over_1e-5": int((rel > 1e-5).sum())}


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
    other = "f64" if args.prec == "f32" else "f32"
    log(f"device {torch.cuda.get_device_name(0)}; workload 2^{args.n_log2} particles, d={d}")

    at_256 = None
    if not args.no_backward:
        # the fwd+bwd ratio at its own target size, measured before anything
        # large is allocated: 256^3 Zel'dovich particles (k^-2, 1-cell rms,
        # seed 1) on a 256^3 grid in a 250 Mpc/h box, target P from seed 2
        n3, d3, box3 = 256, 256, 250.0
        pos3 = synth.zeldovich_torch(n3, box3, d3, seed=1)
        tq = genpk_amd.Plan(d3, args.prec)
        tq.deposit(synth.zeldovich_torch(n3, box3, d3, seed=2), box3, mass=1.0)
        target3 = tq.power(d3, mass=float(n3 ** 3))[0]
        del tq
        at_256 = fwd_bwd(genpk_amd, torch, pos3, box3, d3, d3, 10, args.prec, target3, warm=3, reps=5)
        at_256["workload"] = f"{n3}^3 Zel'dovich particles (seed 1; target seed 2), d={d3}, box {box3}, nrbins {d3}"
        del pos3
        torch.cuda.empty_cache()
        log(f"fwd+bwd at 256^3: ratio {at_256['ratio']}")

    t_gen = time.perf_counter()
    pos = synth.clumped_torch(n, args.box, d, n_clumps=100000, seed=4, device=dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen

    ms_step, stages, res_main = time_forward(genpk_amd, torch, pos, args.box, d, nbins, args.prec, args.steps,
                                             args.warmup)
    value = n / (ms_step * 1e-3) / 1e9
    roofline, per_stage = forward_report(ms_step, stages, n, d, args.prec, args.steps)
    log(f"forward {args.prec}: {ms_step:.3f} ms = {value:.2f} Gparticles/s")

    other_line, res_other = None, None
    if not args.no_other_prec:
        ms_o, stages_o, res_other = time_forward(genpk_amd, torch, pos, args.box, d, nbins, other, args.steps,
                                                 args.warmup)
        rl_o, ps_o = forward_report(ms_o, stages_o, n, d, other, args.steps)
        other_line = {"dtype": other, "value": round(n / (ms_o * 1e-3) / 1e9, 3), "unit": "Gparticles/s",
                      "ms_per_step": round(ms_o, 3), "steps": args.steps, "warmup": args.warmup,
                      "roofline": rl_o, "stages": ps_o}
        log(f"forward {other}: {ms_o:.3f} ms")

    # forward + backward before the large frees below (timed work right
    # after large frees runs slower); target: the forward P of a second seed
    adj = None
    if not args.no_backward:
        tq = genpk_amd.Plan(d, args.prec)
        pos2 = synth.clumped_torch(n, args.box, d, n_clumps=100000, seed=5, device=dev)
        tq.deposit(pos2, args.box, mass=1.0)
        target = tq.power(nbins, mass=float(n))[0]
        del tq, pos2
        torch.cuda.empty_cache()
        adj = fwd_bwd(genpk_amd, torch, pos, args.box, d, nbins, args.backward_steps, args.prec, target)
        adj["at_256"] = at_256
        log(f"fwd+bwd at 1024^3: ratio {adj['ratio']} ({adj['ms_fwd_bwd_always']} ms)")

    if args.copy_calibration:
        # reference kernels for the HBM byte counters, with the access width of
        # this library's kernels: k_fold (export_density) copies the d^3 cells
        # of the tile store into a padded grid with 4- or 8-byte accesses,
        # reading and writing c*d^3 bytes; a torch elementwise add over 2^30
        # floats (16-byte accesses) reads and writes 4 GiB
        plan = genpk_amd.Plan(d, args.prec)
        plan.deposit(pos, args.box, mass=1.0)
        out = plan.export_density()
        a = torch.empty(1 << 30, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b, out, plan
        torch.cuda.empty_cache()

    cross = None
    if not args.no_cross:
        cross = cross_512(genpk_amd, torch, args.prec)
        log(f"two species 512^3: {cross['ms_per_step']} ms")

    bat = None
    if not args.no_batched:
        bat = batched(genpk_amd, torch, args.prec, args.batch_count)
        log(f"batched: speedup {bat['speedup']}")

    cpu, acc_main, acc_other = None, None, None
    if not args.no_cpu:
        pos_h = pos.cpu().numpy()
        del pos
        torch.cuda.empty_cache()
        cpu, ref = cpu_oracle(pos_h, args.box, d, nbins, args.cpu_sample_log2)
        del pos_h
        if ref is not None:
            acc_main = accuracy(res_main, ref, args.prec)
            if res_other is not None:
                acc_other = accuracy(res_other, ref, other)
                other_line["accuracy"] = acc_other
            if args.per_bin_table:
                po, co, ko = ref
                occ = np.nonzero(co > 0)[0]
                with open(args.per_bin_table, "w") as f:
                    f.write(f"# clustered 2^{args.n_log2} particles, d={d}, box {args.box}: |P_GPU / P_oracle - 1| "
                            f"per occupied bin\n# bin count k_eff P_oracle {args.prec} {other}\n")
                    for b in occ:
                        e1 = abs(res_main[0][b] / po[b] - 1)
                        e2 = abs(res_other[0][b] / po[b] - 1) if res_other is not None else float("nan")
                        f.write(f"{b} {co[b]} {ko[b]:.6e} {po[b]:.6e} {e1:.3e} {e2:.3e}\n")

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
        "roofline": roofline,
        "stages": per_stage,
        "accuracy": acc_main,
        other: other_line,
        "cross_512": cross,
        "fwd_bwd": adj,
        "batched": bat,
        "cpu_baseline": cpu,
        "input_gen_s": round(t_gen, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
