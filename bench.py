#!/usr/bin/env python3
"""Headline benchmark: 3D Laplace FMM on one MI355X.

Workload (the 4M-particle 3D configuration of the metric): N = 4,194,304
particles, positions uniform in [0,1)^3 then charges uniform in [-1,1]
(seeded numpy generator, positions drawn first), sources = targets, uniform
Morton octree with leaves at level 6 (at most 64 particles per leaf).  One
step = one full solve through the C ABI with inputs and outputs resident in
HBM: extent, keys, radix sort, tree levels, P2M, M2M, M2L+L2L, P2P+L2P and the
scatter back to input order.

The metric is quoted at rel-L2 potential error <= 1e-6 against the FP64
direct sum, so the headline runs the smallest expansion order that meets it
on this workload: 13 terms (p = 12), with the M2L of levels >= 4 on the
matrix cores in split FP16 (everything else FP64).  The same line reports the
p = 10 configuration (11 terms) and the all-FP64 path at both orders, each
with its own measured error, all against one direct-sum reference on the same
sampled targets.

Also reported: the Plummer sphere (4M particles, scale radius 0.05) on the
adaptive octree -- fixed-order FP64, fixed-order mixed, and per-box orders
from the order network trained in this run on the GPU -- the 2D 1M
configuration, 512 independent systems, and CPU baselines (the CPU oracle,
median of 5 solves after a warm-up).

Printed: one JSON line with particles/s, the accuracy, the M2L roofline of
the headline on its algorithmic flop count (kernel time from HIP events the
library records on its stream during the timed steps), the issued matrix-core
rate beside it, HBM rates of the P2P, M2M and L2L kernels, and the CPU
baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/sec (3D, 4M) at rel-L2 err ≤1e-6 vs FP64 direct; M2L MFMA util"
FP64_PEAK_TFLOPS = 78.6
FP16_PEAK_TFLOPS = 2500.0  # dense MFMA
HBM_PEAK_GBS = 8000.0


def m2l_flops_per_op(N):
    """Algorithmic FP64 flops of one rotation-based M2L with N terms:
    Rz and Ry forward and back (2 (6 Nx2 + sum_{n<N} (n+1)(4n+2))) plus the
    axial translation (4 sum_{t=1..N} t^2).  10,556 at N = 13."""
    nx2 = N * (N + 1) // 2
    rot = sum((n + 1) * (4 * n + 2) for n in range(N))
    return 2 * (6 * nx2 + rot) + 4 * sum(t * t for t in range(1, N + 1))


def m2l16_mfma_flops_per_op(lfmm, N):
    """Matrix-core flops issued per M2L op (one target column, one direction)
    by k_m2l16r: per K chunk of 16 and row tile of 32, one hi*hi MFMA of
    32x32x16 (2 * 16384 flops for 32 columns) for the chunks below the row
    tile's bound k1, and hi*lo + lo*hi for the chunks below its bound k3;
    averaged over the directions as the 8 child classes use them (each class
    its 189 offsets)."""
    _, _, kpre = lfmm.host_m2l16_pack(N)  # (316, RT, 2) = (k3, k1)
    k3 = kpre[:, :, 0].sum(axis=1)
    k1 = kpre[:, :, 1].sum(axis=1)
    per_dir = (k1 + 2 * k3) * 2 * 32 * 16  # flops per column
    cd = lfmm.host_op_tables(N, "class_dirs").reshape(8, -1)
    return float(per_dir[cd].mean())


def plummer(n, seed, a=0.05, rmax=1.0):
    """Plummer sphere of scale radius a truncated at rmax, centred at
    (1/2, 1/2, 1/2): radii r = a / sqrt(u^(-2/3) - 1) (redrawn beyond rmax),
    isotropic directions; charges U[-1,1]."""
    rng = np.random.default_rng(seed)
    r = np.empty(0)
    while len(r) < n:
        u = rng.random(2 * n)
        rr = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = np.concatenate([r, rr[rr <= rmax]])
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return 0.5 + r[:n, None] * d, rng.uniform(-1.0, 1.0, n)


def batch_systems(n_sys, seed):
    """Independent systems of log-uniform size in [1024, 65536]; even ones
    uniform in the unit cube, odd ones a mixture of 4 Gaussian blobs
    (sigma 0.05, centres U[0.2, 0.8]^3); charges U[-1,1]."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(1024), np.log(65536), n_sys))).astype(np.int64)
    xs, qs = [], []
    for i, n in enumerate(sizes):
        if i % 2:
            c = rng.uniform(0.2, 0.8, (4, 3))
            xs.append(c[rng.integers(0, 4, n)] + rng.normal(0.0, 0.05, (n, 3)))
        else:
            xs.append(rng.random((n, 3)))
        qs.append(rng.uniform(-1.0, 1.0, n))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return np.concatenate(xs), np.concatenate(qs), off


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def median_time(fn, reps=5):
    """Median wall time of reps calls after one warm-up call."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--terms", type=int, default=13)
    ap.add_argument("--depth", type=int, default=7)
    ap.add_argument("--fp64", action="store_true", help="headline on the all-FP64 path (default: mixed M2L)")
    ap.add_argument("--mixed", action="store_true", help="(default) M2L of levels >= 4 on the matrix cores")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--sample", type=int, default=4096, help="targets checked against the direct sum")
    ap.add_argument("--cpu-n", type=int, default=0, help="particles in the CPU-baseline run (0: the full workload)")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed CPU solves per configuration (median)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary configurations")
    ap.add_argument("--no-plummer", action="store_true")
    ap.add_argument("--no-uniform-plummer", action="store_true",
                    help="skip the fixed-depth Plummer solve (2.7 s per solve)")
    ap.add_argument("--workload", choices=("uniform", "plummer", "2d", "batch"), default="uniform",
                    help="particles of the timed steps (profiling runs): the uniform cube of the metric, the "
                         "Plummer sphere on the adaptive octree, (2d) only the 2D solves of 1,048,576 points "
                         "at p = 20, FP64 then FP32 M2L sums, or (batch) only the 512-system batch as one "
                         "adaptive octree, FP64; the last two print one JSON line of their own")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="one device copy of this many MiB before the steps (calibrates HBM counters)")
    args = ap.parse_args()
    mixed = not args.fp64

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (the path runs on a single MI355X); "
              "--gpus %d / WORLD_SIZE=%d is not supported" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge

    lfmm = ge._import_package()
    dev = torch.device("cuda:0")
    if args.workload == "2d":
        print(json.dumps(two_d_steps(args, lfmm, torch, dev)), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(batch_steps(args, lfmm, torch, dev)), flush=True)
        return
    rng = np.random.default_rng(args.seed)
    x = rng.random((args.n, 3))
    q = rng.uniform(-1.0, 1.0, args.n)
    adaptive = args.workload == "plummer"
    if adaptive:  # self term excluded on this workload (sampled errors need no correction)
        x, q = plummer(args.n, args.seed + 11)
    xd = torch.from_numpy(x).to(dev)
    qd = torch.from_numpy(q).to(dev)
    pot = torch.empty(args.n, dtype=torch.float64, device=dev)
    field = torch.empty((args.n, 3), dtype=torch.float64, device=dev)
    ctx = lfmm.Context(0)
    idx = np.random.default_rng(args.seed + 1).choice(args.n, min(args.sample, args.n), replace=False)
    idx_d = torch.from_numpy(idx).to(dev)

    if args.copy_mib > 0:
        a_ = torch.empty(args.copy_mib << 17, dtype=torch.float64, device=dev)
        b_ = torch.empty_like(a_)
        b_.copy_(a_)
        torch.cuda.synchronize()
        del a_, b_

    def run(terms, mix, steps, warmup):
        def step(timings=False, count=False):
            if adaptive:
                return lfmm.fmm3d(xd, qd, terms=terms, leaf_max=64, adaptive=True, exclude_self=True, ctx=ctx,
                                  out=(pot, field), timings=timings, count_ops=count, mixed=mix)[2]
            return lfmm.fmm3d(xd, qd, terms=terms, tree_depth=args.depth, ctx=ctx, out=(pot, field),
                              timings=timings, count_ops=count, mixed=mix)[2]

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        reports = []
        t0 = time.perf_counter()
        for _ in range(steps):
            reports.append(step(timings=True))
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        sample = (pot[idx_d].cpu().numpy(), field[idx_d].cpu().numpy())
        log("terms=%d %s: %d steps, %.3f ms/step" % (terms, "mixed" if mix else "fp64", steps, 1e3 * elapsed / steps))
        return {"ms": 1e3 * elapsed / steps, "value": args.n / (elapsed / steps), "reports": reports,
                "sample": sample, "step": step}

    def mean_of(reports, key):
        return float(np.mean([r[key] for r in reports]))

    head = run(args.terms, mixed, args.steps, args.warmup)
    # exact op counts (after the timed region)
    counts = head["step"](timings=True, count=True)
    reports = head["reports"]
    m2l_ms = mean_of(reports, "t_m2l_kernels_ms")
    leaf_ms = mean_of(reports, "t_leaf_kernel_ms")
    alg = m2l_flops_per_op(args.terms)
    if mixed:
        # dominant kernel: the matrix-core M2L of levels >= 4, on its
        # algorithmic (rotation-based) flop count; the MFMA work it issues
        # (dense operators, three split products) beside it
        mx_ms = mean_of(reports, "t_m2l_mixed_ms")
        mx_launches = reports[0]["m2l_mixed_launches"]
        mx_ops = counts["m2l_ops_mixed"]
        ach = mx_ops * alg / (mx_ms * 1e-3) / 1e12
        issued = mx_ops * m2l16_mfma_flops_per_op(lfmm, args.terms) / (mx_ms * 1e-3) / 1e12
        roofline = {"kernel": "k_m2l16r<N> (split-FP16 MFMA M2L, levels >= 4, one launch per level)", "bound": "mfma",
                    "achieved": round(ach, 3), "peak": FP16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(ach / FP16_PEAK_TFLOPS, 4), "traffic": None,
                    "flops_counted": "algorithmic rotation-based M2L flops (m2l_ops x %d)" % alg,
                    "m2l_ops": mx_ops, "flops_per_op": alg, "launches": mx_launches,
                    "avg_launch_ms": round(mx_ms / max(mx_launches, 1), 4), "m2l_kernels_ms_per_step": round(mx_ms, 4),
                    "mfma_issued": {"tflops": round(issued, 1), "frac_of_peak": round(issued / FP16_PEAK_TFLOPS, 4),
                                    "flops_per_op": round(m2l16_mfma_flops_per_op(lfmm, args.terms), 1)},
                    "precision": "fp16 hi/lo operands (22 bits), fp32 accumulation",
                    "fp64_levels": {"m2l_ops": counts["m2l_ops"] - mx_ops, "ms_per_step": round(m2l_ms - mx_ms, 4)}}
    else:
        launches = reports[0]["m2l_launches"]
        ach = counts["m2l_ops"] * alg / (m2l_ms * 1e-3) / 1e12
        roofline = {"kernel": "k_rtr<N,M2L> (M2L+L2L, one launch per level)", "bound": "mfma",
                    "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
                    "flops_counted": "algorithmic rotation-based M2L flops", "m2l_ops": counts["m2l_ops"],
                    "flops_per_op": alg, "launches": launches, "avg_launch_ms": round(m2l_ms / max(launches, 1), 4),
                    "m2l_kernels_ms_per_step": round(m2l_ms, 4), "precision": "fp64 VALU"}
    prof = os.path.join(ROOT, "profiles", "m2l_traffic.json")
    if os.path.exists(prof):
        try:
            t = json.load(open(prof))
            if (t.get("n") == args.n and t.get("terms") == args.terms and t.get("tree_depth") == args.depth
                    and bool(t.get("mixed", False)) == mixed):
                roofline["traffic"] = t.get("hbm_bytes_per_launch")
                roofline["traffic_source"] = "profiles/m2l_traffic.json (rocprofv3 FETCH_SIZE, WRITE_SIZE)"
                if t.get("mfma_busy") is not None:
                    roofline["mfma_busy_by_level"] = t.get("mfma_busy")
                    roofline["mfma_busy_source"] = "profiles/mfma_busy_fmm3d_4m_p12_d7_mixed.md (rocprofv3 SQ_VALU_MFMA_BUSY_CYCLES)"
        except (OSError, ValueError):
            pass
    # HBM rates of the P2P, M2M and L2L kernels on their algorithmic bytes
    n_out = args.n
    p2p_bytes = 32.0 * counts["p2p_source_reads"] + args.n * (32.0 + 32.0) + n_out * 32.0

    def rate(nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else None
        return {"bytes": nbytes, "ms_per_step": round(ms, 4), "gbs": round(gbs, 1) if gbs else None,
                "frac_of_hbm": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    hbm = {"peak_gbs": HBM_PEAK_GBS,
           "p2p": dict(rate(p2p_bytes, mean_of(reports, "t_p2p_ms")),
                       kernel="k_p2p_multi<2> with its lane setup",
                       bytes_counted="32 B per near-list source per leaf + 32 B target + 32 B far field + 32 B out"),
           "m2m": dict(rate(counts["m2m_bytes"], mean_of(reports, "t_m2m_ms")),
                       kernel="k_m2l16r<N,CTW,true> (M2M into levels >= 4 on the matrix cores) + k_rtr<N,M2M> (FP64, levels 2-3)",
                       bytes_counted="every multipole of levels >= 3 read once, levels >= 2 written once"),
           "l2l": dict(rate(counts["l2l_bytes"], mean_of(reports, "t_l2l_ms")), kernel="k_rtr<N,L2L>",
                       bytes_counted="parent locals read once, child locals written once (levels >= 4)")}
    p2p_pairs = counts["p2p_pairs"]
    leaf = {"kernel": "k_l2p<N> + k_p2p (L2P, then P2P)", "ms_per_step": round(leaf_ms, 4), "p2p_pairs": p2p_pairs,
            "p2p_gflops": round(20.0 * p2p_pairs / (mean_of(reports, "t_p2p_ms") * 1e-3) / 1e9, 1),
            "p2p_frac_of_fp64_peak": round(20.0 * p2p_pairs / (mean_of(reports, "t_p2p_ms") * 1e-3) / 1e12 /
                                           FP64_PEAK_TFLOPS, 4)}
    stages = {k: round(mean_of(reports, k), 4) for k in ("t_p2m_ms", "t_m2m_ms", "t_l2l_ms", "t_m2l_kernels_ms",
                                                         "t_l2p_ms", "t_p2p_ms")}

    # secondary configurations: p = 10 (11 terms) and the all-FP64 path
    extras = []
    if not args.no_extra:
        for terms, mix in ((11, True), (args.terms, False), (11, False)):
            if (terms, mix) == (args.terms, mixed):
                continue
            r = run(terms, mix, 3, 1)
            c = r["step"](count=True)
            e = {"terms": terms, "m2l": "split-FP16 MFMA (levels >= 4)" if mix else "FP64", "r": r}
            if not mix:  # the FP64 path's own fraction of its peak (FP64 VALU = FP64 MFMA = 78.6 TF)
                tm = mean_of(r["reports"], "t_m2l_kernels_ms")
                tf = c["m2l_ops"] * m2l_flops_per_op(terms) / (tm * 1e-3) / 1e12
                e["m2l_fp64"] = {"tflops": round(tf, 2), "frac_of_fp64_peak": round(tf / FP64_PEAK_TFLOPS, 4),
                                 "m2l_kernels_ms": round(tm, 3)}
            extras.append(e)
        roofline["fp64_path"] = next((e["m2l_fp64"] for e in extras if e["terms"] == args.terms and "m2l_fp64" in e),
                                     None)

    # Plummer sphere on the adaptive octree: fixed-order FP64 (the plain
    # method), fixed-order mixed, and learned per-box orders (mixed)
    plum = None
    if not args.no_extra and not args.no_plummer:
        plum = plummer_leg(args, lfmm, torch, ctx, dev, idx, idx_d, pot, field)

    # 2D, 1,048,576 uniform points, p = 20 (21 terms), leaf_max 64: FP64
    # against FP32 per-direction M2L sums at levels >= 4
    runs2d = []
    if not args.no_extra:
        r2 = np.random.default_rng(args.seed + 31)
        x2, q2 = r2.random((1048576, 2)), r2.uniform(-1.0, 1.0, 1048576)
        x2d, q2d = torch.from_numpy(x2).to(dev), torch.from_numpy(q2).to(dev)
        for mix in (False, True):
            def step2():
                return lfmm.fmm2d(x2d, q2d, terms=21, leaf_max=64, exclude_self=True, ctx=ctx, mixed=mix)
            p2_, f2_, rep2 = step2()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                step2()
            torch.cuda.synchronize()
            ms2 = 1e3 * (time.perf_counter() - t0) / 5
            runs2d.append({"m2l": "FP32 sums at levels >= 4, FP64 accumulation" if mix else "FP64",
                           "tree_depth": rep2["tree_depth"], "ms_per_step": round(ms2, 3),
                           "value": 1048576 / (ms2 * 1e-3), "sample": (p2_.cpu().numpy(), f2_.cpu().numpy())})
            log("2d 1M p=20 %s %.3f ms/step" % ("fp32-m2l" if mix else "fp64", ms2))

    # many independent systems in one call (p = 8, leaf_max 32), against a
    # loop of single solves over the same device-resident systems
    batch = None
    if not args.no_extra:
        xb, qb, offb = batch_systems(512, args.seed + 21)
        xbd, qbd = torch.from_numpy(xb).to(dev), torch.from_numpy(qb).to(dev)
        pb = torch.empty(len(qb), dtype=torch.float64, device=dev)
        fb = torch.empty((len(qb), 3), dtype=torch.float64, device=dev)

        def run_batch(forest, mix=False):
            def bstep():
                lfmm.fmm3d_batch(xbd, qbd, offb, terms=9, leaf_max=32, ctx=ctx, out=(pb, fb), adaptive=forest,
                                 mixed=mix)
            bstep()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                bstep()
            torch.cuda.synchronize()
            return 1e3 * (time.perf_counter() - t0) / 3, (pb.cpu().numpy(), fb.cpu().numpy())
        # one adaptive octree for all systems, then the worker pool of
        # uniform-tree solves, then a plain loop of single solves
        fms, fsample = run_batch(True)
        mms, msample = run_batch(True, True)
        wms, wsample = run_batch(False)
        t0 = time.perf_counter()
        for s_ in range(512):
            a_, b_ = int(offb[s_]), int(offb[s_ + 1])
            lfmm.fmm3d(xbd[a_:b_], qbd[a_:b_], terms=9, leaf_max=32, ctx=ctx, out=(pb[a_:b_], fb[a_:b_]))
        torch.cuda.synchronize()
        sms = 1e3 * (time.perf_counter() - t0)
        batch = {"workload": "512 independent systems, sizes log-uniform in [1024, 65536] (%d particles in all), "
                             "half uniform in the unit cube, half 4 Gaussian blobs (sigma 0.05, centres "
                             "U[0.2,0.8]^3), q~U[-1,1], p=8 (9 terms), leaf_max 32, FP64" % len(qb),
                 "method": "one adaptive octree for all systems (each in its own cell, one solve)",
                 "particles": int(len(qb)), "ms_per_batch": round(fms, 3), "value": len(qb) / (fms * 1e-3),
                 "unit": "particles/s", "loop_of_single_solves_ms": round(sms, 3),
                 "speedup_vs_loop": round(sms / fms, 2),
                 "forest_mixed": {"method": "the same forest with the split-FP16 matrix-core M2L at each system's "
                                           "levels >= 4", "ms_per_batch": round(mms, 3),
                                 "value": len(qb) / (mms * 1e-3)},
                 "worker_pool": {"method": "8 worker contexts, one uniform-tree solve per system",
                                 "ms_per_batch": round(wms, 3), "value": len(qb) / (wms * 1e-3)},
                 "samples": {"forest": fsample, "forest_mixed": msample, "worker_pool": wsample}}
        log("batch of 512 systems: forest %.3f ms (mixed M2L %.3f ms), worker pool %.3f ms, loop of single solves "
            "%.3f ms" % (fms, mms, wms, sms))

    accuracy = None
    other = []
    if not args.no_error:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t1 = time.perf_counter()
        dp, df = oracle.direct3d(x, q, tgt_index=idx, include_self=False)
        log("direct sum on %d targets took %.1f s" % (len(idx), time.perf_counter() - t1))

        def errors(sample):
            p, f = sample
            # remove the reference's clamped self term q_i / 1e-4 (the Plummer workload excludes it)
            phys = p if adaptive else p - q[idx] * 1e4
            return (float(np.linalg.norm(phys - dp) / np.linalg.norm(dp)),
                    float(np.linalg.norm(f - df) / np.linalg.norm(df)))

        ep, ef = errors(head["sample"])
        accuracy = {"targets": int(len(idx)), "rel_l2_potential": ep, "rel_l2_field": ef,
                    "meets_1e-6": bool(ep <= 1e-6),
                    "reference_convention": dict(zip(("max_err", "l2_err", "rel_err"),
                                                     map(float, oracle.read_error(dp + q[idx] * 1e4,
                                                                                  head["sample"][0] +
                                                                                  (q[idx] * 1e4 if adaptive else 0)))))}
        for e in extras:
            e["rel_l2_potential"], e["rel_l2_field"] = errors(e["r"]["sample"])
        if plum is not None:
            t1 = time.perf_counter()
            pdp, pdf = oracle.direct3d(plum["x"], plum["q"], tgt_index=idx, include_self=False)
            log("plummer direct sum took %.1f s" % (time.perf_counter() - t1))
            for e in plum["runs"]:
                p_, f_ = e.pop("sample")
                e["rel_l2_potential"] = float(np.linalg.norm(p_ - pdp) / np.linalg.norm(pdp))
                e["rel_l2_field"] = float(np.linalg.norm(f_ - pdf) / np.linalg.norm(pdf))
                e["meets_1e-6"] = bool(e["rel_l2_potential"] <= 1e-6)
    if plum is not None:
        for e in plum["runs"]:
            e.pop("sample", None)
        plum.pop("x")
        plum.pop("q")
        runs = plum["runs"]

        def fastest(pred):
            ok = [e for e in runs if pred(e) and e.get("meets_1e-6")]
            return min(ok, key=lambda e: e["ms_per_step"]) if ok else None
        base = fastest(lambda e: e["m2l"] == "FP64" and e["tree"] == "adaptive" and e["orders"] == "fixed")
        base_u = fastest(lambda e: e["m2l"] == "FP64" and e["tree"] == "uniform depth")
        learned = fastest(lambda e: e["orders"] == "learned")
        fixed_mx = fastest(lambda e: e["m2l"] != "FP64" and e["orders"] == "fixed" and e["tree"] == "adaptive")
        plum["speedup_at_1e-6"] = {
            "learned_mixed_vs_fixed_fp64_adaptive": round(base["ms_per_step"] / learned["ms_per_step"], 2)
            if base and learned else None,
            "learned_mixed_vs_fixed_fp64_uniform_depth": round(base_u["ms_per_step"] / learned["ms_per_step"], 2)
            if base_u and learned else None,
            "fixed_mixed_vs_fixed_fp64_adaptive": round(base["ms_per_step"] / fixed_mx["ms_per_step"], 2)
            if base and fixed_mx else None}
    if runs2d and not args.no_error:
        i2 = np.random.default_rng(args.seed + 32).choice(1048576, 1024, replace=False)
        dp2, df2 = oracle.direct2d(x2, q2, tgt_index=i2, include_self=False)
        for e in runs2d:
            p_, f_ = e["sample"]
            e["rel_l2_potential"] = float(np.linalg.norm(p_[i2] - dp2) / np.linalg.norm(dp2))
            e["rel_l2_force"] = float(np.linalg.norm(f_[i2] - df2) / np.linalg.norm(df2))
    for e in runs2d:
        e.pop("sample")
    if batch is not None:
        samples_ = batch.pop("samples")
    if batch is not None and not args.no_error:
        rs_ = np.random.default_rng(args.seed + 22)
        picks_ = []
        for s_ in rs_.choice(512, 8, replace=False):
            a_, b_ = int(offb[s_]), int(offb[s_ + 1])
            ix = rs_.choice(b_ - a_, min(256, b_ - a_), replace=False)
            dp_, df_ = oracle.direct3d(xb[a_:b_], qb[a_:b_], tgt_index=ix, include_self=False)
            picks_.append((a_, b_, ix, dp_, df_))
        for key_, (pbh, fbh) in samples_.items():
            eps_, efs_ = [], []
            for a_, b_, ix, dp_, df_ in picks_:
                phys_ = pbh[a_:b_][ix] - qb[a_:b_][ix] * 1e4  # remove the clamped self term q_i / 1e-4
                eps_.append(float(np.linalg.norm(phys_ - dp_) / np.linalg.norm(dp_)))
                efs_.append(float(np.linalg.norm(fbh[a_:b_][ix] - df_) / np.linalg.norm(df_)))
            dst_ = batch if key_ == "forest" else batch[key_]
            dst_["rel_l2_potential_max_over_8_systems"] = max(eps_)
            dst_["rel_l2_field_max_over_8_systems"] = max(efs_)
    for e in extras:
        r = e.pop("r")
        e.update({"ms_per_step": round(r["ms"], 3), "value": r["value"],
                  "meets_1e-6": bool(e.get("rel_l2_potential", 1.0) <= 1e-6)})
        other.append(e)

    cpu, cpu_all = None, None
    if not args.no_cpu_baseline:
        cpu, cpu_all = cpu_baselines(args, x, q)

    out = {"metric": METRIC, "value": head["value"], "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64+f16x2" if mixed else "f64", "data": "synthetic",
           "config": {"workload": ("3D Laplace FMM, %d particles uniform in [0,1)^3, q~U[-1,1], p=%d (%d terms), "
                                   "uniform octree depth %d (leaves at level %d)" % (args.n, args.terms - 1, args.terms,
                                                                                      args.depth, args.depth - 1))
                      if not adaptive else
                      ("3D Laplace FMM, %d particles in a Plummer sphere (a = 0.05, r <= 1), q~U[-1,1], p=%d "
                       "(%d terms), adaptive octree leaf_max 64" % (args.n, args.terms - 1, args.terms)),
                      "n": args.n, "terms": args.terms, "tree_depth": args.depth, "max_leaf": counts["max_leaf"],
                      "m2l": ("split-FP16 MFMA at levels >= 4, FP64 at levels 2-3" if mixed else "FP64"),
                      "parallelism": "single GPU"},
           "accuracy": accuracy, "roofline": roofline, "hbm": hbm, "stage_ms": stages, "leaf_kernel": leaf,
           "other_configs": other, "plummer": plum,
           "config2d": {"workload": "2D Laplace FMM, 1048576 points uniform in [0,1)^2, q~U[-1,1], p=20 (21 terms), "
                                    "leaf_max 64, self term excluded", "runs": runs2d} if runs2d else None,
           "batch": batch, "cpu_baseline": cpu, "cpu_baselines": cpu_all}
    print(json.dumps(out), flush=True)


def two_d_steps(args, lfmm, torch, dev):
    """Timed 2D solves only (profiling runs): 1,048,576 points uniform in
    the unit square, p = 20 (21 terms), leaf_max 64, FP64 and FP32 M2L sums;
    args.warmup untimed and args.steps timed solves of each."""
    ctx = lfmm.Context(0)
    r2 = np.random.default_rng(args.seed + 31)
    x2, q2 = r2.random((1048576, 2)), r2.uniform(-1.0, 1.0, 1048576)
    x2d, q2d = torch.from_numpy(x2).to(dev), torch.from_numpy(q2).to(dev)
    out = {"workload": "2D Laplace FMM, 1048576 points uniform in [0,1)^2, q~U[-1,1], p=20 (21 terms), leaf_max 64",
           "steps": args.steps, "runs": []}
    for mix in (False, True):
        rep = None
        for _ in range(args.warmup):
            lfmm.fmm2d(x2d, q2d, terms=21, leaf_max=64, exclude_self=True, ctx=ctx, mixed=mix)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rep = lfmm.fmm2d(x2d, q2d, terms=21, leaf_max=64, exclude_self=True, ctx=ctx, mixed=mix)[2]
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
        out["runs"].append({"m2l": "FP32 sums" if mix else "FP64", "ms_per_step": round(ms, 3),
                            "tree_depth": rep["tree_depth"] if rep else None})
        log("2d %s: %.3f ms/step" % ("fp32-m2l" if mix else "fp64", ms))
    ctx.close()
    return out


def batch_steps(args, lfmm, torch, dev):
    """Timed batch solves only (profiling runs): the 512 systems of the
    batch workload as one adaptive octree (FP64, p = 8, leaf_max 32);
    args.warmup untimed and args.steps timed batches."""
    ctx = lfmm.Context(0)
    xb, qb, offb = batch_systems(512, args.seed + 21)
    xbd, qbd = torch.from_numpy(xb).to(dev), torch.from_numpy(qb).to(dev)
    pb = torch.empty(len(qb), dtype=torch.float64, device=dev)
    fb = torch.empty((len(qb), 3), dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        lfmm.fmm3d_batch(xbd, qbd, offb, terms=9, leaf_max=32, ctx=ctx, out=(pb, fb), adaptive=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lfmm.fmm3d_batch(xbd, qbd, offb, terms=9, leaf_max=32, ctx=ctx, out=(pb, fb), adaptive=True)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
    log("batch forest: %.3f ms per batch" % ms)
    ctx.close()
    return {"workload": "512 systems, %d particles, p=8, leaf_max 32, one adaptive octree, FP64" % len(qb),
            "steps": args.steps, "ms_per_batch": round(ms, 3), "particles_per_s": len(qb) / (ms * 1e-3)}


def plummer_leg(args, lfmm, torch, ctx, dev, idx, idx_d, pot, field):
    """4M-particle Plummer sphere (a = 0.05, truncated at r = 1, centred in
    the unit cube; the points outside it are kept, the bounding box grows) on
    the adaptive octree (leaf_max 64): fixed orders in FP64 and mixed, and
    per-box orders from the order network trained here on the GPU on other
    Plummer spheres (labels from this FMM at every order).  Samples are
    taken without the self term (exclude_self)."""
    xp, qp = plummer(args.n, args.seed + 11)
    xpd, qpd = torch.from_numpy(xp).to(dev), torch.from_numpy(qp).to(dev)
    runs = []

    def timed(label, fn, steps=3):
        rep = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            rep = fn()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        log("plummer %s: %.3f ms/step" % (label, ms))
        return ms, rep

    for terms, mix in ((args.terms, False), (args.terms, True), (args.terms - 1, True)):
        ms, rep = timed("adaptive terms=%d %s" % (terms, "mixed" if mix else "fp64"),
                        lambda: lfmm.fmm3d(xpd, qpd, terms=terms, leaf_max=64, adaptive=True, exclude_self=True,
                                           ctx=ctx, out=(pot, field), mixed=mix)[2])
        runs.append({"tree": "adaptive", "orders": "fixed", "terms": terms,
                     "m2l": "split-FP16 MFMA (levels >= 4)" if mix else "FP64",
                     "tree_depth": rep["tree_depth"], "leaves": rep["num_leaves"], "max_leaf": rep["max_leaf"],
                     "ms_per_step": round(ms, 3), "value": args.n / (ms * 1e-3),
                     "sample": (pot[idx_d].cpu().numpy(), field[idx_d].cpu().numpy())})
    # learned per-box orders: training data from three other spheres
    t0 = time.perf_counter()
    sets, label_info = [], []
    for k, n_train in enumerate((1048576, 2097152, args.n)):
        xt, qt = plummer(n_train, args.seed + 101 + k)
        f_, y_, info = lfmm.order.make_labels(lfmm, xt, qt, leaf_max=64, tol=1e-6, safety=0.7, mixed=True, ctx=ctx)
        sets.append((f_, y_))
        label_info.append({"n": n_train, "leaves": info["leaves"], "planned_rel_l2": info["planned_rel_l2"],
                           "label_histogram": np.bincount(y_, minlength=13).tolist()})
    t_labels = time.perf_counter() - t0
    t0 = time.perf_counter()
    model, hist = lfmm.order.train(lfmm, sets, epochs=40, ctx=ctx)
    t_train = time.perf_counter() - t0
    log("order network: labels %.1f s, training %.1f s, loss %.3f -> %.3f" % (t_labels, t_train, hist[0], hist[-1]))
    lfmm.set_order_model(model, confidence=0.5, check_tol=1e-6, check_samples=256, ctx=ctx)
    ms, rep = timed("adaptive learned orders mixed",
                    lambda: lfmm.fmm3d(xpd, qpd, leaf_max=64, learned_order=True, exclude_self=True, ctx=ctx,
                                       out=(pot, field), mixed=True)[2])
    runs.append({"tree": "adaptive", "orders": "learned", "terms": [rep["terms_min"], rep["terms"]],
                 "m2l": "split-FP16 MFMA (levels >= 4)", "tree_depth": rep["tree_depth"], "leaves": rep["num_leaves"],
                 "max_leaf": rep["max_leaf"], "ms_per_step": round(ms, 3), "value": args.n / (ms * 1e-3),
                 "check_rel_l2": rep["check_rel_l2"], "fallback": rep["fallback"],
                 "sample": (pot[idx_d].cpu().numpy(), field[idx_d].cpu().numpy())})
    fl = lfmm.fmm3d(xpd, qpd, leaf_max=64, learned_order=True, exclude_self=True, ctx=ctx, out=(pot, field),
                    mixed=True, count_ops=True)[2]
    fx = lfmm.fmm3d(xpd, qpd, terms=args.terms, leaf_max=64, adaptive=True, exclude_self=True, ctx=ctx,
                    out=(pot, field), mixed=True, count_ops=True)[2]
    runs[-1]["m2l_flops"] = fl["m2l_flops"]
    runs[-1]["m2l_flops_vs_fixed_%d_terms" % args.terms] = round(fl["m2l_flops"] / fx["m2l_flops"], 3)
    lfmm.set_order_model(None, ctx=ctx)
    if not args.no_uniform_plummer:  # the reference's kind of tree: one depth for all leaves
        ms, rep = timed("uniform depth terms=%d fp64" % args.terms,
                        lambda: lfmm.fmm3d(xpd, qpd, terms=args.terms, leaf_max=64, exclude_self=True, ctx=ctx,
                                           out=(pot, field))[2], steps=1)
        runs.append({"tree": "uniform depth", "orders": "fixed", "terms": args.terms, "m2l": "FP64",
                     "tree_depth": rep["tree_depth"], "leaves": rep["num_leaves"], "max_leaf": rep["max_leaf"],
                     "ms_per_step": round(ms, 3), "value": args.n / (ms * 1e-3),
                     "sample": (pot[idx_d].cpu().numpy(), field[idx_d].cpu().numpy())})
    return {"workload": "%d particles, Plummer sphere a = 0.05 truncated at r = 1, centred in the unit cube "
                        "(bounding box grows), q~U[-1,1], adaptive octree leaf_max 64, self term excluded" % args.n,
            "runs": runs, "x": xp, "q": qp,
            "order_network": {"features": 8, "hidden": [64, 64], "classes": 13, "training_sets": label_info,
                              "epochs": len(hist), "loss_first_last": [float(hist[0]), float(hist[-1])],
                              "labels_s": round(t_labels, 2), "training_s": round(t_train, 2),
                              "decision": "smallest order with cumulative probability >= 0.5",
                              "label_rule": "global error budget 0.7e-6 spread by M2L cost (Lagrangian)"}}


def cpu_baselines(args, x, q):
    """The CPU oracle (C/OpenMP restatement of the reference CPU FMM, FP64)
    on all host cores: median of cpu_reps solves after one warm-up, on the
    headline workload and on the 65K 3D and 1M 2D configurations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cores = oracle.num_threads()
    cn = args.cpu_n if args.cpu_n > 0 else args.n
    if cn == args.n:  # the same particles
        xs, qs = x, q
    else:
        rs = np.random.default_rng(args.seed + 7)
        xs = rs.random((cn, 3))
        qs = rs.uniform(-1.0, 1.0, cn)
    cdepth = max(2, args.depth - int(round(np.log(args.n / cn) / np.log(8.0))))
    tc, ts = median_time(lambda: oracle.fmm3d(xs, qs, args.terms, cdepth), args.cpu_reps)
    log("cpu baseline %.1f s (median of %d)" % (tc, len(ts)))
    head = {"value": cn / tc, "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": "CPU oracle (C/OpenMP restatement of the reference CPU FMM, FP64) on %d uniform particles, "
                      "%d terms, depth %d: median of %d solves after a warm-up, %.2f s" %
                      (cn, args.terms, cdepth, len(ts), tc)}
    rng = np.random.default_rng(args.seed + 41)
    x1, q1 = rng.random((65536, 3)), rng.uniform(-1.0, 1.0, 65536)
    t1, _ = median_time(lambda: oracle.fmm3d(x1, q1, 9, 5), args.cpu_reps)
    x2, q2 = rng.random((1048576, 2)), rng.uniform(-1.0, 1.0, 1048576)
    t2, _ = median_time(lambda: oracle.fmm2d(x2, q2, 21, 9), args.cpu_reps)
    log("cpu 65K p=8 %.3f s, 2D 1M p=20 %.3f s" % (t1, t2))
    all_ = {"cores": cores, "protocol": "median of %d solves after one warm-up, all host cores" % args.cpu_reps,
            "3d_65k_p8": {"seconds": round(t1, 4), "particles_per_s": 65536 / t1, "depth": 5},
            "2d_1m_p20": {"seconds": round(t2, 4), "particles_per_s": 1048576 / t2, "depth": 9},
            "3d_4m_p12": {"seconds": round(tc, 4), "particles_per_s": cn / tc, "depth": cdepth}}
    return head, all_


if __name__ == "__main__":
    main()
