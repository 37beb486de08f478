"""Benchmark: poses scored per second, including the per-pose c1/c2 fit, of
the FFT dimer SAXS scoring path on one MI355X, in both forms: FP64
throughout, and mixed precision (FP32 rotated tables and cross-term
contraction, FP64 Bessel / Wigner-d / reductions / fit).  The headline value
is the mixed-precision form; the FP64 form and the ratio are reported beside.

Workload (one step = one scoring pass over the whole pose set, inputs
resident in HBM): synthetic receptor of 3,000 atoms and ligand of 1,500
(compact random-walk chains, C/N/O/S form factors, Shrake-Rupley SASA),
L = 15, 51 q on [0, 0.5] 1/A, z = 25..60 A in 1 A steps (36), beta_rec /
beta_lig in 10 degree steps (19 points), 64-point alpha_lig / gamma grids,
210,000 distinct grid poses, c1 in [0.95, 1.05], c2 in [0, 4].  The
'experimental' curve is the first pose's I(q) at c1 = c2 = 1 plus 1 %
Gaussian noise (sigma = 5 % I).

Prints one JSON line with the throughput, the roofline position of the
dominant kernel and a CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# The bench workloads (synthetic dimers; the default is the one the headline
# value is quoted on).
WORKLOADS = {
    "standard": dict(npose=210000, L=15, qnum=51, qmax=0.5, zlo=25, zhi=60, zstep=1.0,
                     nbeta=19, ngrid=64, n_rec=3000, n_lig=1500),
    # large single complex: L = 25, 81 q on [0, 0.6], z in 0.5 A steps over
    # 20..70 A, 5 degree betas, 128-point alpha/gamma grids, 5e6 poses
    "large": dict(npose=5000000, L=25, qnum=81, qmax=0.6, zlo=20, zhi=70, zstep=0.5,
                  nbeta=37, ngrid=128, n_rec=10000, n_lig=4000),
    # the small CPU-runnable case: L = 10, 31 q on [0, 0.3], z 20..50 A, the
    # native grid of L = 10, 2000 poses
    "small": dict(npose=2000, L=10, qnum=31, qmax=0.3, zlo=20, zhi=50, zstep=1.0,
                  nbeta=11, ngrid=21, n_rec=1200, n_lig=800),
}


def build_workload(seed=0, npose=210000, L=15, qnum=51, qmax=0.5, zlo=25, zhi=60, zstep=1.0,
                   nbeta=19, ngrid=64, n_rec=3000, n_lig=1500, c1=(0.95, 1.05), c2=(0.0, 4.0),
                   water=True):
    import libfmftsaxs_amd as F
    from libfmftsaxs_amd import synthetic as S
    t0 = time.time()
    rec, lig = S.make_dimer_inputs(n_rec, n_lig, 1001 + seed, 2001 + seed, water=water)
    q = F.mkarray(0.0, qmax, qnum)
    Ar, Ai, rmA = F.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
    Br, Bi, rmB = F.atoms2spf(lig.coords, lig.vac, lig.dum, lig.sa, lig.vdw, q, L)
    zv = np.arange(float(zlo), float(zhi) + 0.5 * zstep, zstep)
    poses = S.random_grid_poses(npose, zv.size, nbeta, ngrid, ngrid, 3001 + seed)
    rm = (rmA * rec.natoms + rmB * lig.natoms) / (rec.natoms + lig.natoms)
    # experimental curve from the first pose
    probe = F.ScoringProblem(L, q, zv, Ar, Ai, Br, Bi, np.ones(6 * qnum), 1.0, 1.0,
                             nbeta=nbeta, nalpha=ngrid, ngamma=ngrid, c1=c1, c2=c2)
    ct = F.score(probe, poses[:1], want_cross=True)["cross"][0]
    I = ct[0] - ct[1] + ct[2] + ct[3] - ct[4] + ct[5]
    rng = np.random.default_rng(4001 + seed)
    Iexp = I * (1 + 0.01 * rng.normal(size=qnum))
    a, mult, peak = F.fit_params(q, Iexp, 0.05 * Iexp, q, rm)
    pb = F.ScoringProblem(L, q, zv, Ar, Ai, Br, Bi, a, mult, peak, nbeta=nbeta, nalpha=ngrid,
                          ngamma=ngrid, c1=c1, c2=c2)
    log(f"workload built in {time.time() - t0:.1f} s")
    return dict(problem=pb, poses=poses, q=q, zv=zv, L=L, A=(Ar, Ai), B=(Br, Bi), rm=rm,
                a=a, mult=mult, peak=peak, exp=(q, Iexp, 0.05 * Iexp), nbeta=nbeta,
                ngrid=ngrid, c1=c1, c2=c2)


def run_batch(args):
    """The batched form: NCOMPLEX synthetic complexes (receptor atoms
    log-uniform over 500..6000, ligand 300..3000, L from {10, 15, 20}, poses
    log-uniform over 1e4..1e6; 51 q on [0, 0.5], z 25..60 A, 10 degree betas,
    64-point alpha/gamma grids; water shell from Shrake-Rupley SASA), scored
    by one fsx_score_batch call per step.  The first call of each form tunes
    the cross-term tile rows per expansion-order bucket and captures the
    batch into a HIP graph; the timed steps replay it."""
    import torch
    import libfmftsaxs_amd as F
    rng = np.random.default_rng(5000)
    ws = []
    for i in range(args.complexes):
        n_rec = int(np.exp(rng.uniform(np.log(500), np.log(6000))))
        n_lig = int(np.exp(rng.uniform(np.log(300), np.log(3000))))
        L = int(rng.choice([10, 15, 20]))
        npose = int(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        ws.append(build_workload(seed=100 + i, npose=npose, L=L, n_rec=n_rec, n_lig=n_lig,
                                 water=True))
        log(f"complex {i}: {n_rec} / {n_lig} atoms, L={L}, {npose} poses")
    dev = torch.device("cuda:0")
    pbs = [w["problem"] for w in ws]
    ns = [len(w["poses"]) for w in ws]
    d_poses = [torch.from_numpy(w["poses"]).to(dev) for w in ws]
    outs = [[torch.full((n,), float("nan"), dtype=torch.float64, device=dev) for _ in range(3)]
            for n in ns]
    torch.cuda.synchronize()

    def measure(mixed):
        def step():
            return F.score_batch_device(pbs, [p.data_ptr() for p in d_poses], ns,
                                        [o[0].data_ptr() for o in outs],
                                        [o[1].data_ptr() for o in outs],
                                        [o[2].data_ptr() for o in outs], mixed=mixed)
        tag = "mixed" if mixed else "fp64"
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        first = (time.perf_counter() - t0) * 1e3
        log(f"{tag} first call (tuning, capture): {first:.1f} ms")
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [step() for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        assert all(int(torch.isnan(o[0]).sum()) == 0 for o in outs)
        last = stats[-1]
        assert sum(s["n_scored"] for s in last) == sum(ns)
        evals = sum(s["fit_evals"] for s in last) / sum(ns)
        log(f"{tag}: {dt * 1e3 / args.steps:.1f} ms per step")
        return {"value": sum(ns) / (dt / args.steps), "ms_per_step": dt * 1e3 / args.steps,
                "first_call_ms": first, "evaluations_per_fit": evals,
                "cross_flops": sum(s["cross_flops"] for s in last)}

    fp64 = measure(False)
    mixed = measure(True)
    print(json.dumps({
        "metric": "poses scored/sec incl. c1,c2 fit: mixed-precision vs FP64 GPU form, 1 GPU",
        "value": mixed["value"], "unit": "poses/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": mixed["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 contraction, f64 fit",
        "data": "synthetic",
        "config": {"workload": f"{len(ns)} complexes (receptor 500..6000, ligand 300..3000 "
                               "atoms log-uniform, L from {10, 15, 20}, 1e4..1e6 poses "
                               "log-uniform; 51 q on [0,0.5], z 25..60, 19 betas, 64-point "
                               "alpha/gamma grids, water shell) in one fsx_score_batch call "
                               "(one graph launch)",
                   "name": "batch", "poses": sum(ns),
                   "L": [int(w["L"]) for w in ws]},
        "roofline": None,
        "roofline_note": "stage times are not measured inside the batch graph; the per-kernel "
                         "durations are in the rocprofv3 summary of this workload",
        "fp64_form": fp64, "mixed_form": mixed,
        "mixed_over_fp64": mixed["value"] / fp64["value"], "cpu_baseline": None}))


def run_atoms(args):
    """The atom expansion (fsx_atoms2spf, pdb2spf.c:24-152) of the large
    workload's receptor (10,000 atoms with water shell, L = 25, 81 q on
    [0, 0.6]) in both forms: FP64 MFMA sums and the split-BF16 MFMA sums of
    the mixed form.  Host arrays in and out (the call uploads the atoms and
    downloads the amplitudes); reported per call, with the largest difference
    between the forms relative to each (X, q) block's largest amplitude."""
    import libfmftsaxs_amd as F
    from libfmftsaxs_amd import synthetic as S
    wk = WORKLOADS["large"]
    rec, _ = S.make_dimer_inputs(wk["n_rec"], 300, 1001, 2001, water=True)
    q = F.mkarray(0.0, wk["qmax"], wk["qnum"])
    L = wk["L"]
    res = {}
    for mixed in (False, True):
        for _ in range(args.warmup):
            F.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L, mixed=mixed)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = F.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L, mixed=mixed)
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        res["mixed" if mixed else "fp64"] = (ms, out)
        log(f"{'mixed' if mixed else 'fp64'} atom expansion: {ms:.2f} ms per call")
    (r64, i64, _), (r16, i16, _) = res["fp64"][1], res["mixed"][1]
    scale = np.maximum(np.abs(r64).max(axis=2, keepdims=True), np.abs(i64).max(axis=2, keepdims=True))
    err = float((np.maximum(np.abs(r16 - r64), np.abs(i16 - i64)) / scale).max())
    print(json.dumps({"metric": "atom expansion (fsx_atoms2spf) per call, host arrays in and out",
                      "config": {"workload": f"{rec.natoms} atoms with water shell, L = {L}, "
                                             f"{wk['qnum']} q on [0, {wk['qmax']}]"},
                      "steps": args.steps, "warmup": args.warmup,
                      "fp64_ms": res["fp64"][0], "mixed_ms": res["mixed"][0],
                      "mixed_max_rel_error": err}))


def _oracle_rate(w, sample, nthr):
    import oracle as O
    O.set_pi(O.PI_TRUNC)
    L, nb, ng = w["L"], w["nbeta"], w["ngrid"]
    P = O.Params(w["a"], w["rm"], w["peak"], c1=w["c1"], c2=w["c2"])
    P.s.mult = w["mult"]
    t = time.time()
    r = O.compute_saxs_scores(sample, w["A"][0], w["A"][1], w["B"][0], w["B"][1], P, w["q"],
                              w["zv"], L, nbeta=nb, na=ng, ng=ng, nthreads=nthr)
    dt = time.time() - t
    assert not np.any(np.isnan(r["score"]))
    return len(sample) / dt, dt


def cpu_baseline(w, small=None, max_threads=16):
    """The oracle (FP64 CPU restatement of the reference algorithm: per-slab
    second sums, transform to the requested grid points, one L-BFGS-B fit
    per pose, OpenMP over slabs) on bounded samples of the workloads, on
    this host's cores: all cores (at most max_threads) on the poses of the
    z = 40 A plane, one core on the part of that plane with beta_rec = 0,
    and all cores on every pose of the small CPU-runnable workload."""
    L, nb, ng = w["L"], w["nbeta"], w["ngrid"]
    poses = w["poses"]
    size5d = nb * nb * ng ** 3
    zsel = int(np.argmin(np.abs(w["zv"] - 40.0)))
    sample = poses[poses // size5d == zsel]
    nthr = max(1, min(max_threads, os.cpu_count() or 1))
    rate, dt = _oracle_rate(w, sample, nthr)
    out = {"value": rate, "unit": "poses/s", "cores": nthr, "kind": "port",
           "sample": f"{len(sample)} poses of this workload (all of the z = 40 A plane, "
                     f"{nb * nb} slabs), oracle with OpenMP over slabs on {nthr} threads, "
                     f"{dt:.1f} s", "cpu_model": cpu_model()}
    b1 = (sample % size5d) // (nb * ng ** 3)
    one = sample[b1 == 0]
    rate1, dt1 = _oracle_rate(w, one, 1)
    out["single_core"] = {"value": rate1, "unit": "poses/s", "cores": 1,
                          "sample": f"{len(one)} poses (the z = 40 A plane, beta_rec = 0), "
                                    f"{dt1:.1f} s"}
    if small is not None:
        rs, dts = _oracle_rate(small, small["poses"], nthr)
        out["small_workload_full"] = {"value": rs, "unit": "poses/s", "cores": nthr,
                                      "sample": f"all {len(small['poses'])} poses of the small "
                                                f"CPU-runnable workload (L = {small['L']}), "
                                                f"{dts:.1f} s"}
    return out


def w_alg(poses, L, Q, nb, ng):
    """W_alg of the reference's slab algorithm for these poses (compute_sum1
    per active (z, beta2), compute_sum2 per active (z, beta1, beta2) slab;
    fftsaxs.c:768-799, 416-524)."""
    C1 = sum(1 for m in range(-L, L + 1) for m2 in range(-L, L + 1)
             for l in range(abs(m), L + 1) for l1 in range(max(abs(m), abs(m2)), L + 1))
    C2 = sum(L + 1 - max(abs(m), abs(m1)) for m in range(-L, L + 1)
             for m1 in range(-L, L + 1)) * (2 * L + 1)
    size5d = nb * nb * ng ** 3
    z = poses // size5d
    r = poses % size5d
    b1 = r // (nb * ng ** 3)
    b2 = (r // ng ** 3) % nb
    zb2 = np.unique(z * nb + b2).size
    slabs = np.unique((z * nb + b1) * nb + b2).size
    return 24.0 * Q * C1 * zb2 + 72.0 * Q * C2 * slabs


def load_traffic(workload):
    """HBM bytes per launch of the main kernels on this workload, from the
    committed rocprofv3 --pmc summary of the same workload (FETCH_SIZE and
    WRITE_SIZE scaled by what they read for a plain copy of known size); {}
    when no summary of this workload is committed."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("name") != workload:
            return {}
        return {k: v["bytes_per_launch"] for k, v in d["kernels"].items()}
    except (OSError, KeyError, ValueError):
        return {}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--poses", type=int, default=0, help="pose count (default: the workload's)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["batch", "atoms"], default="standard",
                    help="standard (the headline workload), large (L = 25 single complex), "
                         "small (the CPU-runnable case), batch (many complexes in one "
                         "call) or atoms (the atom expansion in both forms)")
    ap.add_argument("--complexes", type=int, default=64, help="complexes of the batch workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed", action="store_true",
                    help="time only the mixed-precision form (FP32 contraction, FP64 fit)")
    ap.add_argument("--fp64", action="store_true", help="time only the FP64 form")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device three times (a plain copy "
                         "kernel of known size, to calibrate HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build targets single-GPU boxes "
              "(the path shards by z / (z, beta2) buckets with no exchange step)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        nb = args.copy_probe << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nb, "reads_per_copy": nb, "writes_per_copy": nb}))
        return
    import libfmftsaxs_amd as F
    if F.device_count() < 1:
        raise SystemExit("no HIP device")
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "atoms":
        run_atoms(args)
        return
    wk = dict(WORKLOADS[args.workload])
    if args.poses:
        wk["npose"] = args.poses
    w = build_workload(**wk)
    pb = w["problem"]
    n = len(w["poses"])
    dev = torch.device("cuda:0")
    d_poses = torch.from_numpy(w["poses"]).to(dev)
    d_score = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    d_c1 = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    d_c2 = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def measure(mixed):
        """a cold first call, warmup and timed steps of one form; returns its
        record"""
        def step():
            return F.score_device(pb, d_poses.data_ptr(), n, d_score.data_ptr(),
                                  d_c1.data_ptr(), d_c2.data_ptr(), mixed=mixed)
        tag = "mixed" if mixed else "fp64"
        # cold: device buffers and host tables (Wigner-d, d-symbols) released,
        # so the call builds, allocates and uploads them again
        F.release()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        cold_ms = (time.perf_counter() - t0) * 1e3
        log(f"{tag} cold first call: {cold_ms:.1f} ms")
        for i in range(args.warmup):
            st = step()
            log(f"{tag} warmup {i}: {st['ms_total']:.1f} ms")
        torch.cuda.synchronize()
        stats = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            stats.append(step())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        assert int(torch.isnan(d_score).sum()) == 0
        keys = ["ms_tables", "ms_bessel_t", "ms_rotate", "ms_group", "ms_cross", "ms_fit_prep",
                "ms_fit"]
        avg = {k: float(np.mean([s[k] for s in stats])) for k in keys}
        log(f"{tag} per-stage ms: " + ", ".join(f"{k[3:]}={v:.3f}" for k, v in avg.items()))
        cb = float(np.mean([s["cross_bytes"] for s in stats]))
        cf = float(np.mean([s["cross_flops"] for s in stats]))
        # Roofline of the two main kernels; the stage times come from HIP
        # events the library records on the stream its kernels run on.  The
        # cross-term kernel is vector (VALU) arithmetic on operands held in
        # LDS and registers: its roof is the VALU peak of its type, 78.6
        # TFLOP/s for FP64 and for FP32 as built here (packed FP32 is
        # disabled, so FP32 issues at the FP64 rate).  The fit stream is
        # priced against HBM (its profile read once, three results written).
        peak = 78.6
        achieved = cf / (avg["ms_cross"] * 1e-3) / 1e12
        tr = traffic.get("k_cross_" + ("f32" if mixed else "f64"))
        roof_cross = {"kernel": "k_cross", "bound": "valu", "achieved": achieved, "peak": peak,
                      "peak_basis": "VALU FMA rate, FP64 and non-packed FP32",
                      "unit": "TFLOP/s", "frac": achieved / peak, "traffic": tr,
                      "ms_per_launch": avg["ms_cross"], "algorithmic_flops": cf,
                      "algorithmic_bytes": cb, "arith": "f32" if mixed else "f64"}
        if tr:
            # the measured HBM bytes per launch as a rate.  The algorithmic
            # bytes (one receptor row per pose and q) overstate what must come
            # from HBM wherever L2 serves repeated rows (L = 15: 40.6 GB asked,
            # 15.7 GB fetched), so HBM is taken as the binding roof only when
            # the measured rate is clearly nearer its peak than the flops are
            # to theirs (receptor rows that no longer fit L2, e.g. L = 25), and
            # then priced with the smaller of the two byte counts: bytes the
            # caches served never count as HBM throughput
            rate = tr / (avg["ms_cross"] * 1e-3) / 1e9
            roof_cross["traffic_gbs"] = rate
            roof_cross["traffic_frac"] = rate / 8000.0
            if rate / 8000.0 > 1.25 * roof_cross["frac"]:
                ab = min(cb, tr) / (avg["ms_cross"] * 1e-3) / 1e9
                roof_cross.update({"bound": "hbm", "achieved": ab, "peak": 8000.0,
                                   "peak_basis": "HBM3E", "unit": "GB/s", "frac": ab / 8000.0,
                                   "valu_tflops": achieved, "valu_frac": achieved / peak})
        # each pose's six rescaled profile terms read once (FP32 in the mixed
        # form's lane-group fit, FP64 otherwise) and score, c1, c2 written
        fb = n * (6 * pb.qnum * (4 if mixed else 8) + 24)
        achieved = fb / (avg["ms_fit"] * 1e-3) / 1e9
        roof_fit = {"kernel": "k_fit", "bound": "hbm", "achieved": achieved, "peak": 8000.0,
                    "unit": "GB/s", "frac": achieved / 8000.0,
                    "traffic": traffic.get("k_fit_" + ("f32" if mixed else "f64")),
                    "ms_per_launch": avg["ms_fit"], "algorithmic_bytes": fb,
                    "evaluations_per_fit": float(np.mean([s["fit_evals"] for s in stats]))
                    / max(1, stats[-1]["n_scored"]),
                    # loop iterations of the fit kernel: summed over its wavefronts, of
                    # those wavefront-wide (the last few fits of a wavefront evaluated
                    # together), and the longest-running wavefront's
                    "wavefront_steps": float(np.mean([s["fit_steps"] for s in stats])),
                    "wide_steps": float(np.mean([s["fit_wide_steps"] for s in stats])),
                    "longest_wavefront_steps": float(np.mean([s["fit_max_steps"] for s in stats]))}
        # the longer of the two is the dominant kernel; the other is reported beside
        roof, other = ((roof_cross, roof_fit) if avg["ms_cross"] >= avg["ms_fit"]
                       else (roof_fit, roof_cross))
        ms = dt * 1e3 / args.steps
        return {"value": n / (dt / args.steps), "ms_per_step": ms, "stage_ms": avg,
                "cold_first_call_ms": cold_ms, "roofline": roof, "roofline_other": other}

    traffic = load_traffic(args.workload)
    walg = w_alg(w["poses"], w["L"], int(pb.qnum), w["nbeta"], w["ngrid"])
    fp64 = None if args.mixed else measure(False)
    mixed = None if args.fp64 else measure(True)
    head = mixed or fp64
    cpu = None
    if args.workload == "large":
        log("no CPU baseline for the large workload (one z plane is hours of CPU work)")
    elif not args.no_cpu_baseline:
        try:
            small = build_workload(**WORKLOADS["small"]) if args.workload == "standard" else None
            cpu = cpu_baseline(w, small)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed:", repr(e))
    out = {"metric": "poses scored/sec incl. c1,c2 fit: mixed-precision vs FP64 GPU form, "
                     "1 GPU",
           "value": head["value"], "unit": "poses/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 contraction, f64 fit" if mixed else "f64", "data": "synthetic",
           "config": {"workload": f"receptor {wk['n_rec']} / ligand {wk['n_lig']} atoms, "
                                  f"L={wk['L']}, {wk['qnum']} q on [0,{wk['qmax']}], z "
                                  f"{wk['zlo']}..{wk['zhi']} step {wk['zstep']}, "
                                  f"{wk['nbeta']} betas ({180 / (wk['nbeta'] - 1):g} deg), "
                                  f"{wk['ngrid']}-point alpha/gamma grids, {n} poses, "
                                  "c1 [0.95,1.05], c2 [0,4]",
                      "name": args.workload,
                      "form": "mixed-precision" if mixed else "fp64",
                      "poses": n, "L": w["L"], "qnum": int(pb.qnum), "nbeta": w["nbeta"],
                      "ngrid": w["ngrid"], "znum": int(w["zv"].size)},
           "stage_ms": head["stage_ms"], "cold_first_call_ms": head["cold_first_call_ms"],
           "roofline": head["roofline"], "roofline_other": head["roofline_other"],
           # the reference's slab algorithm's flops for the same poses (per
           # active (z, beta2) 24 Q C1(L) flops of first sums, per active
           # (z, beta1, beta2) slab 72 Q C2(L) of second sums) over this
           # step's time: a rate of the reference algorithm's work, not a
           # roofline figure (this path runs a different, cheaper algorithm)
           "reference_algorithm": {"w_alg_flops": walg,
                                   "effective_tflops": walg / (head["ms_per_step"] * 1e-3) / 1e12},
           "fp64_form": fp64, "mixed_over_fp64": (mixed["value"] / fp64["value"])
           if (mixed and fp64) else None,
           "cpu_baseline": cpu}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
