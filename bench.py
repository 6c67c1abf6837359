"""Benchmark: FDK of a 512^3 volume from 720 views of 1024^2 on one MI355X.

Workload (the projector-pair configuration at full size): 512^3 float32 volume,
0.5 mm voxels, centred; 720 views over 360 deg, SID 1000 mm, SDD 1536 mm, 1024^2
detector at 0.4 mm, per-view detector offsets jittered +-2 mm from the seed;
projections are the analytic line integrals of a seeded 20-ellipsoid phantom on a
0.02 mm^-1 base.  One step = one full FDK pass (fused cosine weighting + FFT ramp
filter over the stack, then the voxel-driven FDK backprojection of all 720 views)
with the projections already resident in HBM.

value = voxel updates per second of the whole FDK step (GUPS, 512^3 * 720 updates per
step).  roofline = the backprojection kernel (timed with HIP events on the stream it
runs on) against the unit that bounds it, VALU issue: its VALU lane instructions per
voxel update (from the committed rocprofv3 counters of this kernel) times its update
rate, over MI355X's VALU issue peak.  Beside it: the counter-measured HBM traffic per
launch as a fraction of 8 TB/s, and the algorithmic 16 bytes per update (four FP32
taps, served from LDS, so that rate is not an HBM figure).  cpu_baseline = the
oracle's multithreaded FDK (a line-by-line restatement of RTK's CPU filters) on a
bounded sample of the same workload (all voxels, the first views), timed on this
host's cores available to the job.

The second half of the metric is reported beside it (skip with --no-cg):
"cg" times plain CG at 256^3 from 360 noisy views of 512^2 (30 iterations with RTK's
default Joseph FP + voxel BP pair and FP32, FP16 and BF16 stacks, each compared with
the double-precision oracle's 30-iteration result stored in tests/golden/cg_c3.npz;
then time to a relative residual of 1e-3 with that pair and with the matched Joseph
pair); "cg_warm_start" times CG to 1e-3 at 256^3 from 90 noisy views of 512^2 on
held-out phantom seeds, from zero, from FDK and from the committed warm-start
network's CNN(FDK), start costs included; "cg_warm_start_at_scale" does the same at
512^3 from 180 views of 1024^2 (I0 = 5e3, 30 ellipsoids) with the network run as
128^3 tiles with 8-voxel halos (skip with --no-scale).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX, NU, NVIEWS = 512, 1024, 720
SEED = 20260001
# The FDK backprojection kernel on this workload (FP32 stack), from the rocprofv3
# counters of the committed kernel (profiles/bp_fdk_512_720_pmc.md; re-measure when the
# kernel changes): HBM bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors
# from a 1 GiB device copy) and VALU lane instructions per voxel update
# (SQ_INSTS_VALU x 64 / updates).
BP_TRAFFIC_BYTES = 1.554e11
BP_VALU_PER_UPDATE = 12.58
# MI355X VALU issue peak: 256 CUs x 4 SIMDs, one 64-lane instruction per SIMD every
# 4 cycles at 2.4 GHz = 3.93e13 lane instructions per second
VALU_PEAK_LANE_INSTR = 1024 * 2.4e9 / 4 * 64
HBM_PEAK_GBS = 8000.0
# C3 (plain CG) workload: phantom and noise seeds of tests/golden/cg_c3.npz
C3_PHANTOM_SEED, C3_NOISE_SEED = 20260004, 20260005
C3_FIXTURE = os.path.join(ROOT, "tests", "golden", "cg_c3.npz")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(rtk, np):
    rng = np.random.default_rng(SEED)
    offs = rng.uniform(-2., 2., size=(NVIEWS, 2))
    g = rtk.ThreeDCircularProjectionGeometry()
    rows = []
    for i in range(NVIEWS):
        r = [1000., 1536., i * 360. / NVIEWS, offs[i, 0], offs[i, 1], 0., 0., 0., 0.]
        g.AddProjection(*r)
        rows.append(r)
    half_fov = NX * 0.5 / 2
    q = rtk.phantoms.random_ellipsoids(SEED, n=20, half_fov=half_fov)
    det_origin = -(NU - 1) * 0.4 / 2
    proj = rtk.project_quadrics(g, rtk.ConstantImageSource((NU, NU, NVIEWS), (0.4, 0.4, 1.),
                                                          (det_origin, det_origin, 0.)), q)
    vol = rtk.ConstantImageSource((NX, NX, NX), (0.5, 0.5, 0.5), (-(NX - 1) * 0.5 / 2,) * 3)
    return np.array(rows), g, q, proj, vol


def fdk_with_upload(np, torch, fdk, proj, reps=2):
    """The FDK step including the upload of the stack from pinned host memory: what a
    caller that hands over host buffers sees (reported beside `value`, never as it)."""
    host = torch.empty(proj.tensor.shape, dtype=torch.float32, pin_memory=True)
    host.copy_(proj.tensor)
    dev = torch.empty_like(proj.tensor)
    fdk.SetInput(1, proj.like(dev))
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev.copy_(host, non_blocking=True)
        fdk.Update()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    fdk.SetInput(1, proj)
    dt = min(times)
    gb = host.numel() * 4 / 1e9
    return {"ms": round(dt * 1e3, 2), "gups": round(float(NX) ** 3 * NVIEWS / dt / 1e9, 1), "stack_gb": round(gb, 2),
            "note": "stack uploaded from pinned host memory inside the timed region, best of 2"}


def cpu_share():
    """Host cores this job may use: its CPU affinity set, capped by the OMP_NUM_THREADS
    share the host grants the job when it sets one."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(np, rows, proj, vol, target_s=6.0):
    """The oracle's FDK on the first views of the stack, all voxels.  The number of views
    is set from an 8-view calibration run so that one run takes about `target_s`."""
    from oracle import rtkref
    threads = cpu_share()
    ov = rtkref.image((NX,) * 3, vol.origin, vol.spacing)

    def run(views):
        st = rtkref.image((NU, NU, views), proj.origin, proj.spacing)
        p = proj.tensor[:views].cpu().numpy()
        v = np.zeros((NX,) * 3, np.float32)
        t0 = time.perf_counter()
        rtkref.fdk(rows[:views], ov, v, st, p, threads=threads)
        return time.perf_counter() - t0

    t8 = run(8)
    views = int(max(8, min(NVIEWS, round(8 * target_s / max(t8, 1e-3)))))
    times = [run(views) for _ in range(3)]  # median of three runs, warm caches
    dt = float(np.median(times))
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(line.split(":", 1)[1].strip() for line in f if line.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": round(NX ** 3 * views / dt / 1e9, 4), "unit": "GUPS", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model,
            "sample": f"FDK (weight+ramp+BP, float) of the full 512^3 volume from the first {views} of the 720 "
                      f"views, median of 3 runs {dt:.2f} s on {threads} threads (one per core of the job's "
                      f"CPU share) of {model}"}


def projector_pair_measurement(rtk, np, torch, g, q, proj, vol, reps=2):
    """The projector pair alone at the FDK shape: Joseph forward projection of the
    voxelised phantom (512^3) onto the 720 x 1024^2 stack, and the plain voxel-based
    backprojection of the analytic stack into 512^3."""
    phantom = rtk.draw_quadrics(vol, q)
    fp = rtk.HipForwardProjectionImageFilter()
    fp.SetInput(0, proj.like(torch.zeros_like(proj.tensor)))
    fp.SetInput(1, phantom)
    fp.SetGeometry(g)
    fp.Update(count_samples=True)
    samples = fp.GetNumberOfSamples()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fp.Update()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_fp = min(ts)
    del fp
    bp = rtk.HipBackProjectionImageFilter()
    bp.SetInput(0, vol)
    bp.SetInput(1, proj)
    bp.SetGeometry(g)
    bp.Update()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bp.Update()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t_bp = min(ts)
    updates = float(NX) ** 3 * NVIEWS
    out = {"workload": "Joseph FP 512^3 -> 720 x 1024^2 and voxel BP back, same geometry as the FDK",
           "joseph_fp": {"seconds": round(t_fp, 4), "samples": int(samples),
                         "gsamples_per_s": round(samples / t_fp / 1e9, 2),
                         "alg_gb_per_s": round(samples * 16 / t_fp / 1e9, 1)},
           "voxel_bp": {"seconds": round(t_bp, 4), "gups": round(updates / t_bp / 1e9, 1),
                        "alg_gb_per_s": round(updates * 16 / t_bp / 1e9, 1)}}
    log(f"projector pair: {out}")
    return out


def ramp_measurement(rtk, np, torch, proj, reps=5):
    """The ramp filter alone (Hann window 0.8, zero padding to 2048 columns) of the
    720 x 1024^2 stack: the stack in and the filtered stack out.  Kernel time from HIP
    events on the launch stream, best of `reps`."""
    f = rtk.HipFFTRampImageFilter()
    f.SetInput(0, proj)
    f.SetHannCutFrequency(0.8)
    f.Update()
    nbytes = proj.tensor.numel() * 4 * 2
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        f.Update()
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t = min(ms)
    out = {"workload": "ramp filter of the 720 x 1024^2 FP32 stack (Hann 0.8, padded to 2048)",
           "ms": round(t, 3), "bytes": int(nbytes), "gb_per_s": round(nbytes / t / 1e6, 1),
           "frac_of_8tb": round(nbytes / t / 1e6 / 8000., 3)}
    log(f"ramp filter: {out}")
    return out


def weighting_measurement(rtk, np, torch, proj, reps=3):
    """Parker short-scan and displaced-detector weighting of the 720 x 1024^2 stack
    (HBM streams): a 240 degree arc for Parker (stack in, stack out), a detector
    shifted by 60 mm for the displaced-detector filter (stack in, twice as wide out).
    Kernel time from HIP events on the launch stream."""
    out = {"workload": "720 x 1024^2 FP32 stack: Parker weights (240 deg arc); displaced-detector weights "
                       "(+60 mm detector shift, zero-padded to 2048 columns)"}
    n = proj.tensor.shape[0]
    for name, rows, cls in (
            ("parker", [[1000., 1536., i * 240. / n, 0., 0., 0., 0., 0., 0.] for i in range(n)],
             rtk.HipParkerShortScanImageFilter),
            ("displaced_detector", [[1000., 1536., i * 360. / n, 60., 0., 0., 0., 0., 0.] for i in range(n)],
             rtk.HipDisplacedDetectorImageFilter)):
        g = rtk.ThreeDCircularProjectionGeometry()
        for r in rows:
            g.AddProjection(*r)
        f = cls()
        f.SetInput(proj)
        f.SetGeometry(g)
        f.Update()
        nbytes = proj.tensor.numel() * 4 + f.GetOutput().tensor.numel() * 4
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            f.Update()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        t = min(ms)
        out[name] = {"ms": round(t, 3), "bytes": int(nbytes), "gb_per_s": round(nbytes / t / 1e6, 1),
                     "frac_of_8tb": round(nbytes / t / 1e6 / 8000., 3)}
        del f
    # a plain device copy of the stack for scale (what a pure stream reaches here)
    dst = torch.empty_like(proj.tensor)
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(proj.tensor)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    out["device_copy_gb_per_s"] = round(2 * proj.tensor.numel() * 4 / min(ms) / 1e6, 1)
    del dst
    log(f"weighting: {out}")
    return out


def c3_projections(rtk, np, torch, g, nu, n):
    """C3's noisy stack: FP64 analytic line integrals of the 20-ellipsoid phantom (float32),
    Poisson counts at I0 = 1e4 drawn per pixel by inversion of numpy uniforms on the host
    (the draw the double-precision oracle's fixture used; the HIP line integrals equal the
    oracle's to FP32 rounding), converted back to line integrals."""
    q = rtk.phantoms.random_ellipsoids(C3_PHANTOM_SEED, n=20, half_fov=256 * 1.0 / 2)
    det0 = -(nu - 1) * 0.8 / 2
    clean = rtk.project_quadrics(g, rtk.ConstantImageSource((nu, nu, n), (0.8, 0.8, 1.), (det0, det0, 0.)), q)
    p, counts = rtk.phantoms.poisson_line_integrals(clean.numpy(), 1e4, C3_NOISE_SEED, return_counts=True)
    return clean.like(torch.from_numpy(p).cuda()), int(counts.sum())


def c3_error_vs_oracle(np, torch, x, res, fx):
    """A 30-iteration CG result against the double-precision oracle's (tests/golden/cg_c3.npz):
    relative L2 on the two central planes and on the 8^3 block means, and the largest
    relative deviation of the residual history ||R_k|| / ||B||."""
    a = x.double()
    nx = a.shape[0]
    bm = a.reshape(nx // 8, 8, nx // 8, 8, nx // 8, 8).mean(dim=(1, 3, 5)).cpu().numpy()

    def rl2(u, v):
        return float(np.linalg.norm(u - v) / np.linalg.norm(v))
    r = np.asarray(res[:len(fx["residual"])])
    return {"rel_l2_slice_y": rl2(a[:, nx // 2, :].cpu().numpy(), fx["slice_y"]),
            "rel_l2_slice_z": rl2(a[nx // 2].cpu().numpy(), fx["slice_z"]),
            "rel_l2_block_means": rl2(bm, fx["block_means"]),
            "residual_history_max_rel_dev": float(np.max(np.abs(r / fx["residual"][:len(r)] - 1.)))}


def c3_float_oracle_error(np, fx):
    """The float oracle's own 30-iteration result (RTK's CPU filters in float, CG scalars
    in float) against the double one, from the fixture: the FP32 scale of this problem."""
    if "slice_y_f32" not in fx:
        return None

    def rl2(u, v):
        return float(np.linalg.norm(np.asarray(u, np.float64) - v) / np.linalg.norm(v))
    return {"rel_l2_slice_y": rl2(fx["slice_y_f32"], fx["slice_y"]),
            "rel_l2_slice_z": rl2(fx["slice_z_f32"], fx["slice_z"]),
            "rel_l2_block_means": rl2(fx["block_means_f32"], fx["block_means"]),
            "residual_history_max_rel_dev": float(np.max(np.abs(fx["residual_f32"] / fx["residual"] - 1.)))}


def cg_measurement(rtk, np, torch, max_iter, tol):
    """Plain CG on 256^3 @1 mm from 360 views of 512^2 @0.8 mm with Poisson noise
    (I0 = 1e4): 30 iterations from zero with RTK's default pair (Joseph FP + voxel
    BP, FP64 scalars) with FP32, FP16 and BF16 stacks, each against the double-precision
    oracle's 30 iterations on the same input; then time to a relative residual `tol`
    from zero with that pair and with the matched Joseph pair."""
    n, nu, nx = 360, 512, 256
    g = rtk.ThreeDCircularProjectionGeometry()
    for i in range(n):
        g.AddProjection(1000., 1536., i * 360. / n)
    proj, counts_sum = c3_projections(rtk, np, torch, g, nu, n)
    fx = dict(np.load(C3_FIXTURE)) if os.path.exists(C3_FIXTURE) else None
    # the fixture's input: the same counts up to the few pixels a rounding step moves
    # across a count boundary
    matches = fx is not None and abs(int(fx["counts_sum"]) - counts_sum) <= 1e-9 * int(fx["counts_sum"])
    if fx is not None and not matches:
        log(f"cg: Poisson counts sum {counts_sum} differs from the oracle fixture's {int(fx['counts_sum'])}")
    vol = rtk.ConstantImageSource((nx, nx, nx), (1., 1., 1.), (-(nx - 1) / 2.,) * 3)
    out = {}
    runs = [("fixed_30_voxel_bp", 30, 0., None, "f32"), ("fixed_30_voxel_bp_f16_stack", 30, 0., None, "f16"),
            ("fixed_30_voxel_bp_bf16_stack", 30, 0., None, "bf16"), ("to_tol_voxel_bp", max_iter, tol, None, "f32"),
            ("to_tol_joseph_bp", max_iter, tol, "joseph", "f32")]
    ref = None
    for name, niter, t, bp, dtype in runs:
        cg = rtk.HipConjugateGradientConeBeamReconstructionFilter()
        cg.SetInput(0, vol)
        cg.SetInput(1, proj)
        cg.SetGeometry(g)
        if bp == "joseph":
            cg.SetBackProjectionFilter(cg.BP_JOSEPH)
        cg.SetStackDtype(dtype)
        cg.SetNumberOfIterations(niter)
        cg.SetResidualTolerance(t)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cg.Update()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res = cg.GetResiduals()
        it = int(cg.GetNumberOfIterationsDone())
        out[name] = {"seconds": round(dt, 3), "iterations": it, "final_rel_residual": float(res[-1]),
                     "ms_per_iteration": round(dt * 1e3 / max(1, it), 2)}
        if t:
            out[name]["reached"] = bool(res[-1] <= t)
        x = cg.GetOutput().tensor
        if name.startswith("fixed_30"):
            if matches:
                out[name]["vs_fp64_oracle"] = c3_error_vs_oracle(np, torch, x, res, fx)
            if name == "fixed_30_voxel_bp":
                ref = x.double()
            else:
                # the 16-bit stack's own error: relative L2 against the FP32-stack run
                out[name]["rel_l2_vs_f32_stack"] = float(torch.linalg.norm(x.double() - ref) /
                                                         torch.linalg.norm(ref))
        log(f"cg {name}: {out[name]}")
    return {"workload": "CG 256^3 @1 mm <- 360 x 512^2 @0.8 mm, Poisson I0=1e4, zero start; FP32/FP16/BF16 stacks "
                        "read by the backprojector (FP32 arithmetic, FP64 CG scalars)",
            "target_rel_residual": tol, "max_iterations": max_iter,
            "input_matches_oracle_fixture": None if fx is None else bool(matches),
            "float_oracle_vs_fp64_oracle": c3_float_oracle_error(np, fx) if fx is not None else None,
            **out}


def warm_start_measurement(rtk, np, torch, seeds, tol):
    """Learned warm start (sparse view): 256^3 @1 mm from 90 views of 512^2, Poisson
    I0 = 1e4, held-out phantom seeds.  CG (Joseph pair, FP64 scalars) to a relative
    residual `tol` from zero, from FDK and from CNN(FDK) with the committed network;
    each start's own cost is included in its time."""
    from rtk_amd import warmstart_train as W
    from rtk_amd.warmstart import WarmStartCNN, default_weights_path
    net = WarmStartCNN.load(default_weights_path())
    g = W.geometry()
    W.evaluate_seed(net, g, max(seeds) + 1, tol=tol)  # warm-up (first-use allocations), not reported
    rows = []
    for seed in seeds:
        rows.append(W.evaluate_seed(net, g, seed, tol=tol))
        log(f"warm start seed {seed}: " + ", ".join(f"{k} {rows[-1][k]['iterations']} it "
                                                   f"{rows[-1][k]['seconds']:.2f} s" for k in ("zero", "fdk", "cnn")))
    med = {k: float(np.median([r[k]["seconds"] for r in rows])) for k in ("zero", "fdk", "cnn")}
    return {"workload": "CG 256^3 @1 mm <- 90 x 512^2 @0.8 mm, Poisson I0=1e4, held-out seeds",
            "target_rel_residual": tol, "seeds": [int(s) for s in seeds],
            "median_seconds": {k: round(v, 3) for k, v in med.items()},
            "median_iterations": {k: float(np.median([r[k]["iterations"] for r in rows]))
                                  for k in ("zero", "fdk", "cnn")},
            "median_rmse_vs_phantom": {k: float(np.median([r[k]["rmse_vs_phantom"] for r in rows]))
                                       for k in ("zero", "fdk", "cnn")},
            "speedup_cnn_vs_zero": round(med["zero"] / med["cnn"], 3),
            "speedup_fdk_vs_zero": round(med["zero"] / med["fdk"], 3),
            "per_seed": rows}


def warm_start_at_scale_measurement(rtk, np, torch, tol):
    """Learned warm start at scale: 512^3 @0.5 mm from 180 views of 1024^2 @0.4 mm,
    30 ellipsoids, Poisson I0 = 5e3, one held-out seed.  The committed at-scale network
    runs as 128^3 tiles with 8-voxel halos; CG (Joseph pair, FP32 projectors, FP64
    scalars) runs to a relative residual `tol` from zero, FDK and CNN(FDK)."""
    from rtk_amd import warmstart_train as W
    from rtk_amd.warmstart import WarmStartCNN, scale_weights_path
    net = WarmStartCNN.load(scale_weights_path())
    row = W.evaluate_at_scale(net, W.EVAL_SEED0 + 100, tol=tol)
    out = {"workload": "CG 512^3 @0.5 mm <- 180 x 1024^2 @0.4 mm, 30 ellipsoids, Poisson I0=5e3, held-out seed "
                       f"{row['seed']}; CNN as 128^3 tiles with 8-voxel halos",
           "target_rel_residual": tol}
    for k in ("zero", "fdk", "cnn"):
        out[k] = {f: row[k][f] for f in ("seconds", "start_seconds", "iterations", "final_rel_residual",
                                          "reached", "rmse_vs_phantom") if f in row[k]}
        if "rms_diff_vs_zero_start" in row[k]:
            out[k]["rms_diff_vs_zero_start"] = row[k]["rms_diff_vs_zero_start"]
    out["speedup_cnn_vs_zero"] = round(row["zero"]["seconds"] / row["cnn"]["seconds"], 3)
    out["speedup_fdk_vs_zero"] = round(row["zero"]["seconds"] / row["fdk"]["seconds"], 3)
    log(f"warm start at scale: {out}")
    return out


def joseph_pair_measurement(rtk, np, torch, reps=5):
    """The CG operator's projectors at the sparse-view shape: Joseph forward projection
    of a 20-ellipsoid phantom (256^3 @1 mm) onto 90 x 512^2 @0.8 mm, and the Joseph
    backprojection (its exact transpose) of the noisy stack.  Kernel-side time from
    HIP events on the launch stream, best of `reps`."""
    from rtk_amd import warmstart_train as W
    g = W.geometry()
    proj = W.noisy_projections(g, W.EVAL_SEED0)
    vol = W.volume_source()
    x = rtk.draw_quadrics(vol, rtk.phantoms.random_ellipsoids(W.EVAL_SEED0, n=20, half_fov=100.))
    fp = rtk.HipForwardProjectionImageFilter()
    fp.SetInput(0, proj.like(torch.zeros_like(proj.tensor)))
    fp.SetInput(1, x)
    fp.SetGeometry(g)
    fp.Update(count_samples=True)
    samples = fp.GetNumberOfSamples()
    bp = rtk.HipJosephBackProjectionImageFilter()
    bp.SetInput(0, vol)
    bp.SetInput(1, proj)
    bp.SetGeometry(g)
    bp.Update()
    out = {"workload": "Joseph FP 256^3 @1 mm -> 90 x 512^2 @0.8 mm and Joseph BP back (the CG operator's pair)",
           "samples": int(samples)}
    for name, f in (("joseph_fp", fp), ("joseph_bp", bp)):
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            f.Update()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        t = min(ms)
        out[name] = {"ms": round(t, 3), "gsamples_per_s": round(samples / t / 1e6, 1),
                     "alg_gb_per_s": round(samples * 16 / t / 1e6, 1)}
    log(f"joseph pair: {out}")
    return out


def _warm_seeds(args):
    from rtk_amd import warmstart_train as W
    return [W.EVAL_SEED0 + i for i in range(args.warm_seeds)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stack-dtype", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--cpu-seconds", type=float, default=6.0,
                    help="target length of one CPU-baseline run (the view count is calibrated to it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG and warm-start measurements")
    ap.add_argument("--no-pair", action="store_true", help="skip the Joseph FP / voxel BP measurement")
    ap.add_argument("--no-weights", action="store_true", help="skip the Parker / displaced-detector measurement")
    ap.add_argument("--no-scale", action="store_true", help="skip the warm start at 512^3 / 180 views")
    ap.add_argument("--cg-tol", type=float, default=1e-3)
    ap.add_argument("--cg-iter", type=int, default=100)
    ap.add_argument("--warm-seeds", type=int, default=16, help="held-out seeds for the warm-start measurement")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--cg-only", action="store_true", help="skip the FDK measurement")
    ap.add_argument("--joseph-only", action="store_true",
                    help="only the Joseph FP / BP pair at the sparse-view CG shape")
    ap.add_argument("--ramp-only", action="store_true", help="only the ramp filter of the FDK stack")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: rtkhip has no multi-GPU path "
              "(the single-GPU kernels run at the per-GPU share's full shape).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import rtk_amd as rtk

    rtk.lib()
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.calib_copy_mib:
        a = torch.ones(args.calib_copy_mib * 262144, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy of {args.calib_copy_mib} MiB done")
        del a, b
    if args.joseph_only:
        print(json.dumps({"joseph_pair_sparse_view": joseph_pair_measurement(rtk, np, torch)}), flush=True)
        return
    if args.cg_only:
        res = {"cg": cg_measurement(rtk, np, torch, args.cg_iter, args.cg_tol),
               "cg_warm_start": warm_start_measurement(rtk, np, torch, _warm_seeds(args), args.cg_tol)}
        if not args.no_scale:
            res["cg_warm_start_at_scale"] = warm_start_at_scale_measurement(rtk, np, torch, args.cg_tol)
        print(json.dumps(res), flush=True)
        return
    log("building workload (analytic projections of the 20-ellipsoid phantom)")
    rows, g, q, proj, vol = build_workload(rtk, np)
    torch.cuda.synchronize()
    if args.ramp_only:
        print(json.dumps({"ramp_filter": ramp_measurement(rtk, np, torch, proj)}), flush=True)
        return

    fdk = rtk.HipFDKConeBeamReconstructionFilter()
    fdk.SetInput(0, vol)
    fdk.SetInput(1, proj)
    fdk.SetGeometry(g)
    fdk.SetStackDtype(args.stack_dtype)
    for w in range(args.warmup):
        fdk.Update()
        log(f"warmup {w + 1}/{args.warmup}: BP {fdk.GetBackProjectionTimeMs():.1f} ms")
    torch.cuda.synchronize()
    bp_ms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        fdk.Update()
        bp_ms.append(fdk.GetBackProjectionTimeMs())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    for s, b in enumerate(bp_ms):
        log(f"step {s + 1}: BP kernel {b:.1f} ms")
    out = fdk.GetOutput().tensor
    if not torch.isfinite(out).all():
        raise RuntimeError("non-finite FDK output")
    with_upload = fdk_with_upload(np, torch, fdk, proj)
    log(f"FDK step with the stack upload: {with_upload}")

    updates = float(NX) ** 3 * NVIEWS
    ms_step = dt / args.steps * 1e3
    bp_avg = float(np.mean(bp_ms))
    bp_s = bp_avg * 1e-3
    bytes_per_update = 16 if args.stack_dtype == "f32" else 8
    f32 = args.stack_dtype == "f32"
    # VALU issue: lane instructions per update are the committed profile's (FP32 stack)
    lane_rate = updates * BP_VALU_PER_UPDATE / bp_s if f32 else None
    hbm_gbs = BP_TRAFFIC_BYTES / bp_s / 1e9 if f32 else None
    roofline = {"bound": "valu",
                "achieved": round(lane_rate / 1e12, 2) if f32 else None,
                "peak": round(VALU_PEAK_LANE_INSTR / 1e12, 2),
                "unit": "T VALU lane-instructions/s",
                "frac": round(lane_rate / VALU_PEAK_LANE_INSTR, 4) if f32 else None,
                "traffic": BP_TRAFFIC_BYTES if f32 else None,
                "kernel": "bp_lds_kernel (FDK backprojection, all 720 views in one launch)",
                "per_unit": f"{BP_VALU_PER_UPDATE} VALU lane instructions per voxel update (SQ_INSTS_VALU x 64 / "
                            "updates of this kernel)",
                "hbm": {"achieved": round(hbm_gbs, 1) if f32 else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(hbm_gbs / HBM_PEAK_GBS, 4) if f32 else None,
                        "note": "counter-measured HBM bytes per launch (FETCH_SIZE x 2 + WRITE_SIZE) / kernel time"},
                "algorithmic_bytes": {"per_update": bytes_per_update,
                                      "gb_per_s": round(updates * bytes_per_update / bp_s / 1e9, 1),
                                      "note": "four taps per update, read from LDS tiles staged once per view and "
                                              "workgroup: an effective rate, not HBM traffic"}}
    result = {
        "metric": "FDK GUPS at 512\u00b3; CG time-to-residual 1e-3, learned warm start vs plain",
        "value": round(updates * args.steps / dt / 1e9, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if f32 else f"f32 math, {args.stack_dtype} filtered stack",
        "data": "synthetic: analytic FP64 projections of a seeded 20-ellipsoid phantom",
        "config": {"workload": "FDK 512^3 @0.5 mm <- 720 x 1024^2 @0.4 mm, SID 1000, SDD 1536, +-2 mm offsets",
                   "volume": [NX, NX, NX], "views": NVIEWS, "detector": [NU, NU],
                   "bp_kernel_gups": round(updates / bp_s / 1e9, 3),
                   "bp_kernel_ms": round(bp_avg, 3)},
        "fdk_with_stack_upload": with_upload,
        "roofline": roofline,
    }
    if not args.no_pair:
        result["projector_pair"] = projector_pair_measurement(rtk, np, torch, g, q, proj, vol)
    if not args.no_weights:
        result["ramp_filter"] = ramp_measurement(rtk, np, torch, proj)
        result["short_scan_weighting"] = weighting_measurement(rtk, np, torch, proj)
    if not args.no_cg:
        result["joseph_pair_sparse_view"] = joseph_pair_measurement(rtk, np, torch)
        result["cg"] = cg_measurement(rtk, np, torch, args.cg_iter, args.cg_tol)
        result["cg_warm_start"] = warm_start_measurement(rtk, np, torch, _warm_seeds(args), args.cg_tol)
        if not args.no_scale:
            result["cg_warm_start_at_scale"] = warm_start_at_scale_measurement(rtk, np, torch, args.cg_tol)
    if not args.no_cpu:
        log(f"cpu baseline: oracle FDK on {cpu_share()} threads")
        result["cpu_baseline"] = cpu_baseline(np, rows, proj, vol, args.cpu_seconds)
    # release the device before printing, so the process ends right after its line
    del fdk, out, proj, vol
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
