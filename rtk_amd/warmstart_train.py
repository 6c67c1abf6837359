"""Training and evaluation of the learned CG warm start on GPU-generated pairs.

Sparse-view configuration (256^3 @1 mm, 90 views of 512^2 @0.8 mm, SID 1000 mm,
SDD 1536 mm, Poisson noise I0 = 1e4): for each seed, the random-ellipsoid phantom is
projected analytically, noise is applied, and the pair (FDK volume, volume after 40
CG iterations from zero) is produced by this package's own HIP filters.  Forty
iterations take the relative residual to about 7e-4, below the 1e-3 tolerance the
start is evaluated at.  Unregularised CG keeps fitting noise after that: a network asked
for a later iterate starts CG from a noisier volume (its end point drifts from the
phantom), and one asked for the 100-iteration volume must guess noise it cannot see
(a smaller gain).  `--cg-iter` selects the iterate.  CG uses the
Joseph forward projector with its matched Joseph backprojector, so the normal
operator BP W FP is symmetric (RTK's CG with the voxel-based backprojector is not,
and stalls at this view count).  The CNN is trained on random crops with Adam and
evaluated on held-out seeds as CG time to a relative residual ||b - A x|| / ||b||
of 1e-3 from three starts: zero, FDK, CNN(FDK); the start's own cost (FDK, CNN) is
included in its time.

Usage: python -m rtk_amd.warmstart_train --train 500 --steps 8000 --eval 16 --out w.npz
       python -m rtk_amd.warmstart_train --scale-eval 1   (512^3 / 180 views, tiled CNN)
       python -m rtk_amd.warmstart_train --scale-train 32 --steps 4000 --scale-eval 1 --out w5.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

import rtk_amd as rtk
from rtk_amd.warmstart import WarmStartCNN

NX, NU, NVIEWS, I0 = 256, 512, 90, 1e4
TRAIN_SEED0, EVAL_SEED0 = 1000, 5000
SELECT_SEED0 = 7000  # checkpoint selection (--select), disjoint from training and evaluation seeds


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def geometry(views=NVIEWS):
    g = rtk.ThreeDCircularProjectionGeometry()
    for i in range(views):
        g.AddProjection(1000., 1536., i * 360. / views)
    return g


def volume_source(nx=NX):
    return rtk.ConstantImageSource((nx, nx, nx), (1., 1., 1.), (-(nx - 1) / 2.,) * 3)


def noisy_projections(g, seed, nu=NU, views=NVIEWS, i0=I0, nx=NX):
    q = rtk.phantoms.random_ellipsoids(seed, n=20, half_fov=nx / 2.)
    det0 = -(nu - 1) * 0.8 / 2
    clean = rtk.project_quadrics(g, rtk.ConstantImageSource((nu, nu, views), (0.8, 0.8, 1.), (det0, det0, 0.)), q)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(int(seed))
    counts = torch.poisson(i0 * torch.exp(-clean.tensor.double()), generator=gen)
    return clean.like((-torch.log(torch.clamp(counts, min=1.) / i0)).float().contiguous())


def fdk(g, proj, vol):
    f = rtk.HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, vol)
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.Update()
    return f.GetOutput()


def cg(g, proj, x0, niter, tol=0.):
    f = rtk.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, x0)
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.SetBackProjectionFilter(f.BP_JOSEPH)
    f.SetNumberOfIterations(niter)
    f.SetResidualTolerance(tol)
    f.Update()
    return f.GetOutput(), f.GetResiduals(), f.GetNumberOfIterationsDone()


def make_pair(g, seed, cg_iter=40):
    vol = volume_source()
    proj = noisy_projections(g, seed)
    x_fdk = fdk(g, proj, vol)
    x_cg, res, _ = cg(g, proj, vol, cg_iter)
    return x_fdk.tensor, x_cg.tensor, proj, float(res[-1])


def random_crops(pairs, crop, batch, rng):
    xs, ts = [], []
    for _ in range(batch):
        x, t = pairs[rng.integers(len(pairs))][:2]
        z, y, w = (int(rng.integers(0, n - crop + 1)) for n in x.shape)
        xs.append(x[z:z + crop, y:y + crop, w:w + crop])
        ts.append(t[z:z + crop, y:y + crop, w:w + crop])
    return torch.stack(xs).contiguous(), torch.stack(ts).contiguous()


def train(net, pairs, steps, crop=64, batch=2, lr=1e-3, seed=0, val=None, val_every=500):
    """Adam on random crops with a cosine learning rate.  With `val` (a function of the
    network, lower is better) the parameters are scored every `val_every` steps and at
    the end, and the best-scoring ones are kept."""
    rng = np.random.default_rng(seed)
    losses = []
    best, best_params = float("inf"), None
    for s in range(steps):
        x, t = random_crops(pairs, crop, batch, rng)
        step_lr = lr * 0.5 * (1 + np.cos(np.pi * s / steps))  # cosine decay
        losses.append(net.train_step(x, t, lr=step_lr))
        if (s + 1) % 100 == 0:
            log(f"train step {s + 1}/{steps}: loss {np.mean(losses[-100:]):.4g}")
        if val is not None and ((s + 1) % val_every == 0 or s + 1 == steps):
            score = val(net)
            log(f"step {s + 1}: validation score {score:.4g}")
            if score < best:
                best, best_params = score, net.params()
    if best_params is not None:
        net.set_params(best_params)
        log(f"kept the parameters of validation score {best:.4g}")
    return losses


def validation_sample(g, seed):
    """(FDK volume, B = BP_Joseph(p), the stack image) of one seed."""
    vol = volume_source()
    proj = noisy_projections(g, seed)
    x_fdk = fdk(g, proj, vol)
    bp = rtk.HipJosephBackProjectionImageFilter()
    bp.SetInput(0, vol)
    bp.SetInput(1, proj)
    bp.SetGeometry(g)
    bp.Update()
    return x_fdk.tensor, bp.GetOutput().tensor.clone(), proj


def start_residual(net, g, samples):
    """Median over validation samples (FDK volume, B = BP(p), stack) of the relative
    residual ||B - A CNN(FDK)|| / ||B|| the network's start gives CG."""
    r = []
    for x_fdk, b, proj in samples:
        x0 = volume_source().like(net(x_fdk).contiguous())
        ax = apply_normal_operator(g, x0, proj)
        r.append(float(torch.linalg.norm((b - ax).double()) / torch.linalg.norm(b.double())))
    return float(np.median(r))


def apply_normal_operator(g, x, stack_like):
    """A x = BP_Joseph(FP_Joseph(x)) with unit weights: the operator of the CG runs
    (ReconstructionConjugateGradientOperator with the matched Joseph pair)."""
    fp = rtk.HipForwardProjectionImageFilter()
    fp.SetInput(0, stack_like.like(torch.zeros_like(stack_like.tensor)))
    fp.SetInput(1, x)
    fp.SetGeometry(g)
    fp.Update()
    bp = rtk.HipJosephBackProjectionImageFilter()
    bp.SetInput(0, x.like(torch.zeros_like(x.tensor)))
    bp.SetInput(1, fp.GetOutput())
    bp.SetGeometry(g)
    bp.Update()
    return bp.GetOutput().tensor


def make_full_sample(g, seed, cg_iter=40):
    """(FDK, B = BP_Joseph(p), cg_iter-iteration CG volume, the stack image) of one seed."""
    vol = volume_source()
    proj = noisy_projections(g, seed)
    x_fdk = fdk(g, proj, vol)
    bp = rtk.HipJosephBackProjectionImageFilter()
    bp.SetInput(0, vol)
    bp.SetInput(1, proj)
    bp.SetGeometry(g)
    bp.Update()
    x_cg, res, _ = cg(g, proj, vol, cg_iter)
    return x_fdk.tensor, bp.GetOutput().tensor.clone(), x_cg.tensor, proj


def finetune_residual(net, g, samples, steps, lam_img=1.0, lam_res=1.0, lr=2e-4, seed=0, crop=128, halo=6):
    """Whole-volume objective loss = lam_img mean((s (y - t))^2) + lam_res ||B - A y||^2 / ||B||^2
    with y = net(FDK), t the CG target and A the CG operator (the matched Joseph pair):
    the second term is the start's squared relative residual, the quantity CG's stopping
    test reads.  Its gradient, -2 lam_res A (B - A y) / ||B||^2 (A is symmetric), costs two
    applications of A to the whole-volume output per step.  The network's parameters
    then follow that output gradient on a random crop^3 block (a stochastic estimate of
    the whole-volume parameter gradient, scaled to it): the block is run with `halo`
    voxels of context (the network's receptive radius is 6, so its centre equals the
    whole-volume output) and the gradient enters the network's own backward pass,
    which differentiates the first term's form, through the target t' that gives the
    same output gradient, t' = y - (n / (2 s^2)) dL/dy (t' = y on the halo)."""
    rng = np.random.default_rng(seed)
    stats = []
    vol_like = volume_source()
    s2 = net.scale * net.scale
    for it in range(steps):
        x, b, t, stack = samples[rng.integers(len(samples))]
        y = net.forward(x)
        r = b - apply_normal_operator(g, vol_like.like(y), stack)
        ar = apply_normal_operator(g, vol_like.like(r.contiguous()), stack)
        bb = float(torch.dot(b.ravel().double(), b.ravel().double()))
        rel2 = float(torch.dot(r.ravel().double(), r.ravel().double())) / bb
        n_full = y.numel()
        grad = (lam_img * 2. * s2 / n_full) * (y - t) - (2. * lam_res / bb) * ar  # dL/dy
        img = float(torch.mean((net.scale * (y - t)) ** 2))
        Z, Y, X = y.shape
        c = [min(crop, n) for n in (Z, Y, X)]
        o = [int(rng.integers(0, n - cn + 1)) for n, cn in zip((Z, Y, X), c)]
        lo = [max(q - halo, 0) for q in o]
        hi = [min(q + cn + halo, n) for q, cn, n in zip(o, c, (Z, Y, X))]
        xc = x[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].contiguous()
        yc = net.forward(xc)
        inner = tuple(slice(q - l, q - l + cn) for q, l, cn in zip(o, lo, c))
        gc = torch.zeros_like(yc)
        gc[inner] = grad[o[0]:o[0] + c[0], o[1]:o[1] + c[1], o[2]:o[2] + c[2]] * (n_full / float(np.prod(c)))
        net.backward_to((yc - (yc.numel() / (2. * s2)) * gc).contiguous())
        step_lr = lr * 0.5 * (1 + np.cos(np.pi * it / steps))
        net.adam(step_lr)
        stats.append((img, rel2))
        if (it + 1) % 20 == 0:
            m = np.mean(stats[-20:], axis=0)
            log(f"finetune step {it + 1}/{steps}: image loss {m[0]:.4g}, start rel residual {np.sqrt(m[1]):.4g}")
    return stats


def evaluate_seed(net, g, seed, tol=1e-3, max_iter=200):
    """CG time to relative residual `tol` from zero, FDK and CNN(FDK) starts on one
    phantom seed; the start's own cost (FDK, CNN inference) is part of its time."""
    vol = volume_source()
    proj = noisy_projections(g, seed)
    # the voxelised phantom: every end point's distance to the truth, to show that
    # the faster start does not buy its speed with a worse volume
    truth = rtk.draw_quadrics(vol, rtk.phantoms.random_ellipsoids(seed, n=20, half_fov=NX / 2.)).tensor
    torch.cuda.synchronize()
    out = {"seed": int(seed)}
    vols = {}
    for name in ("zero", "fdk", "cnn"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if name == "zero":
            x0 = vol
        else:
            x0 = fdk(g, proj, vol)
            if name == "cnn":
                x0 = x0.like(net(x0.tensor))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x, res, it = cg(g, proj, x0, max_iter, tol)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out[name] = {"seconds": round(t2 - t0, 4), "start_seconds": round(t1 - t0, 4), "iterations": int(it),
                     "start_rel_residual": float(res[0]), "final_rel_residual": float(res[-1]),
                     "reached": bool(res[-1] <= tol),
                     "rmse_vs_phantom": float(torch.sqrt(torch.mean((x.tensor - truth) ** 2)).item())}
        vols[name] = x.tensor
    # the end points all sit at the same residual; their spread is the solution's own slack
    for name in ("fdk", "cnn"):
        d = vols[name] - vols["zero"]
        out[name]["rms_diff_vs_zero_start"] = float(torch.sqrt(torch.mean(d * d)).item())
    return out


def evaluate(net, g, seeds, tol=1e-3, max_iter=200):
    """evaluate_seed over held-out seeds."""
    rows = []
    for seed in seeds:
        out = evaluate_seed(net, g, seed, tol, max_iter)
        log(f"seed {seed}: " + ", ".join(f"{k} {out[k]['iterations']} it {out[k]['seconds']:.2f} s"
                                         for k in ("zero", "fdk", "cnn")))
        rows.append(out)
    return rows


def scale_problem(seed):
    """The at-scale workload: 512^3 @0.5 mm from 180 views of 1024^2 @0.4 mm (SID
    1000 mm, SDD 1536 mm), 30 random ellipsoids, Poisson noise I0 = 5e3.  Returns
    (geometry, noisy projections, volume source, quadrics)."""
    views, nu, nx = 180, 1024, 512
    g = geometry(views)
    vol = rtk.ConstantImageSource((nx, nx, nx), (0.5, 0.5, 0.5), (-(nx - 1) * 0.5 / 2,) * 3)
    q = rtk.phantoms.random_ellipsoids(seed, n=30, half_fov=nx * 0.5 / 2)
    det0 = -(nu - 1) * 0.4 / 2
    clean = rtk.project_quadrics(g, rtk.ConstantImageSource((nu, nu, views), (0.4, 0.4, 1.), (det0, det0, 0.)), q)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(int(seed))
    counts = torch.poisson(5e3 * torch.exp(-clean.tensor.double()), generator=gen)
    proj = clean.like((-torch.log(torch.clamp(counts, min=1.) / 5e3)).float().contiguous())
    del clean, counts
    return g, proj, vol, q


def make_scale_pair(seed, cg_iter=40, tile=128, halo=8):
    """(FDK, cg_iter-iteration CG volume) of one at-scale phantom."""
    g, proj, vol, _ = scale_problem(seed)
    x_fdk = fdk(g, proj, vol)
    x_cg, res, _ = cg(g, proj, vol, cg_iter)
    return x_fdk.tensor, x_cg.tensor, float(res[-1])


def evaluate_at_scale(net, seed, tol=1e-3, max_iter=300, tile=128, halo=8):
    """The learned start at scale (scale_problem).  The network runs as overlapping
    tile^3 blocks with `halo` voxels of context, then CG (Joseph pair, FP64 scalars)
    runs to relative residual `tol` from zero, FDK and CNN(FDK).  Returns the same
    fields as evaluate_seed."""
    g, proj, vol, q = scale_problem(seed)
    truth = rtk.draw_quadrics(vol, q).tensor
    out = {"seed": int(seed)}
    vols = {}
    for name in ("zero", "fdk", "cnn"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if name == "zero":
            x0 = vol
        else:
            x0 = fdk(g, proj, vol)
            if name == "cnn":
                x0 = x0.like(net.forward_tiled(x0.tensor, tile=tile, halo=halo))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x, res, it = cg(g, proj, x0, max_iter, tol)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out[name] = {"seconds": round(t2 - t0, 3), "start_seconds": round(t1 - t0, 4), "iterations": int(it),
                     "start_rel_residual": float(res[0]), "final_rel_residual": float(res[-1]),
                     "reached": bool(res[-1] <= tol),
                     "rmse_vs_phantom": float(torch.sqrt(torch.mean((x.tensor - truth) ** 2)).item())}
        vols[name] = x.tensor
        log(f"scale seed {seed} {name}: {out[name]}")
    for name in ("fdk", "cnn"):
        d = vols[name] - vols["zero"]
        out[name]["rms_diff_vs_zero_start"] = float(torch.sqrt(torch.mean(d * d)).item())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=500)
    ap.add_argument("--eval", type=int, default=16)
    ap.add_argument("--eval-seed0", type=int, default=EVAL_SEED0,
                    help="first evaluation seed (held-out test seeds from 5000; other ranges for model selection)")
    ap.add_argument("--seed", type=int, default=0, help="network initialisation and crop order of --train")
    ap.add_argument("--select", type=int, default=0,
                    help="keep the training checkpoint whose start residual is lowest on this many seeds from 7000")
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--cg-iter", type=int, default=40)
    ap.add_argument("--out", default="warmstart_c4.npz")
    ap.add_argument("--report", default="warmstart_c4.json")
    ap.add_argument("--eval-only", action="store_true",
                    help="evaluate the shipped network on --eval held-out seeds, without training")
    ap.add_argument("--scale-eval", type=int, default=0,
                    help="evaluate at 512^3 / 180 views on this many held-out seeds (after --scale-train if given)")
    ap.add_argument("--scale-train", type=int, default=0,
                    help="train a network on this many 512^3 / 180-view pairs (toward the --cg-iter volume)")
    ap.add_argument("--weights", default=None, help="network for --eval-only / --scale-eval (default: shipped)")
    ap.add_argument("--finetune", type=int, default=0,
                    help="whole-volume fine-tuning steps of the --weights network with the residual term")
    ap.add_argument("--ft-samples", type=int, default=48, help="training seeds for --finetune")
    ap.add_argument("--lam-img", type=float, default=1.0)
    ap.add_argument("--lam-res", type=float, default=1.0)
    ap.add_argument("--ft-crop", type=int, default=128, help="block size of --finetune's parameter gradient")
    args = ap.parse_args()
    rtk.lib()
    if args.scale_train or args.scale_eval:
        from rtk_amd.warmstart import default_weights_path, scale_weights_path
        if args.scale_train:
            t0 = time.perf_counter()
            pairs = []
            for i in range(args.scale_train):
                x, t, r = make_scale_pair(TRAIN_SEED0 + 100 + i, args.cg_iter)
                pairs.append((x, t))
                log(f"scale pairs {i + 1}/{args.scale_train} ({time.perf_counter() - t0:.0f} s), last CG residual {r:.3g}")
            scale = float(1.0 / torch.stack([p[1].abs().mean() for p in pairs]).mean().item())
            net = WarmStartCNN(scale=scale, seed=0)
            t1 = time.perf_counter()
            losses = train(net, pairs, args.steps, args.crop, args.batch, args.lr)
            torch.cuda.synchronize()
            log(f"scale training: {args.scale_train} pairs {t1 - t0:.0f} s, {args.steps} steps "
                f"{time.perf_counter() - t1:.0f} s, loss {np.mean(losses[:20]):.3g} -> {np.mean(losses[-20:]):.3g}")
            del pairs
            net.save(args.out)
        else:
            path = args.weights or (scale_weights_path() if os.path.exists(scale_weights_path())
                                    else default_weights_path())
            net = WarmStartCNN.load(path)
        rows = [evaluate_at_scale(net, EVAL_SEED0 + 100 + i) for i in range(args.scale_eval)]
        print(json.dumps({"workload": "CG 512^3 @0.5 mm <- 180 x 1024^2 @0.4 mm, 30 ellipsoids, I0=5e3",
                          "eval": rows}), flush=True)
        return
    g = geometry()
    if args.finetune:
        from rtk_amd.warmstart import default_weights_path
        net = WarmStartCNN.load(args.weights or default_weights_path())
        t0 = time.perf_counter()
        samples = [make_full_sample(g, TRAIN_SEED0 + i, args.cg_iter) for i in range(args.ft_samples)]
        t1 = time.perf_counter()
        log(f"{args.ft_samples} whole-volume samples in {t1 - t0:.0f} s")
        stats = finetune_residual(net, g, samples, args.finetune, args.lam_img, args.lam_res, args.lr,
                                  crop=args.ft_crop)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        del samples
        net.save(args.out)
        rows = evaluate(net, g, [args.eval_seed0 + i for i in range(args.eval)])
        rep = {"finetune_steps": args.finetune, "samples": args.ft_samples, "lam_img": args.lam_img,
               "lam_res": args.lam_res, "lr": args.lr, "cg_iter": args.cg_iter, "data_seconds": t1 - t0,
               "train_seconds": t2 - t1, "first": list(np.mean(stats[:20], axis=0)),
               "last": list(np.mean(stats[-20:], axis=0)), "eval": rows}
        for name in ("zero", "fdk", "cnn"):
            rep[f"median_{name}_seconds"] = float(np.median([r[name]["seconds"] for r in rows]))
            rep[f"median_{name}_iterations"] = float(np.median([r[name]["iterations"] for r in rows]))
            rep[f"median_{name}_rmse_vs_phantom"] = float(np.median([r[name]["rmse_vs_phantom"] for r in rows]))
        rep["speedup_cnn_vs_zero"] = rep["median_zero_seconds"] / rep["median_cnn_seconds"]
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
        log(json.dumps({k: v for k, v in rep.items() if k != "eval"}))
        return
    if args.eval_only:
        from rtk_amd.warmstart import default_weights_path
        net = WarmStartCNN.load(args.weights or default_weights_path())
        rows = evaluate(net, g, [args.eval_seed0 + i for i in range(args.eval)])
        rep = {"workload": "CG 256^3 @1 mm <- 90 x 512^2 @0.8 mm, Poisson I0=1e4, held-out seeds", "eval": rows}
        for name in ("zero", "fdk", "cnn"):
            rep[f"median_{name}_seconds"] = float(np.median([r[name]["seconds"] for r in rows]))
            rep[f"median_{name}_iterations"] = float(np.median([r[name]["iterations"] for r in rows]))
            rep[f"median_{name}_rmse_vs_phantom"] = float(np.median([r[name]["rmse_vs_phantom"] for r in rows]))
        rep["speedup_cnn_vs_zero"] = rep["median_zero_seconds"] / rep["median_cnn_seconds"]
        rep["speedup_fdk_vs_zero"] = rep["median_zero_seconds"] / rep["median_fdk_seconds"]
        print(json.dumps(rep), flush=True)
        return
    t0 = time.perf_counter()
    pairs = []
    for i in range(args.train):
        x, t, _, r = make_pair(g, TRAIN_SEED0 + i, args.cg_iter)
        pairs.append((x, t))
        if (i + 1) % 8 == 0:
            log(f"pairs {i + 1}/{args.train} ({time.perf_counter() - t0:.0f} s), last CG residual {r:.3g}")
    scale = float(1.0 / torch.stack([p[1].abs().mean() for p in pairs]).mean().item())
    log(f"input scale {scale:.4g}")
    net = WarmStartCNN(scale=scale, seed=args.seed)
    val = None
    if args.select:
        samples = [validation_sample(g, SELECT_SEED0 + i) for i in range(args.select)]
        val = lambda n: start_residual(n, g, samples)  # noqa: E731
    t1 = time.perf_counter()
    losses = train(net, pairs, args.steps, args.crop, args.batch, args.lr, seed=args.seed, val=val)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    net.save(args.out)
    rows = evaluate(net, g, [args.eval_seed0 + i for i in range(args.eval)])
    rep = {"train_pairs": args.train, "steps": args.steps, "crop": args.crop, "batch": args.batch,
           "data_seconds": t1 - t0, "train_seconds": t2 - t1, "first_loss": float(np.mean(losses[:20])),
           "last_loss": float(np.mean(losses[-20:])), "scale": scale, "eval": rows}
    for name in ("zero", "fdk", "cnn"):
        rep[f"median_{name}_seconds"] = float(np.median([r[name]["seconds"] for r in rows]))
        rep[f"median_{name}_iterations"] = float(np.median([r[name]["iterations"] for r in rows]))
        rep[f"median_{name}_rmse_vs_phantom"] = float(np.median([r[name]["rmse_vs_phantom"] for r in rows]))
    rep["speedup_cnn_vs_zero"] = rep["median_zero_seconds"] / rep["median_cnn_seconds"]
    with open(args.report, "w") as f:
        json.dump(rep, f, indent=1)
    log(json.dumps({k: v for k, v in rep.items() if k != "eval"}))


if __name__ == "__main__":
    main()
