"""Inverse rendering with the path-replay adjoint: recover the shared base-colour texture and the
per-sphere roughness of the sphere grid from images.

    python -m mitrace_amd.inverse optimize [--iters 500] [--json out.json]
    python -m mitrace_amd.inverse gradcheck [--json out.json]

optimize: targets from the ground-truth parameters at 8 seeded 512x512 poses (1024 spp);
initial texture = truth + N(0, 0.15) clamped to [0, 1], roughness = truth + U(-0.3, 0.3)
(clamped to [0.05, 1]); per iteration one 32-spp render of every pose with its own seed, loss =
mean squared error over the 8 images, Adam (lr 1e-2) on texture and roughness.

gradcheck: gradients of that loss at the initial parameters against central differences
(h = 1e-3): per parameter for 64 random texels and every sphere's roughness and metallic with
the sampling frozen against the detached adjoint (the estimator it differentiates), and along
random directions with common random numbers only (4096 spp), which also sees the changed
sampling decisions, against the full adjoint.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from . import scenegen
from .diff import DiffRenderer
from .tracer import GpuScene

GROUPS = ("tex", "rough", "metal", "emis")


def setup(width=512, height=512, seed=4):
    sc, truth = scenegen.inverse_scene(seed=seed)
    cams = scenegen.inverse_poses(8, seed=seed)
    gs = GpuScene(sc)
    return sc, truth, cams, gs


def to_dev(params):
    return {k: torch.tensor(np.ascontiguousarray(params[k], np.float32), device="cuda")
            for k in GROUPS}


def perturbed(truth, seed=4):
    rng = np.random.default_rng(seed + 77)
    init = dict(truth)
    init["tex"] = np.clip(truth["tex"] + rng.normal(0.0, 0.15, truth["tex"].shape), 0.0, 1.0)
    r = truth["rough"].copy()
    r[:49] = np.clip(r[:49] + rng.uniform(-0.3, 0.3, 49), 0.05, 1.0)
    init["rough"] = r
    return {k: (np.asarray(v, np.float32) if k in GROUPS else v) for k, v in init.items()}


def targets(gs, truth, cams, spp=1024, width=512, height=512):
    st = scenegen.inverse_settings(width=width, height=height, samples=spp)
    dr = DiffRenderer(gs, st, truth["tex_first"], truth["tex_count"])
    d = to_dev(truth)
    out = [dr.forward(c, 900_000 + i, **d) for i, c in enumerate(cams)]
    dr.close()
    return out


def optimize(iters=500, spp=32, lr=1e-2, log_every=10, target_spp=1024, decorrelated=True,
             tex_lr=None, eval_spp=128):
    """decorrelated: the residual (image - target) comes from an independent render (another
    seed) than the one whose adjoint is taken.  With one render for both, the gradient of the
    squared error carries 2 Cov(I, dI/dtheta), which for a base colour (I roughly linear in it)
    is positive and steadily darkens the texture."""
    sc, truth, cams, gs = setup()
    t0 = time.perf_counter()
    tg = targets(gs, truth, cams, spp=target_spp)
    t_targets = time.perf_counter() - t0
    init = perturbed(truth)
    st = scenegen.inverse_settings(samples=spp)
    dr = DiffRenderer(gs, st, truth["tex_first"], truth["tex_count"])
    dev = to_dev(init)
    tex = dev["tex"].clone().requires_grad_(True)
    rough = dev["rough"].clone().requires_grad_(True)
    metal, emis = dev["metal"], dev["emis"]
    # Adam normalises every texel's step to about lr whatever its gradient's signal-to-noise
    # ratio, so texels seen by few noisy pixels random-walk; tex_lr sets their own step size
    opt = torch.optim.Adam([{"params": [tex], "lr": tex_lr or lr}, {"params": [rough], "lr": lr}])
    gt = to_dev(truth)
    log = []

    seen = torch.zeros(tex.shape[:2], dtype=torch.int32, device="cuda")

    def errors():
        with torch.no_grad():
            return (float(torch.sqrt(torch.mean((tex - gt["tex"]) ** 2))),
                    float(torch.mean(torch.abs(rough[:49] - gt["rough"][:49]))))

    def image_rmse(spp_eval=eval_spp):
        """RMSE of the 8 poses rendered with the current parameters against the targets
        (None when eval_spp is 0)"""
        if not spp_eval:
            return None
        ev = DiffRenderer(gs, scenegen.inverse_settings(samples=spp_eval), truth["tex_first"],
                          truth["tex_count"])
        with torch.no_grad():
            err = [float(torch.mean((ev.forward(c, 700_000 + i, tex.detach(), rough.detach(),
                                                metal, emis) - tg[i]) ** 2))
                   for i, c in enumerate(cams)]
        ev.close()
        return float(np.sqrt(np.mean(err)))

    e0 = errors()
    img0 = image_rmse()
    it_times = []
    for it in range(iters):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        opt.zero_grad()
        total = 0.0
        for i, cam in enumerate(cams):
            seed = 1 + 2 * (it * len(cams) + i)
            if decorrelated:
                with torch.no_grad():
                    resid = dr.forward(cam, seed + 1, tex.detach(), rough.detach(), metal, emis) - tg[i]
                img = dr.render(cam, seed, tex, rough, metal, emis)
                # surrogate whose gradient is 2 * resid * dI/dtheta / N
                loss = torch.sum(2.0 * resid * img) / resid.numel() / len(cams)
                loss.backward()
                total += float(torch.mean(resid ** 2)) / len(cams)
            else:
                img = dr.render(cam, seed, tex, rough, metal, emis)
                loss = torch.mean((img - tg[i]) ** 2) / len(cams)
                loss.backward()
                total += float(loss.detach())
        with torch.no_grad():
            seen += (tex.grad.abs().sum(-1) > 0).int()
        opt.step()
        with torch.no_grad():
            tex.clamp_(0.0, 1.0)
            rough.clamp_(0.05, 1.0)
        torch.cuda.synchronize()
        it_times.append(time.perf_counter() - ts)
        if it % log_every == 0 or it == iters - 1:
            te, re = errors()
            log.append({"iter": it, "loss": total, "tex_rmse": te, "rough_mae": re})
            print(f"iter {it}: loss {total:.5f} texture RMSE {te:.4f} roughness MAE {re:.4f}",
                  flush=True)
    te, re = errors()
    img1 = image_rmse()
    with torch.no_grad():
        # texels with a gradient in at least half of the iterations
        m = seen >= max(1, iters // 2)
        init_t = torch.tensor(init["tex"], device="cuda")
        seen_rmse = lambda t: float(torch.sqrt(torch.mean(((t - gt["tex"])[m]) ** 2)))
        seen_frac = float(m.float().mean())
        s0, s1 = seen_rmse(init_t), seen_rmse(tex)
    return {"iters": iters, "spp": spp, "poses": len(cams), "lr": lr, "tex_lr": tex_lr or lr,
            "decorrelated": decorrelated,
            "targets_spp": target_spp, "targets_s": t_targets,
            "initial": {"tex_rmse": e0[0], "tex_rmse_observed": s0, "rough_mae": e0[1],
                        "image_rmse": img0},
            "final": {"tex_rmse": te, "tex_rmse_observed": s1, "rough_mae": re,
                      "image_rmse": img1},
            "observed_texel_fraction": seen_frac,
            "ms_per_iter": 1e3 * float(np.median(it_times)), "log": log}


def gradcheck(h=1e-3, n_texels=64, frozen_spp=64, crn_spp=4096, seed=5, crn_seeds=2):
    sc, truth, cams, gs = setup()
    tg = targets(gs, truth, cams[:1], spp=256)[0]
    init = perturbed(truth)
    rng = np.random.default_rng(seed)
    report = {"h": h}

    def loss_fn(img):
        return float(torch.sum((img.double() - tg.double()) ** 2))

    # per-parameter check with the sampling frozen
    st = scenegen.inverse_settings(samples=frozen_spp)
    dr = DiffRenderer(gs, st, truth["tex_first"], truth["tex_count"])
    d = to_dev(init)
    img = dr.forward(cams[0], seed, **d)
    g = {k: v.cpu().numpy().reshape(-1)
         for k, v in dr.backward(cams[0], seed, 2.0 * (img - tg), detach_sampling=True, **d).items()}
    # texels that some pixel sees, so that the derivative is not trivially zero
    seen = np.flatnonzero(g["tex"] != 0.0)
    cases = [("tex", int(i)) for i in rng.choice(seen, min(n_texels, len(seen)), replace=False)]
    cases += [("rough", i) for i in range(49) if init["rough"][i] > 0.05 + 3 * h]
    cases += [("metal", i) for i in range(49)]
    rows = []
    for group, i in cases:
        vals = []
        for step in (h, 3 * h):
            flat = d[group].reshape(-1)
            a, b = flat.clone(), flat.clone()
            a[i] += step
            b[i] -= step
            pp, pm = dict(d), dict(d)
            pp[group] = a.reshape(d[group].shape)
            pm[group] = b.reshape(d[group].shape)
            lp = loss_fn(dr.forward(cams[0], seed, sampling=d, **pp))
            lm = loss_fn(dr.forward(cams[0], seed, sampling=d, **pm))
            vals.append((lp - lm) / (2 * step))
        adj = float(g[group][i])
        rows.append({"group": group, "index": i, "adjoint": adj, "fd": vals[0], "fd_3h": vals[1]})
    dr.close()
    for group in ("tex", "rough", "metal"):
        r = [x for x in rows if x["group"] == group]
        scale = max(abs(x["adjoint"]) for x in r) + 1e-30
        rel = [abs(x["adjoint"] - x["fd"]) / max(abs(x["fd"]), 1e-3 * scale) for x in r]
        report[f"frozen_{group}"] = {"n": len(r), "within_5pct": int(sum(v <= 0.05 for v in rel)),
                                     "median_rel": float(np.median(rel)), "max_rel": float(max(rel))}
    report["frozen_spp"] = frozen_spp
    # directional derivatives with common random numbers only
    st = scenegen.inverse_settings(samples=crn_spp, paths_in_flight=1 << 24)
    dr = DiffRenderer(gs, st, truth["tex_first"], truth["tex_count"])
    for group in ("tex", "rough", "metal"):
        v = np.zeros(np.size(init[group]), np.float32)
        if group == "tex":
            v[:] = rng.normal(size=v.size)
        else:
            v[:49] = rng.normal(size=49)
            v[np.asarray(init[group]) <= 0.05 + 3 * h] = 0.0
        vt = torch.tensor(v.reshape(np.shape(init[group])), device="cuda")
        adj, fd = [], []
        for s in range(crn_seeds):
            img = dr.forward(cams[0], 50 + s, **d)
            gg = dr.backward(cams[0], 50 + s, 2.0 * (img - tg), **d)[group]
            adj.append(float((gg * vt).sum()))
            pp, pm = dict(d), dict(d)
            pp[group] = (d[group] + h * vt).contiguous()
            pm[group] = (d[group] - h * vt).contiguous()
            fd.append((loss_fn(dr.forward(cams[0], 50 + s, **pp)) -
                       loss_fn(dr.forward(cams[0], 50 + s, **pm))) / (2 * h))
        report[f"crn_{group}"] = {"adjoint": adj, "fd": fd,
                                  "rel_of_means": abs(np.mean(adj) - np.mean(fd)) / abs(np.mean(fd))}
        print(f"crn {group}: {report[f'crn_{group}']}", flush=True)
    report["crn_spp"] = crn_spp
    dr.close()
    report["rows"] = rows
    return report


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("optimize", "gradcheck"))
    ap.add_argument("--iters", type=int, default=500)
    ap.add_argument("--spp", type=int, default=32)
    ap.add_argument("--crn-spp", type=int, default=4096)
    ap.add_argument("--json", default=None)
    ap.add_argument("--tex-lr", type=float, default=None, help="Adam step for the texture")
    ap.add_argument("--target-spp", type=int, default=1024, help="samples of the target images")
    ap.add_argument("--eval-spp", type=int, default=128,
                    help="samples of the image-RMSE evaluation before and after (0: none)")
    ap.add_argument("--correlated", action="store_true",
                    help="one render per pose for both the residual and the adjoint")
    a = ap.parse_args(argv)
    if a.mode == "optimize":
        out = optimize(iters=a.iters, spp=a.spp, decorrelated=not a.correlated, tex_lr=a.tex_lr,
                       target_spp=a.target_spp, eval_spp=a.eval_spp)
    else:
        out = gradcheck(crn_spp=a.crn_spp)
    summary = {k: v for k, v in out.items() if k not in ("log", "rows")}
    print(json.dumps(summary), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
