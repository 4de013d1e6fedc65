"""Training and evaluation of the learned light cut (ARC_CUT_LEARNED).

Training rows come from the exact gather: arc_gather_cut_samples walks each
8x8 tile's cut with exact labels (a node is "refine" when some pixel of the
tile sees the node's estimate off the exact sum of its AVPLs by more than the
pixel's threshold) and records the features of every internal node it meets.
The scenes are seeded variants of the frame workload (Cornell walls + random
triangles, AVPL radiance log-uniform over [1e-3, 10], 30 % antiradiance) at
mixed resolutions and AVPL counts, with the error bound drawn per scene, so
the network sees the thresholds and tile sizes it will meet.  The network is
then trained on the GPU (arc_cutnet_train) and saved as an .npz holding the
ARC_CUTNET_PARAMS weights and the training record.

    python -m antiradiancecuts_amd.learned_cut train --out antiradiancecuts_amd/data/cutnet_clutter.npz
    python -m antiradiancecuts_amd.learned_cut evaluate
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np

from . import ffi
from . import scene as S
from .runtime import Context, CutNet, DeviceAvpls, Gather, default_params, device_array

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "data" / "cutnet_clutter.npz"


@dataclass
class TrainConfig:
    scenes: int = 32
    width_lo: int = 320           # width log-uniform in [width_lo, width_hi], height = width * 9 / 16
    width_hi: int = 1280
    avpls_lo: int = 32768         # AVPL count log-uniform
    avpls_hi: int = 524288
    tris: int = 200_000
    eps_lo: float = 0.003         # error bound log-uniform
    eps_hi: float = 0.03
    rows: int = 5_000_000         # training rows over all scenes
    epochs: int = 12
    batch: int = 4096
    lr: float = 2e-3
    seed: int = 2024
    val_scenes: int = 4


def _scene(cfg: TrainConfig, k: int):
    rng = np.random.default_rng(cfg.seed * 7919 + k)
    w = int(round(math.exp(rng.uniform(math.log(cfg.width_lo), math.log(cfg.width_hi))) / 8)) * 8
    h = max(8, int(round(w * 9 / 16 / 8)) * 8)
    n = int(round(math.exp(rng.uniform(math.log(cfg.avpls_lo), math.log(cfg.avpls_hi)))))
    eps = float(math.exp(rng.uniform(math.log(cfg.eps_lo), math.log(cfg.eps_hi))))
    sc = S.clutter_1080p(seed=cfg.seed * 1000 + k, width=w, height=h, num_avpls=n, num_tris=cfg.tris)
    return sc, eps


def collect(ctx: Context, cfg: TrainConfig, first: int, count: int, rows: int, log=print):
    """Rows of scenes first..first+count-1, about rows / count kept per scene."""
    import torch

    per_scene = max(1, rows // max(count, 1))
    Xs, ys = [], []
    gen = torch.Generator(device=f"cuda:{ctx.device}")
    gen.manual_seed(cfg.seed + first)
    for k in range(first, first + count):
        t0 = time.time()
        sc, eps = _scene(cfg, k)
        g = Gather(ctx, sc.width, sc.height, sc.materials)
        g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
        av = DeviceAvpls.from_scene(sc)
        p = default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, errorBound=eps)
        # every row of the scene (a few thousand internal nodes per tile at
        # the tightest bounds), then the scene's share drawn from them
        tiles = ((sc.width + 7) // 8) * ((sc.height + 7) // 8)
        keep = 1.0
        cap = 4096 * tiles
        X, y, met = g.cut_samples(av, sc.cam, sc.extent, p, capacity=cap, seed=cfg.seed + k, keep=keep)
        if met > cap:
            log(f"[learned_cut] scene {k}: {met} rows, {met - cap} beyond the capacity dropped")
        take = torch.randperm(X.shape[0], device=X.device, generator=gen)[:per_scene]
        Xs.append(X[take].clone())
        ys.append(y[take].clone())
        log(f"[learned_cut] scene {k}: {sc.width}x{sc.height}, {sc.num_avpls} AVPLs, eps {eps:.4f}: "
            f"{met} rows at keep {keep:.3f} (refine {float(y.mean()):.3f}), taken {len(take)}, "
            f"{time.time() - t0:.1f} s")
        del X, y
        g.close()
    return torch.cat(Xs), torch.cat(ys)


def train(ctx: Context, cfg: TrainConfig, log=print):
    """Returns (CutNet, record)."""
    import torch

    t0 = time.time()
    X, y = collect(ctx, cfg, 0, cfg.scenes, cfg.rows, log)
    Xv, yv = collect(ctx, cfg, cfg.scenes, cfg.val_scenes, max(cfg.rows // 10, 1), log)
    t_rows = time.time() - t0
    mean = X.mean(0).cpu().numpy()
    std = X.std(0).cpu().numpy()
    net = CutNet(ctx)
    net.init(cfg.seed, mean, np.where(std > 1e-6, std, 1.0))
    t1 = time.time()
    loss = net.train(X, y, epochs=cfg.epochs, batch=cfg.batch, lr=cfg.lr, seed=cfg.seed)
    ctx.sync()
    t_train = time.time() - t1

    def stats(Xa, ya):
        p = net.forward(Xa)
        yb = ya > 0.5
        conf = (p >= 0.9) | (p <= 0.1)
        return {"rows": int(Xa.shape[0]), "refine_share": float(ya.mean()),
                "accuracy": float(((p > 0.5) == yb).float().mean()),
                "confident_share": float(conf.float().mean()),
                "confident_accuracy": float(((p > 0.5) == yb)[conf].float().mean()) if bool(conf.any()) else None,
                "false_keep_confident": float(((p <= 0.1) & yb).float().mean())}

    record = {"config": asdict(cfg), "loss_per_epoch": [float(v) for v in loss], "train": stats(X, y),
              "validation": stats(Xv, yv), "rows_seconds": round(t_rows, 1), "train_seconds": round(t_train, 2)}
    log(f"[learned_cut] loss {np.round(loss, 4).tolist()}")
    log(f"[learned_cut] train {record['train']}")
    log(f"[learned_cut] validation {record['validation']}")
    return net, record


def save(net: CutNet, record: dict, path: Path) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, weights=net.get_weights(), record=np.array(json.dumps(record)))


def load(ctx: Context, path: Path = DEFAULT_WEIGHTS) -> CutNet:
    """A CutNet with the saved weights; raises when the file is missing."""
    if not Path(path).exists():
        raise FileNotFoundError(f"no learned-cut weights at {path} (python -m antiradiancecuts_amd.learned_cut train)")
    with np.load(path, allow_pickle=False) as z:
        w = z["weights"]
    net = CutNet(ctx)
    net.set_weights(w)
    return net


def evaluate(ctx: Context, net: CutNet, eps_list=(0.003, 0.01, 0.03), width=1920, height=1080,
             avpls=1_000_000, tris=200_000, stride=4, reps=5, log=print) -> dict:
    """Radiance-kernel time and result error of the error-bound tile cut and
    the learned cut on the frame workload, against the exact FP64 gather on
    every `stride`-th pixel in x and y."""
    import torch

    sc = S.clutter_1080p(width=width, height=height, num_avpls=avpls, num_tris=tris)
    g = Gather(ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    g.set_cutnet(net)
    av = DeviceAvpls.from_scene(sc)
    ys, xs = np.arange(0, sc.height, stride), np.arange(0, sc.width, stride)
    pix = (ys[:, None] * sc.width + xs[None, :]).ravel()
    sub = S.Scene(**{**sc.__dict__, "width": len(xs), "height": len(ys), "pos4": sc.pos4[pix].copy(),
                     "nrm4": sc.nrm4[pix].copy(), "albedo4": sc.albedo4[pix].copy()})
    gs = Gather(ctx, sub.width, sub.height, sub.materials)
    gs.set_gbuffer(device_array(sub.pos4), device_array(sub.nrm4), device_array(sub.albedo4))
    base = default_params(useLightCuts=1, antiradianceModel=ffi.ARC_ANTIRAD_DISK)
    r64, a64 = gs.exact_f64(av, sc.cam, sc.extent, base)
    torch.cuda.synchronize()
    ref = (r64 - a64).cpu().numpy()
    rref = r64.cpu().numpy()
    gs.close()
    rows = []
    for mode in ("error-bound", "learned"):
        for eps in eps_list:
            p = default_params(useLightCuts=1, antiradianceModel=ffi.ARC_ANTIRAD_DISK, errorBound=eps,
                               cutMode=ffi.ARC_CUT_LEARNED if mode == "learned" else ffi.ARC_CUT_ERROR_BOUND,
                               cutPerTile=1)
            flags = ffi.ARC_RUN_TIMELINE | ffi.ARC_TRUST_INPUT
            for _ in range(2):
                g.run(av, sc.cam, sc.extent, p, flags)
            for _ in range(reps):
                g.run(av, sc.cam, sc.extent, p, flags)
            kern = np.zeros((reps, 2), np.float32)
            g.lib.arc_gather_kernel_history(g.h, reps, kern.ctypes.data)
            hist = np.zeros((reps, 6), np.float32)
            g.lib.arc_gather_profile_history(g.h, reps, hist.ctypes.data)
            st = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_COUNTERS | ffi.ARC_TRUST_INPUT)
            torch.cuda.synchronize()
            res = g.result[:, :3].double().cpu().numpy()[pix]
            rad = g.radiance[:, :3].double().cpu().numpy()[pix]
            row = {"mode": mode, "eps": eps, "cut_ms": float(np.median(kern[:, 0])),
                   "frame_ms": float(np.median(hist[:, 5])),
                   "rel_l2": float(np.linalg.norm(res - ref) / np.linalg.norm(ref)),
                   "radiance_rel_l2": float(np.linalg.norm(rad - rref) / np.linalg.norm(rref)),
                   "nodes_per_px": float(st.evaluatedNodes) / sc.npix,
                   "refined_per_px": float(st.visitedNodes) / sc.npix,
                   # (tile, internal node) decisions left to the criterion, over all tile decisions
                   "fallback_share": float(st.fallbackNodes) * 64.0 / max(1.0, float(st.visitedNodes +
                                                                                   st.evaluatedNodes))}
            rows.append(row)
            log(f"[learned_cut] {json.dumps(row)}")
    g.close()
    return {"workload": f"clutter {width}x{height}, {avpls} AVPLs, {tris} triangles", "rows": rows,
            "equal_error": equal_error(rows)}


def equal_error(rows) -> list:
    """For each learned point, the error-bound cut's radiance-kernel time at
    the same result error (log-log interpolation between its points)."""
    eb = sorted([r for r in rows if r["mode"] == "error-bound"], key=lambda r: r["rel_l2"])
    out = []
    for r in rows:
        if r["mode"] != "learned":
            continue
        e = r["rel_l2"]
        t = None
        for a, b in zip(eb, eb[1:]):
            if a["rel_l2"] <= e <= b["rel_l2"] and b["rel_l2"] > a["rel_l2"]:
                w = (math.log(e) - math.log(a["rel_l2"])) / (math.log(b["rel_l2"]) - math.log(a["rel_l2"]))
                t = math.exp(math.log(a["cut_ms"]) + w * (math.log(b["cut_ms"]) - math.log(a["cut_ms"])))
        out.append({"eps": r["eps"], "rel_l2": e, "learned_cut_ms": r["cut_ms"], "error_bound_cut_ms_at_equal_error": t,
                    "speedup": (t / r["cut_ms"]) if t else None})
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", type=Path, default=DEFAULT_WEIGHTS)
    t.add_argument("--scenes", type=int, default=TrainConfig.scenes)
    t.add_argument("--rows", type=int, default=TrainConfig.rows)
    t.add_argument("--epochs", type=int, default=TrainConfig.epochs)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", type=Path, default=DEFAULT_WEIGHTS)
    e.add_argument("--out", type=Path, default=None)
    args = ap.parse_args(argv)
    ctx = Context(0)
    if args.cmd == "train":
        cfg = TrainConfig(scenes=args.scenes, rows=args.rows, epochs=args.epochs)
        net, record = train(ctx, cfg, log=lambda s: print(s, flush=True))
        save(net, record, args.out)
        print(json.dumps(record), flush=True)
        net.close()
    else:
        net = load(ctx, args.weights)
        out = evaluate(ctx, net, log=lambda s: print(s, flush=True))
        print(json.dumps(out), flush=True)
        if args.out:
            args.out.write_text(json.dumps(out, indent=1))
        net.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
