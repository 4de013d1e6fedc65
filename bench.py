#!/usr/bin/env python3
"""Frame benchmark of the AVPL gather on one MI355X.

A step is one frame of CudaGather::run (cudaGather.cu:836-1051) on the
1920x1080 x 1,000,000-AVPL workload: the AvplBvh over the frame's AVPLs is
rebuilt, radiance is gathered over a light cut of it, antiradiance is
gathered, and result = radiance - antiradiance is written.  All inputs are in
HBM before the timed region.

Prints one JSON line: pixel-AVPL pairs/s (W*H*N per frame time), ms/frame,
the error of the frame against the exact FP64 gather on a pixel sample, the
roofline of the dominant kernel (timed with HIP events on the library's
stream inside the timed region) and the CPU baseline (the oracle timed on
this host's cores on a bounded sample of the same frame).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pixel-AVPL pairs/s + ms/frame at 1080p×1M AVPLs; rel-L2 vs FP64 oracle"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (published; the same for the vector and the matrix units)
# The roofline's compute bound in the line's vocabulary ("hbm" or "mfma"):
# the gather's kernels are bound by FP32 compute, which they do on the
# vector units ("engine": "VALU"); the FP32 peak is the same for both units.
COMPUTE_BOUND = "mfma"
PEAK_HBM_GBS = 8000.0
# Algorithmic FLOP per unit (FMA = 2, transcendental = 1):
#   radiance pair 31; disk antiradiance pair 37; cone antiradiance pair 33;
#   solid-angle cut: visited inner node 32, evaluated node 31 + 31;
#   error-bound cut: node test (criterion, counted per (pixel, inner node)
#   test) 45, node estimate (per kept node or, with sums D levels down, per
#   descendant summed) 37;
#   clustered antiradiance: tile-filter ray-box test 35 per list entry and
#   tile, plus the pair tests.
FLOP_RAD_PAIR, FLOP_DISK_PAIR, FLOP_CONE_PAIR = 31, 37, 33
FLOP_VISIT, FLOP_EVAL = 32, 31
FLOP_EB_TEST, FLOP_EB_EST = 45, 37
FLOP_TILE_FILTER = 35
# moment estimator: Lightcuts-bound test per (pixel, inner node) 32 (the
# box distance and tangent-plane height, 32 as a visited node); the
# expansion criterion adds the centroid distance and two products, 45; a
# moment estimate 161 (four pixel-side points 4 x 18 + 4, the direction to
# the centroid 10 and its six products 6, per channel the normal moments'
# form 20 and the clamp, the product and the add 3); a leaf summed exactly
# is a radiance pair, 31; the cone filter of an antiradiance tile list 30.
FLOP_MOM_TEST, FLOP_MOM_TEST_EXP, FLOP_MOM_EST = 32, 45, 161
FLOP_CONE_FILTER = 30
# AvplBvh rebuild, algorithmic HBM bytes per leaf: codes, perm, sorted
# position / normal, nodes, parents, ranges, node data and random numbers
# (276).  The light cuts' rebuild: AVPL input (52), codes, perm, sorted
# position, nodes, parents, ranges (72), merge sums (48 per slot, 2 slots per
# leaf) and node records (64 per slot): 348.
BYTES_BVH_LEAF, BYTES_CUT_BVH_LEAF = 276, 348
BYTES_SIMPLE_LEAF = 80  # SimpleBvh: codes, perm, sorted positions, nodes, parents, ranges
PHASES = ("upload", "bvh", "radiance", "antiradiance", "combine", "total")


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--avpls", type=int, default=1_000_000)
    ap.add_argument("--tris", type=int, default=200_000)
    ap.add_argument("--error-stride", type=int, default=4,
                    help="FP64 error on every k-th pixel in x and y (0 = skip)")
    ap.add_argument("--cpu-sample", type=int, default=1024,
                    help="CPU baseline on 1/k of the pixels, 1/(16 k) for its one-thread runs (0 = skip)")
    ap.add_argument("--cut", choices=("error-bound", "solid-angle"), default="error-bound",
                    help="light cut of the radiance: the analytic error bound or kernel_bvh's solid angle")
    ap.add_argument("--error-bound", type=float, default=0.01,
                    help="relative error bound of the cut (the workload's threshold is 1 %%)")
    ap.add_argument("--estimator", choices=("moments", "centroid"), default="moments",
                    help="error-bound cut: how a kept node is summed (second-order moments, or the centroid)")
    ap.add_argument("--criterion", choices=("bound", "estimate", "expansion"), default="bound",
                    help="error-bound cut: the Lightcuts contribution bound (the workload's analytic bound), the "
                         "centroid estimator's error estimate, or the moment expansion's")
    ap.add_argument("--image-floor", type=float, default=0.0,
                    help="error-bound cut: threshold floor as a fraction of the image-mean estimate (0 = none)")
    ap.add_argument("--exact-count", type=int, default=16,
                    help="moment estimator: kept nodes of at most this many AVPLs are summed exactly")
    ap.add_argument("--opening", type=float, default=0.3, help="expansion criterion: largest s / d of a kept node")
    ap.add_argument("--antiradiance", choices=("cone", "disk"), default="cone",
                    help="antiradiance model: 10-degree cones (Gather.frag, the workload's) or disks")
    ap.add_argument("--per-tile", type=int, default=1, help="error-bound cut: one cut per 8x8 tile (1) or per pixel (0)")
    ap.add_argument("--eval-depth", type=int, default=0,
                    help="centroid estimator: a kept node is summed as its descendants this many levels down (0-4)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU baseline: timed runs (median reported)")
    ap.add_argument("--accurate", default="0.0045,4,128",
                    help="a second setting of the same frame, timed and checked against FP64 beside the headline: "
                         "errorBound,cutImageFloor,cutExactCount of the moment estimator with the Lightcuts bound "
                         "('' = skip)")
    ap.add_argument("--estimate-depth", type=int, default=6,
                    help="error-bound cut: depth of the fixed cut that estimates each pixel for its threshold")
    ap.add_argument("--component", choices=("frame", "trace", "batch", "learned", "lbvh", "exact"), default="frame",
                    help="frame: the gather frame (the headline metric); trace: the GPU AVPL shooter and "
                         "G-buffer ray caster on the same scene; batch: many views in one graph launch; "
                         "learned: the learned cut against the error-bound cut; lbvh: the 1M-point LBVH build; "
                         "exact: the 256x256 x 4096 exact gather (side measurements)")
    ap.add_argument("--paths", type=int, default=250_000, help="trace: paths shot per step")
    ap.add_argument("--views", type=int, default=64, help="batch: views per graph launch")
    ap.add_argument("--lanes", type=int, default=4, help="batch: streams the views are dealt over")
    ap.add_argument("--batch-seed", type=int, default=7, help="batch: seed of the view sizes")
    ap.add_argument("--tune", type=int, default=1,
                    help="autotune the frame's (or each batched view's) schedule before timing (1) or not (0)")
    ap.add_argument("--schedule", default="",
                    help="cut_stack,anti_chunk,estimate_pool: a fixed schedule instead of the tuner's "
                         "(e.g. 96,1536,32 estimates the kept nodes on the matrix cores)")
    return ap.parse_args(argv)


def frame_params(args):
    """The gather parameters of the frame the options describe (the
    workload as written by default)."""
    from antiradiancecuts_amd import ffi
    from antiradiancecuts_amd.runtime import default_params

    eb = args.cut == "error-bound"
    mom = eb and args.estimator == "moments"
    crit = {"bound": ffi.ARC_CRITERION_BOUND, "estimate": ffi.ARC_CRITERION_ESTIMATE,
            "expansion": ffi.ARC_CRITERION_EXPANSION}[args.criterion]
    return default_params(useLightCuts=1,
                          antiradianceModel=ffi.ARC_ANTIRAD_CONE if args.antiradiance == "cone" else ffi.ARC_ANTIRAD_DISK,
                          antiradianceMode=ffi.ARC_ANTI_CLUSTERED,
                          cutMode=ffi.ARC_CUT_ERROR_BOUND if eb else ffi.ARC_CUT_SOLID_ANGLE,
                          errorBound=args.error_bound, cutPerTile=args.per_tile, cutEvalDepth=args.eval_depth,
                          cutEstimateDepth=args.estimate_depth, cutCriterion=crit, cutImageFloor=args.image_floor,
                          cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS if mom else ffi.ARC_ESTIMATOR_CENTROID,
                          cutExactCount=args.exact_count, cutOpening=args.opening)


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the gather shards by pixel tiles with no exchange; see the design notes)", file=sys.stderr)
        return 2
    if args.component == "trace":
        return bench_trace(args)
    if args.component == "batch":
        return bench_batch(args)
    if args.component == "learned":
        return bench_learned(args)
    if args.component == "lbvh":
        return bench_lbvh(args)
    if args.component == "exact":
        return bench_exact(args)
    import torch

    from antiradiancecuts_amd import ffi
    from antiradiancecuts_amd import scene as S
    from antiradiancecuts_amd.runtime import Context, DeviceAvpls, Gather, default_params, device_array

    t0 = time.time()
    sc = S.clutter_1080p(width=args.width, height=args.height, num_avpls=args.avpls, num_tris=args.tris)
    t_scene = time.time() - t0
    ctx = Context(0)
    g = Gather(ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    eb = args.cut == "error-bound"
    mom = eb and args.estimator == "moments"
    cone = args.antiradiance == "cone"
    params = frame_params(args)
    flags = ffi.ARC_RUN_TIMELINE | ffi.ARC_TRUST_INPUT
    # the schedule: launch shapes and estimate pool, fixed or tuned on the frame (untimed)
    from antiradiancecuts_amd import autotune
    tune_table = None
    if args.schedule:
        stack, chunk, pool = (int(v) for v in args.schedule.split(","))
        sched = autotune.Schedule(stack, chunk, pool)
        sched.apply(g)
    elif args.tune:
        params, sched, tune_table = autotune.tune(g, av, sc.cam, sc.extent, params)
    else:
        sched = autotune.Schedule()
        sched.apply(g)

    for _ in range(args.warmup):
        g.run(av, sc.cam, sc.extent, params, flags)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        g.run(av, sc.cam, sc.extent, params, flags)
    stop.record()
    torch.cuda.synchronize()
    ms_per_step = start.elapsed_time(stop) / args.steps
    hist = np.zeros((args.steps, 6), np.float32)
    got = g.lib.arc_gather_profile_history(g.h, args.steps, hist.ctypes.data)
    hist = hist[:max(got, 0)]
    phase_ms = {k: float(hist[:, i].mean()) for i, k in enumerate(PHASES)} if len(hist) else {}
    frame_ms = hist[:, 5] if len(hist) else np.array([ms_per_step])
    # the two main kernels of each timed frame, timed with HIP events on the
    # library's stream right around their launches
    kern = np.zeros((args.steps, 2), np.float32)
    got = g.lib.arc_gather_kernel_history(g.h, args.steps, kern.ctypes.data)
    kern = kern[:max(got, 0)]
    kernel_ms = {"radiance": float(kern[:, 0].mean()), "antiradiance": float(kern[:, 1].mean())} if len(kern) else {}

    # untimed: work counters and the error against the exact FP64 gather
    st = g.run(av, sc.cam, sc.extent, params, ffi.ARC_RUN_COUNTERS | ffi.ARC_TRUST_INPUT)
    torch.cuda.synchronize()
    result = g.result[:, :3].double().cpu().numpy()
    err = None
    ref_pix = None
    if args.error_stride > 0:
        k = args.error_stride
        ys, xs = np.arange(0, sc.height, k), np.arange(0, sc.width, k)
        pix = (ys[:, None] * sc.width + xs[None, :]).ravel()
        sub = S.Scene(**{**sc.__dict__, "width": len(xs), "height": len(ys), "pos4": sc.pos4[pix].copy(),
                         "nrm4": sc.nrm4[pix].copy(), "albedo4": sc.albedo4[pix].copy()})
        gs = Gather(ctx, sub.width, sub.height, sub.materials)
        gs.set_gbuffer(device_array(sub.pos4), device_array(sub.nrm4), device_array(sub.albedo4))
        t1 = time.time()
        r64, a64 = gs.exact_f64(av, sc.cam, sc.extent, params)
        torch.cuda.synchronize()
        t_f64 = time.time() - t1
        ref = (r64 - a64).cpu().numpy()
        diff = result[pix] - ref
        ref_pix = (pix, ref)
        err = {"rel_l2": float(np.linalg.norm(diff) / np.linalg.norm(ref)),
               "max_abs": float(np.abs(diff).max()),
               "max_rel_vs_rms": float(np.abs(diff).max() / np.sqrt(np.mean(ref ** 2))),
               "ref_norm_rms": float(np.sqrt(np.mean(ref ** 2))),
               "pixels": int(len(pix)), "sample": f"every {k}th pixel in x and y", "f64_seconds": round(t_f64, 2)}
        gs.close()

    # roofline of the dominant kernel: its algorithmic FLOPs (from the
    # kernel's own counters) over its HIP-event duration; the LBVH rebuild,
    # many small HBM-bound kernels, is taken as one unit.  The cut and the
    # clustered antiradiance run on two streams; when they overlap for a good
    # part of their time, each stretched by the other, they are taken
    # together: their FLOPs over the union of their spans in the frame.
    npix, n = sc.npix, sc.num_avpls
    spans = np.zeros((args.steps, 4), np.float32)
    got = g.lib.arc_gather_kernel_spans(g.h, args.steps, spans.ctypes.data)
    spans = spans[:max(got, 0)]
    flop_anti = float(st.antiEvaluatedNodes) * (FLOP_CONE_PAIR if cone else FLOP_DISK_PAIR) + \
        float(st.antiVisitedNodes) * (FLOP_CONE_FILTER if cone else FLOP_TILE_FILTER)
    if mom:
        exact = float(st.exactLeaves)
        flop_rad = float(st.testedNodes) * (FLOP_MOM_TEST_EXP if args.criterion == "expansion" else FLOP_MOM_TEST) \
            + (float(st.evaluatedNodes) - exact) * FLOP_MOM_EST + exact * FLOP_RAD_PAIR
        rad_name = "k_cut_mom_tile"
    elif eb:
        flop_rad = float(st.testedNodes) * FLOP_EB_TEST + float(st.evaluatedNodes) * FLOP_EB_EST
        rad_name = "k_cut_eb_tile" if args.per_tile else "k_cut_eb_px"
    else:
        flop_rad = float(st.visitedNodes) * FLOP_VISIT + float(st.evaluatedNodes) * (FLOP_EVAL + FLOP_RAD_PAIR)
        rad_name = "k_cut_sa"

    def rate(flop, ms):
        a = flop / (ms * 1e-3) / 1e12
        return {"achieved": round(a, 3), "frac": round(a / PEAK_FP32_TFLOPS, 4), "ms": round(ms, 4)}

    t_bvh = phase_ms.get("bvh", 0.0)
    t_pair = float(np.mean(np.maximum(spans[:, 1], spans[:, 3]) - np.minimum(spans[:, 0], spans[:, 2]))) \
        if len(spans) else 0.0
    t_overlap = float(np.mean(np.maximum(0.0, np.minimum(spans[:, 1], spans[:, 3]) -
                                         np.maximum(spans[:, 0], spans[:, 2])))) if len(spans) else 0.0
    t_rad, t_anti = kernel_ms.get("radiance", 0.0), kernel_ms.get("antiradiance", 0.0)
    # the pair is one unit when each kernel spends a good part of its time
    # beside the other; otherwise the longer of the two is the unit alone
    paired = t_overlap > 0.25 * min(t_rad, t_anti)
    per_kernel = {rad_name: {**rate(flop_rad, t_rad), "flop": flop_rad, "overlap_ms": round(t_overlap, 4)},
                  "k_anti_clustered": {**rate(flop_anti, t_anti), "flop": flop_anti}}
    if not paired and max(t_rad, t_anti) >= t_bvh and max(t_rad, t_anti) > 0.0:
        dom = "radiance" if t_rad >= t_anti else "antiradiance"
        name, flop, t_dom = (rad_name, flop_rad, t_rad) if dom == "radiance" else ("k_anti_clustered", flop_anti, t_anti)
        ach = flop / (t_dom * 1e-3) / 1e12
        roof = {"kernel": name, "bound": COMPUTE_BOUND, "engine": "VALU", "achieved": round(ach, 3),
                "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                "traffic": profiled_traffic(dom, eb, bool(args.per_tile), args.eval_depth, mom, cone), "ms": round(t_dom, 4),
                "kernels": per_kernel}
    elif t_pair >= t_bvh and t_pair > 0.0:
        ach = (flop_rad + flop_anti) / (t_pair * 1e-3) / 1e12
        tr = [profiled_traffic(k, eb, bool(args.per_tile), args.eval_depth, mom, cone) for k in ("radiance", "antiradiance")]
        roof = {"kernel": f"{rad_name} + k_anti_clustered (concurrent on two streams)", "bound": COMPUTE_BOUND,
                "engine": "VALU",
                "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                "traffic": float(sum(tr)) if all(t is not None for t in tr) else None,
                "ms": round(t_pair, 4), "kernels": per_kernel}
    else:
        bytes_ = float(BYTES_CUT_BVH_LEAF if eb else BYTES_BVH_LEAF) * n
        ach = bytes_ / (t_bvh * 1e-3) / 1e9
        roof = {"kernel": "AvplBvh rebuild (sort, Karras, refit)", "bound": "hbm", "achieved": round(ach, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 4),
                "traffic": profiled_traffic("bvh", eb, bool(args.per_tile)), "ms": round(t_bvh, 4)}

    accurate = None
    if args.accurate and eb and mom:
        eb2, fl2, ec2 = (float(v) for v in args.accurate.split(","))
        p2 = default_params(**{f: getattr(params, f) for f, _ in params._fields_})
        p2.errorBound, p2.cutImageFloor, p2.cutExactCount, p2.cutBudget = eb2, fl2, int(ec2), 1 << 20
        for _ in range(args.warmup):
            g.run(av, sc.cam, sc.extent, p2, flags)
        torch.cuda.synchronize()
        start.record()
        for _ in range(args.steps):
            g.run(av, sc.cam, sc.extent, p2, flags)
        stop.record()
        torch.cuda.synchronize()
        ms2 = start.elapsed_time(stop) / args.steps
        h2 = np.zeros((args.steps, 6), np.float32)
        got2 = g.lib.arc_gather_profile_history(g.h, args.steps, h2.ctypes.data)
        f2 = h2[:max(got2, 0), 5] if got2 > 0 else np.array([ms2])
        g.run(av, sc.cam, sc.extent, p2, ffi.ARC_TRUST_INPUT)
        torch.cuda.synchronize()
        e2 = None
        if ref_pix is not None:
            pix2, ref2 = ref_pix
            d2 = g.result[:, :3].double().cpu().numpy()[pix2] - ref2
            e2 = float(np.linalg.norm(d2) / np.linalg.norm(ref2))
        accurate = {"setting": {"criterion": "bound", "error_bound": eb2, "image_floor": fl2,
                                "exact_count": int(ec2), "estimator": "moments"},
                    "ms_per_frame": round(ms2, 4), "ms_per_frame_p50": round(float(np.percentile(f2, 50)), 4),
                    "ms_per_frame_p99": round(float(np.percentile(f2, 99)), 4),
                    "value": npix * float(n) / (ms2 * 1e-3), "rel_l2_vs_f64": e2,
                    "meets_1e-3": bool(e2 is not None and e2 <= 1e-3)}

    cpu = cpu_baseline(sc, params, args.cpu_sample, args.cpu_runs) if args.cpu_sample > 0 else None

    value = npix * float(n) / (ms_per_step * 1e-3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pixel-AVPL pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "ms_per_frame_p50": round(float(np.percentile(frame_ms, 50)), 4),
        "ms_per_frame_p99": round(float(np.percentile(frame_ms, 99)), 4),
        "evaluated_pairs_per_s": (float(st.evaluatedNodes) + float(st.antiEvaluatedNodes)) / (ms_per_step * 1e-3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"clutter_{sc.width}x{sc.height}: Cornell walls + {args.tris} random triangles, "
                        f"{n} AVPLs (bounce 1-4, L log-uniform [1e-3,10], 30% antiradiance); AvplBvh rebuilt "
                        "per frame; light-cut radiance ("
                        + (f"{args.criterion} criterion at {args.error_bound:g}"
                           + (f", image floor {args.image_floor:g}" if args.image_floor > 0 else ", no image floor")
                           + (f", second-order moment estimates, nodes of <= {args.exact_count} AVPLs exact, one cut "
                              "per 8x8 tile" if mom else
                              f", centroid estimates, {'per 8x8 tile' if args.per_tile else 'per pixel'}, kept nodes "
                              f"summed {args.eval_depth} levels down")
                           if eb else "solid angle < pi/10")
                        + f"); {'10-degree cone' if cone else 'disk'} antiradiance over per-cluster AVPL lists (exact)",
            "width": sc.width, "height": sc.height, "avpls": n, "parallelism": "single GPU",
            "phase_ms": {k: round(v, 4) for k, v in phase_ms.items()},
            "kernel_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
            "schedule": {**dataclasses.asdict(sched), "tuned": tune_table is not None,
                         "tried_ms": {c: (round(ms, 3) if ms is not None else None) for c, ms in tune_table}
                         if tune_table else None},
            "cut": {"mode": args.cut, "error_bound": args.error_bound if eb else None,
                    "criterion": args.criterion if eb else None, "estimator": args.estimator if eb else None,
                    "image_floor": args.image_floor if eb else None,
                    "exact_count": args.exact_count if mom else None,
                    "opening": args.opening if mom and args.criterion == "expansion" else None,
                    "eval_depth": args.eval_depth if eb and not mom else None,
                    "estimate_depth": args.estimate_depth if eb else None,
                    "visited_nodes": int(st.visitedNodes), "tested_nodes": int(st.testedNodes),
                    "evaluated_nodes": int(st.evaluatedNodes), "exact_leaves": int(st.exactLeaves),
                    "stack_overflows": int(st.errorPixels), "budget_hits": int(st.budgetHits)},
            "antiradiance_model": "cone 10 deg (Gather.frag:149-188)" if cone else "disk (cudaGather.cu:97-144)",
            "antiradiance": {"pair_tests": int(st.antiEvaluatedNodes), "tile_filter_tests": int(st.antiVisitedNodes)},
            "scene_build_s": round(t_scene, 2),
        },
        "error_vs_f64": err,
        "accuracy_target": {"rel_l2": 1e-3, "met": bool(err is not None and err["rel_l2"] <= 1e-3)},
        "accurate_setting": accurate,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    g.close()
    ctx.close()
    return 0


# the PMC summary of the benchmark frame (moment estimates, cone antiradiance)
TRAFFIC_PROFILE = ROOT / "profiles" / "frame_1080p_1M_cone_moments_traffic.txt"
# LBVH kernels of one rebuild, with their dispatch counts (4 radix passes)
BVH_KERNELS = {"k_bounds_partial": 1, "k_bounds_final": 1, "k_encode": 1, "k_hist": 4, "k_scan_digit": 4,
               "k_scatter": 4, "k_stage_avpls": 1, "k_permute_cut": 1, "k_karras": 1,
               "k_refit_local<2>": 1, "k_refit_mid<2>": 1, "k_refit_top<2>": 1}


def profiled_traffic(dom: str, eb: bool, per_tile: bool, eval_depth: int = 0, mom: bool = False, cone: bool = False):
    """HBM bytes per launch of the dominant kernel on the 1080p x 1M frame,
    from the committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE scaled
    by a plain copy's readings); None when the summary does not cover it."""
    if not TRAFFIC_PROFILE.exists():
        return None
    table = {}
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("#") or "\t" not in line:
            continue
        name, _, _, b = line.split("\t")
        table[name] = float(b)
    if dom == "antiradiance":
        # the kernel's name, with or without its list-block template argument
        stem = f"k_anti_clustered<{1 if cone else 0}, false"
        names = {next((k for k in table if k.startswith(stem)), stem + ">"): 1}
    elif dom == "radiance":
        name = "k_cut_mom_tile" if mom else ("k_cut_eb_tile" if per_tile else "k_cut_eb_px") if eb else "k_cut_sa"
        own = next((k for k in table if k.startswith(f"{name}<false")), None) if mom else None
        if own:  # the traversal that estimates its kept nodes itself (any stack depth)
            name = own
        elif eb and per_tile and f"{name}<{eval_depth}>" in table:
            name = f"{name}<{eval_depth}>"
        names = {name: 1}
    else:
        names = BVH_KERNELS
    if not all(k in table for k in names if names[k] == 1 and k != "k_bounds_final"):
        return None
    return float(sum(table.get(k, 0.0) * c for k, c in names.items()))


def bench_trace(args) -> int:
    """AvplShooter::shoot and the G-buffer on the GPU, on the frame's scene
    (Cornell walls + 200K random triangles, both faces: 400,012 triangles):
    `steps` shoots of --paths paths, then `steps` 1080p G-buffer casts."""
    import math

    import torch

    from antiradiancecuts_amd import scene as S
    from antiradiancecuts_amd.runtime import Context, TraceScene

    ctx = Context(0)
    d = S.desc(2, args.width, args.height, 0, 1, 4, 1e-3, 10.0, True, 0.3, False, args.tris, 1e-4, 1e-2)
    ts = S.triangle_scene(d)
    sc = TraceScene(ctx, ts)
    for k in range(args.warmup):
        sc.shoot(args.paths, 1000 + k)
    torch.cuda.synchronize()
    t0 = time.time()
    n = 0
    for k in range(args.steps):
        n += sc.shoot(args.paths, k).struct.numAvpls
    torch.cuda.synchronize()
    shoot_ms = (time.time() - t0) * 1e3 / args.steps
    cam = (0.5, 0.5, -0.49 / math.tan(math.pi / 8) / max(args.width / args.height, 1.0))
    for _ in range(args.warmup):
        sc.gbuffer(cam, (0, 0, 1), (0, 1, 0), math.pi / 4, args.width, args.height)
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(args.steps):
        sc.gbuffer(cam, (0, 0, 1), (0, 1, 0), math.pi / 4, args.width, args.height)
    torch.cuda.synchronize()
    gb_ms = (time.time() - t0) * 1e3 / args.steps
    line = {"metric": "AVPLs/s shot on the GPU (side measurement; not the headline metric)",
            "value": n / args.steps / (shoot_ms * 1e-3), "unit": "AVPLs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(shoot_ms, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"AvplShooter::shoot, {args.paths} paths, {len(ts.material)} triangles "
                                   f"(tree depth {sc.info()[1]}); G-buffer {args.width}x{args.height}",
                       "avpls_per_step": n / args.steps, "gbuffer_ms": round(gb_ms, 3),
                       "gbuffer_rays_per_s": args.width * args.height / (gb_ms * 1e-3)}}
    print(json.dumps(line), flush=True)
    sc.close()
    ctx.close()
    return 0


def bench_batch(args) -> int:
    """Batched views in one HIP graph: --views per-view seeded variants of the
    frame workload (Cornell walls + --tris random triangles), each with a
    width and height drawn uniformly from [256, 1920] x [256, 1080] (rounded
    to 8) and an AVPL count log-uniform over [16K, 1M], at the frame's
    setting.  Each view's schedule (the estimate-cut depth; for the centroid
    estimator also the cut per tile or per pixel) is picked by the autotuner
    on the view itself, and the batch's (its views' moment cuts as one launch
    over all their tiles, or one launch per view) on the batch; all views are
    captured into one graph over --lanes streams and a step is one
    hipGraphLaunch of the whole batch.  Per-view ms are the same views run
    alone (HIP events, median of 3)."""
    import torch

    from antiradiancecuts_amd import autotune, ffi
    from antiradiancecuts_amd import scene as S
    from antiradiancecuts_amd.runtime import Batch, Context, DeviceAvpls, Gather, default_params, device_array

    rng = np.random.default_rng(args.batch_seed)
    ctx = Context(0)
    b = Batch(ctx, lanes=args.lanes)
    base = frame_params(args)
    views, alone_ms, pairs, tuned = [], [], 0.0, {}
    t0 = time.time()
    for k in range(args.views):
        w = int(rng.integers(256, 1921)) // 8 * 8
        h = int(rng.integers(256, 1081)) // 8 * 8
        n = int(round(math.exp(rng.uniform(math.log(16384), math.log(1_000_000)))))
        sc = S.clutter_1080p(seed=1000 + k, width=w, height=h, num_avpls=n, num_tris=args.tris)
        g = Gather(b.lane(k), w, h, sc.materials)
        g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
        av = DeviceAvpls.from_scene(sc)
        torch.cuda.synchronize()
        p = base
        if args.tune:
            p, sched, table = autotune.tune(g, av, sc.cam, sc.extent, base)
            key = (f"tile={p.cutPerTile},d0={p.cutEstimateDepth},stack={sched.cut_stack},"
                   f"chunk={sched.anti_chunk},pool={sched.estimate_pool}")
            tuned[key] = tuned.get(key, 0) + 1
        ms = [g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_PROFILE | ffi.ARC_TRUST_INPUT).msTotal for _ in range(3)]
        alone_ms.append(float(np.median(ms)))
        pairs += float(w * h) * n
        views.append((g, av, sc.cam, sc.extent, p))
        print(f"[batch] view {k}: {w}x{h}, {n} AVPLs, alone {alone_ms[-1]:.3f} ms", file=sys.stderr, flush=True)
    t_setup = time.time() - t0
    one_launch, batch_table = autotune.tune_batch(b, views) if args.tune else (True, None)
    if not args.tune:
        b.capture(views)
    for _ in range(args.warmup):
        b.launch()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        b.launch()
    stop.record()
    torch.cuda.synchronize()
    ms_per_step = start.elapsed_time(stop) / args.steps
    alone = np.array(alone_ms)
    line = {"metric": "aggregate pixel-AVPL pairs/s of a batch of views in one graph launch (side measurement)",
            "value": pairs / (ms_per_step * 1e-3), "unit": "pixel-AVPL pairs/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{args.views} views, W,H uniform [256,1920]x[256,1080], AVPLs log-uniform "
                                   f"[16K,1M], Cornell walls + {args.tris} random triangles per view (seeded "
                                   f"variants), {args.cut} cut {args.error_bound:g} ({args.criterion} criterion, "
                                   f"{args.estimator} estimates, image floor {args.image_floor:g}, nodes of <= "
                                   f"{args.exact_count} AVPLs exact), {args.antiradiance} antiradiance, {args.lanes} lanes",
                       "views": args.views, "lanes": args.lanes, "pairs_per_batch": pairs,
                       "view_ms_alone_p50": round(float(np.percentile(alone, 50)), 4),
                       "view_ms_alone_p99": round(float(np.percentile(alone, 99)), 4),
                       "sum_view_ms_alone": round(float(alone.sum()), 3),
                       "tuned_schedules": tuned if args.tune else None,
                       "one_cut_launch": one_launch, "batch_schedule_ms": batch_table,
                       "setup_s": round(t_setup, 1)}}
    print(json.dumps(line), flush=True)
    ctx.sync()
    for g, *_ in views:
        g.close()
    b.close()
    ctx.close()
    return 0


def bench_learned(args) -> int:
    """The learned cut (committed weights) and the error-bound tile cut on the
    frame workload at error bounds 0.003, 0.01 and 0.03: radiance-kernel ms
    (HIP events, median of --steps frames), frame ms and the result's
    relative L2 against the exact FP64 gather on every 4th pixel in x and y;
    then, per learned point, the error-bound cut's time at the same error."""
    from antiradiancecuts_amd import learned_cut as L
    from antiradiancecuts_amd.runtime import Context

    ctx = Context(0)
    net = L.load(ctx)
    out = L.evaluate(ctx, net, width=args.width, height=args.height, avpls=args.avpls, tris=args.tris,
                     reps=max(args.steps, 1), log=lambda s: print(s, file=sys.stderr, flush=True))
    eq = [r for r in out["equal_error"] if r["speedup"] is not None]
    line = {"metric": "learned-cut speedup over the error-bound cut at equal result rel-L2 (side measurement)",
            "value": eq[0]["speedup"] if eq else None, "unit": "x", "n_gpus": 1, "steps": args.steps,
            "warmup": 2, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32",
            "data": "synthetic", "config": {"workload": out["workload"], "weights": str(L.DEFAULT_WEIGHTS.name),
                                            "points": out["rows"], "equal_error": out["equal_error"]}}
    print(json.dumps(line), flush=True)
    net.close()
    ctx.close()
    return 0


def _timed(fn, steps: int, warmup: int) -> float:
    """ms per call on torch's current stream (the library's context follows it)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_lbvh(args) -> int:
    """SimpleBvh over 1,000,000 points (half uniform, half a 32-Gaussian
    mixture, sigma 0.01): bounds, Morton codes, radix sort, Karras tree and
    refit per step, inputs in HBM.  The CPU baseline is the oracle's build
    of the same points (one thread)."""
    from antiradiancecuts_amd import scene as S
    from antiradiancecuts_amd.runtime import Bvh, Context, device_array

    ctx = Context(0)
    pos = S.points(1_000_000, seed=1)
    dpos = device_array(pos)
    b = Bvh(ctx)
    ms = _timed(lambda: b.build_points(dpos), args.steps, args.warmup)
    n = len(pos)
    ach = BYTES_SIMPLE_LEAF * n / (ms * 1e-3) / 1e9
    cpu = None
    if args.cpu_sample > 0:
        from oracle import pyoracle as orc

        t0 = time.time()
        orc.lbvh_build(pos)
        t = time.time() - t0
        cpu = {"value": n / t, "unit": "points/s", "cores": 1, "kind": "port", "ms_per_build": round(t * 1e3, 1),
               "sample": "the oracle's full build of the same 1,000,000 points"}
    line = {"metric": "LBVH build of 1M points (side measurement)", "value": n / (ms * 1e-3), "unit": "points/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+u32", "data": "synthetic",
            "config": {"workload": "SimpleBvh, 1,000,000 points: half uniform, half 32 Gaussians (sigma 0.01)"},
            "roofline": {"kernel": "SimpleBvh build (bounds, encode, sort, Karras, refit)", "bound": "hbm",
                         "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None, "ms": round(ms, 4)},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    b.close()
    ctx.close()
    return 0


def bench_exact(args) -> int:
    """The exact gather (kernel_radiance_simple + kernel_antiradiance) on the
    256x256 Cornell scene with 4096 AVPLs and 10-degree antiradiance cones:
    every pixel sums every AVPL.  The CPU baseline is the oracle (FP32,
    OpenMP) on the same frame."""
    import torch

    from antiradiancecuts_amd import ffi
    from antiradiancecuts_amd import scene as S
    from antiradiancecuts_amd.runtime import Context, DeviceAvpls, Gather, default_params, device_array, photon_radius

    ctx = Context(0)
    sc = S.cornell_256(seed=0)
    g = Gather(ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    p = default_params(useLightCuts=0, antiradianceMode=ffi.ARC_ANTI_EXACT, antiradianceModel=ffi.ARC_ANTIRAD_CONE)
    ms = _timed(lambda: g.run(av, sc.cam, sc.extent, p, ffi.ARC_TRUST_INPUT), args.steps, args.warmup)
    g.run(av, sc.cam, sc.extent, p)
    r64, a64 = g.exact_f64(av, sc.cam, sc.extent, p)
    torch.cuda.synchronize()
    ref = (r64 - a64).cpu().numpy()
    res = g.result[:, :3].double().cpu().numpy()
    err = float(np.linalg.norm(res - ref) / np.linalg.norm(ref))
    pairs = float(sc.npix) * sc.num_avpls
    flop = pairs * (FLOP_RAD_PAIR + FLOP_CONE_PAIR)
    ach = flop / (ms * 1e-3) / 1e12
    cpu = None
    if args.cpu_sample > 0:
        from oracle import pyoracle as orc

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ref32 = orc.SceneRef(sc, photon_radius(p, sc.extent), model=ffi.ARC_ANTIRAD_CONE,
                             cone_half_angle=float(p.coneHalfAngle))
        t0 = time.time()
        orc.gather_exact(ref32, 32)
        t = time.time() - t0
        cpu = {"value": pairs / t, "unit": "pixel-AVPL pairs/s", "cores": threads, "kind": "port",
               "ms_per_frame": round(t * 1e3, 1), "sample": "the oracle's REF32 exact gather of the whole frame"}
    line = {"metric": "exact gather pixel-AVPL pairs/s, 256x256 x 4096 (side measurement)",
            "value": pairs / (ms * 1e-3), "unit": "pixel-AVPL pairs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "cornell_256: 256x256 G-buffer, 4096 AVPLs (seed 0), exact radiance + exact "
                                   "cone antiradiance, Kahan-compensated", "rel_l2_vs_f64": err},
            "roofline": {"kernel": "k_exact (radiance and antiradiance pairs)", "bound": COMPUTE_BOUND,
                         "engine": "VALU",
                         "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None, "ms": round(ms, 4)},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    g.close()
    ctx.close()
    return 0


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, params, k: int, runs: int = 5):
    """The oracle (C++ restatement of the reference path, REF32) on this
    host's cores: the full AvplBvh build (bvh.cu:746-808, one thread), then
    kernel_bvh's solid-angle cut radiance (cudaGather.cu:316-444) and the
    exact antiradiance of the frame's model on a pixel sample, extrapolated
    to the frame.  The median of `runs` runs with all the host's worker
    threads on every k-th pixel, and of `runs` runs with one thread on every
    (16 k)-th pixel."""
    from oracle import pyoracle as orc
    from antiradiancecuts_amd.runtime import photon_radius

    L = orc.lib()
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", avail))
    rand = np.random.default_rng(0).random(2 * sc.num_avpls - 1, dtype=np.float32)
    ref = orc.SceneRef(sc, photon_radius(params, sc.extent), model=params.antiradianceModel,
                       cone_half_angle=params.coneHalfAngle)

    def frame_s(pix, nthreads):
        L.orc_set_threads(nthreads)
        t0 = time.time()
        tree = orc.avpl_bvh_build(sc, rand)
        t_bvh = time.time() - t0
        t0 = time.time()
        orc.gather_cut_solid_angle(ref, tree, params.clusterRefinementThreshold, 32, pix)
        orc.gather_exact(ref, 32, pix, what=2)
        t_pix = time.time() - t0
        return t_bvh + t_pix * sc.npix / len(pix), t_bvh + t_pix

    t_all = time.time()
    pix_t = np.arange(0, sc.npix, k, dtype=np.int32)
    pix_1 = np.arange(0, sc.npix, 16 * k, dtype=np.int32)
    multi = [frame_s(pix_t, threads) for _ in range(runs)]
    single = [frame_s(pix_1, 1) for _ in range(runs)]
    L.orc_set_threads(threads)
    f_t = float(np.median([m[0] for m in multi]))
    f_1 = float(np.median([m[0] for m in single]))
    return {"value": sc.npix * float(sc.num_avpls) / f_t, "unit": "pixel-AVPL pairs/s", "cores": threads,
            "kind": "port", "ms_per_frame": round(f_t * 1e3, 1),
            "value_1_thread": sc.npix * float(sc.num_avpls) / f_1, "ms_per_frame_1_thread": round(f_1 * 1e3, 1),
            "cpu_model": _cpu_model(), "nproc": avail, "runs": runs, "statistic": "median",
            "sample": f"per run: AvplBvh build of all {sc.num_avpls} AVPLs (1 thread) + solid-angle cut radiance "
                      f"and exact {'cone' if params.antiradianceModel == 1 else 'disk'} antiradiance on every "
                      f"{k}th pixel ({len(pix_t)} px, {threads} threads) or every {16 * k}th pixel "
                      f"({len(pix_1)} px, 1 thread), extrapolated to the {sc.npix}-pixel frame; "
                      f"{runs} + {runs} runs took {time.time() - t_all:.1f} s"}


if __name__ == "__main__":
    sys.exit(main())
