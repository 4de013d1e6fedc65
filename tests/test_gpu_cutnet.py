"""The learned cut: its MLP kernels against the NumPy restatement, the
training step's gradient, the exact-label traversal's edge cases, and the
learned traversal against the error-bound and exact gathers.

What pins the learned traversal without a trained network: a network whose
output is a constant p (all weights 0, b3 = logit(p)) makes every decision
one way.  p = 0.5 is never confident, so every node goes to the error-bound
criterion and the image equals the error-bound tile cut's; p ~ 1 refines
every node, so the image equals the exact gather (FP64 image, 1e-5).
"""
import math

import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import CutNet, DeviceAvpls, Gather, default_params, device_array
from tests import cutnet_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(gpu_ctx):
    n = CutNet(gpu_ctx)
    yield n
    n.close()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_forward_matches_reference(gpu_ctx, net):
    rng = np.random.default_rng(5)
    w = R.random_weights(rng)
    X = rng.normal(0, 1, (1003, R.IN)).astype(np.float32)
    net.set_weights(w)
    p = net.forward(device_array(X)).cpu().numpy()
    ref = R.forward(w, X)
    d = np.abs(p - ref)
    print(f"[cutnet] forward: max |p - ref| {d.max():.2e}, mean {d.mean():.2e}", flush=True)
    assert d.max() <= 5e-3 and d.mean() <= 2e-4
    np.testing.assert_array_equal(net.get_weights(), w)


def test_one_adam_step_follows_the_gradient(gpu_ctx, net):
    rng = np.random.default_rng(6)
    w = R.random_weights(rng)
    n = 512
    X = rng.normal(0, 1, (n, R.IN)).astype(np.float32)
    y = (rng.random(n) < 0.4).astype(np.float32)
    net.set_weights(w)
    lr = 1e-4
    net.train(device_array(X), device_array(y), epochs=1, batch=n, lr=lr, seed=3)
    dw = net.get_weights().astype(np.float64) - w
    g = R.grad(w, X, y)
    # Adam's first step is -lr * g / (|g| + eps): the sign of every clear gradient entry
    agree = total = 0
    for lo, hi in ((R.W1, R.B1), (R.B1, R.W2), (R.W2, R.B2), (R.B2, R.W3), (R.W3, R.B3), (R.B3, R.B3 + 1)):
        gs, ds = g[lo:hi], dw[lo:hi]
        clear = np.abs(gs) > 1e-2 * np.abs(gs).max()
        agree += int(np.sum(np.sign(ds[clear]) == -np.sign(gs[clear])))
        total += int(clear.sum())
        assert np.all(np.abs(ds) <= lr * 1.001)
    print(f"[cutnet] Adam step signs: {agree}/{total}", flush=True)
    assert agree >= 0.99 * total
    assert np.all(dw[:R.W1] == 0)                     # feature normalisation is fixed
    assert np.all(dw[R.W1 + R.IN * R.HID:R.B1] == 0)  # padded input rows stay zero


def test_training_separates_a_linear_rule(gpu_ctx, net):
    rng = np.random.default_rng(7)
    n = 20000
    X = rng.normal(0, 1, (n, R.IN)).astype(np.float32)
    X[:, 5] = X[:, 5] * 3 + 2
    y = ((X[:, 0] + 0.5 * X[:, 3] - (X[:, 5] - 2) / 3) > 0.2).astype(np.float32)
    net.init(11, X.mean(0), X.std(0))
    Xd, yd = device_array(X), device_array(y)
    loss = net.train(Xd, yd, epochs=10, batch=1024, lr=3e-3, seed=1)
    p = net.forward(Xd).cpu().numpy()
    acc = float(np.mean((p > 0.5) == (y > 0.5)))
    print(f"[cutnet] loss per epoch {np.round(loss, 4).tolist()}, accuracy {acc:.4f}", flush=True)
    assert loss[-1] < 0.5 * loss[0]
    assert acc >= 0.97


# ------------------------------------------------------------ traversal --

@pytest.fixture(scope="module")
def rig(gpu_ctx):
    sc = S.cornell_256(seed=3, width=96, height=80, num_avpls=4096)
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    r64, _ = g.exact_f64(av, sc.cam, sc.extent, default_params())
    yield sc, g, av, r64.cpu().numpy()
    g.close()


def const_net(net, prob):
    w = np.zeros(R.PARAMS, np.float32)
    w[R.ISTD:R.ISTD + R.IN] = 1.0
    w[R.B3] = math.log(prob / (1.0 - prob))
    net.set_weights(w)


def learned(**kw):
    return default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_LEARNED, **kw)


def test_learned_cut_needs_a_network(gpu_ctx, rig):
    sc, g, av, _ = rig
    g.set_cutnet(None)
    with pytest.raises(ffi.ArcError) as e:
        g.run(av, sc.cam, sc.extent, learned())
    assert e.value.status == ffi.ARC_E_STATE


def test_unconfident_network_gives_the_error_bound_cut(gpu_ctx, net, rig):
    sc, g, av, _ = rig
    const_net(net, 0.5)
    g.set_cutnet(net)
    st = g.run(av, sc.cam, sc.extent, learned(errorBound=0.01), ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    rad_l = g.radiance[:, :3].double().cpu().numpy()
    st_eb = g.run(av, sc.cam, sc.extent, default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND,
                                                        errorBound=0.01, cutPerTile=1), ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    rad_e = g.radiance[:, :3].double().cpu().numpy()
    e = rel_l2(rad_l, rad_e)
    print(f"[learned] p=0.5: fallbacks {st.fallbackNodes}, evaluated {st.evaluatedNodes} vs {st_eb.evaluatedNodes}, "
          f"rel-L2 vs error-bound tile cut {e:.2e}", flush=True)
    assert st.fallbackNodes > 0
    assert st.evaluatedNodes == st_eb.evaluatedNodes and st.visitedNodes == st_eb.visitedNodes
    assert e <= 1e-5


def test_always_refine_is_exact(gpu_ctx, net, rig):
    sc, g, av, r64 = rig
    const_net(net, 1.0 - 1e-9)
    g.set_cutnet(net)
    st = g.run(av, sc.cam, sc.extent, learned(), ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    e = rel_l2(g.radiance[:, :3].double().cpu().numpy(), r64)
    print(f"[learned] p=1: rel-L2 vs FP64 exact {e:.2e}, evaluated/px {st.evaluatedNodes / sc.npix:.0f}", flush=True)
    assert st.fallbackNodes == 0 and st.errorPixels == 0
    assert st.evaluatedNodes == sc.npix * sc.num_avpls  # every leaf, nothing else
    assert e <= 1e-5


def test_always_keep_evaluates_the_root(gpu_ctx, net, rig):
    sc, g, av, _ = rig
    const_net(net, 1e-9)
    g.set_cutnet(net)
    st = g.run(av, sc.cam, sc.extent, learned(), ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    assert st.fallbackNodes == 0 and st.visitedNodes == 0
    assert st.evaluatedNodes == sc.npix
    assert np.all(np.isfinite(g.radiance.cpu().numpy()))


def test_labelled_traversal_edges(gpu_ctx, rig):
    sc, g, av, _ = rig
    tiles = math.ceil(sc.width / 8) * math.ceil(sc.height / 8)
    # a huge bound keeps the root everywhere: one row per tile, label 0
    X, y, cnt = g.cut_samples(av, sc.cam, sc.extent, default_params(errorBound=1e6), capacity=10 * tiles)
    yy = y.cpu().numpy()
    assert cnt == tiles and np.all(yy == 0)
    assert np.all(np.isfinite(X.cpu().numpy()))
    # no tolerance: every node that contributes is refined, down to the leaves
    p0 = default_params(errorBound=0.0, cutImageFloor=0.0)
    X, y, cnt = g.cut_samples(av, sc.cam, sc.extent, p0, capacity=tiles * (sc.num_avpls - 1))
    yy = y.cpu().numpy()
    print(f"[labels] exact: {cnt} rows over {tiles} tiles, refine share {yy.mean():.3f}", flush=True)
    assert cnt <= tiles * (sc.num_avpls - 1)
    assert yy.mean() > 0.5
    # keep = 0.25 keeps about a quarter of the same rows
    _, _, c4 = g.cut_samples(av, sc.cam, sc.extent, p0, capacity=1, keep=0.25, seed=9)
    assert abs(c4 / cnt - 0.25) < 0.02


def test_committed_network_runs_the_learned_cut(gpu_ctx):
    """The shipped weights (trained on the cluttered-scene family) load,
    decide a share of the nodes with confidence and give an image of a small
    scene of that family within a loose bound of the exact one (the learned
    cut trades error for nodes; its accuracy is measured on the frame
    workload)."""
    from antiradiancecuts_amd import learned_cut as L

    sc = S.clutter_1080p(seed=77, width=128, height=96, num_avpls=32768, num_tris=20000)
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    r64, _ = g.exact_f64(av, sc.cam, sc.extent, default_params())
    net = L.load(gpu_ctx)
    g.set_cutnet(net)
    st = g.run(av, sc.cam, sc.extent, learned(errorBound=0.003), ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    e = rel_l2(g.radiance[:, :3].double().cpu().numpy(), r64.cpu().numpy())
    decisions = (st.visitedNodes + st.evaluatedNodes) / 64.0
    print(f"[learned] committed weights: rel-L2 {e:.3e}, evaluated/px {st.evaluatedNodes / sc.npix:.1f}, "
          f"fallback share {st.fallbackNodes / decisions:.3f}", flush=True)
    assert 0 < st.fallbackNodes < decisions
    assert e < 0.2
    g.close()
    net.close()


def test_committed_network_on_the_full_frame(gpu_ctx):
    """The shipped weights on the 1920x1080 x 1M-AVPL frame (the inputs the
    network's training scenes were drawn like): the learned cut runs, decides
    a share of the nodes itself and falls back to the criterion for the rest,
    and its radiance stays within a loose bound of the FP64 exact gather on
    every 8th pixel; error and time beside the error-bound cut at the same
    errorBound are printed (the equal-error comparison is bench.py's)."""
    import time

    from antiradiancecuts_amd import learned_cut as L

    sc = S.clutter_1080p()
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    k = 8
    ys, xs = np.arange(0, sc.height, k), np.arange(0, sc.width, k)
    pix = (ys[:, None] * sc.width + xs[None, :]).ravel()
    sub = S.Scene(**{**sc.__dict__, "width": len(xs), "height": len(ys), "pos4": sc.pos4[pix].copy(),
                     "nrm4": sc.nrm4[pix].copy(), "albedo4": sc.albedo4[pix].copy()})
    gs = Gather(gpu_ctx, sub.width, sub.height, sub.materials)
    gs.set_gbuffer(device_array(sub.pos4), device_array(sub.nrm4), device_array(sub.albedo4))
    r64, _ = gs.exact_f64(av, sc.cam, sc.extent, default_params())
    gpu_ctx.sync()
    r64 = r64.cpu().numpy()
    gs.close()
    net = L.load(gpu_ctx)
    g.set_cutnet(net)
    out = {}
    for name, p in (("learned", learned(errorBound=0.01)),
                    ("error-bound", default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, errorBound=0.01,
                                                   cutPerTile=1))):
        g.run(av, sc.cam, sc.extent, p)
        gpu_ctx.sync()
        t0 = time.time()
        st = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_COUNTERS)
        gpu_ctx.sync()
        ms = (time.time() - t0) * 1e3
        e = rel_l2(g.radiance[:, :3].double().cpu().numpy()[pix], r64)
        out[name] = (e, st)
        print(f"[learned] full frame, {name} cut at 0.01: radiance rel-L2 {e:.3e}, evaluated/px "
              f"{st.evaluatedNodes / sc.npix:.1f}, about {ms:.1f} ms with counters", flush=True)
    e, st = out["learned"]
    decisions = (st.visitedNodes + st.evaluatedNodes) / 64.0
    assert 0 < st.fallbackNodes < decisions
    assert e < 0.2
    g.close()
    net.close()
