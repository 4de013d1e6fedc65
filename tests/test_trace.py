"""Ray tracing of the triangle scene: Triangle::intersect (FRONT_FACE), the
closest hit of Scene::IntersectRayScene, the AVPL shooter
(AvplShooter::createPath) and the G-buffer ray caster.

CPU: the oracle's triangle test against closed forms.  GPU: the BVH closest
hit against the oracle's brute force over every triangle (same triangle and
the same FP32 distance bit for bit: both follow the reference's operation
order without contraction); the shooter against the oracle's replay of the
same paths with the same counter-based random numbers (the library math
functions acos / pow / sin / cos may differ by an ulp between the GPU and
the host, and a path diverges once such a difference tips a ray across a
triangle edge: every path's first AVPL must agree, most whole paths agree,
and the AVPL population's statistics agree); AVPL invariants; the G-buffer
against the host ray caster of the scene library."""
import ctypes as C
import math

import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S

vp = C.c_void_p


def _p(a):
    return a.ctypes.data_as(vp)


def f3(*v):
    return np.array(v, np.float32)


# ---------------------------------------------------------------- oracle --
@pytest.fixture(scope="module")
def tri():
    # floor triangle in y = 0 facing up: normal = normalize(cross(p2 - p0, p1 - p0)) = +y
    return f3(0, 0, 0), f3(1, 0, 0), f3(0, 0, 1)


def test_oracle_head_on_hit(orc, tri):
    t = np.zeros(1, np.float32)
    assert orc.lib().orc_tri_intersect(*map(_p, tri), _p(f3(0.25, 2.0, 0.25)), _p(f3(0, -1, 0)), 1e-4, _p(t)) == 1
    assert t[0] == pytest.approx(2.0, rel=1e-6)


def test_oracle_back_face_misses(orc, tri):
    assert orc.lib().orc_tri_intersect(*map(_p, tri), _p(f3(0.25, -2.0, 0.25)), _p(f3(0, 1, 0)), 1e-4, None) == 0


def test_oracle_outside_misses(orc, tri):
    for o in (f3(0.6, 1, 0.6), f3(-0.1, 1, 0.5), f3(0.5, 1, -0.1)):
        assert orc.lib().orc_tri_intersect(*map(_p, tri), _p(o), _p(f3(0, -1, 0)), 1e-4, None) == 0


def test_oracle_det_floor_rejects_small_triangles(orc):
    """The reference's |det| < EPS test is absolute: a triangle of area 2e-5
    seen head-on has det = 4e-5 < 1e-4 and is missed; with no floor it is hit."""
    s = math.sqrt(4e-5)
    p0, p1, p2 = f3(0, 0, 0), f3(s, 0, 0), f3(0, 0, s)
    o, d = f3(s / 4, 1, s / 4), f3(0, -1, 0)
    assert orc.lib().orc_tri_intersect(_p(p0), _p(p1), _p(p2), _p(o), _p(d), 1e-4, None) == 0
    assert orc.lib().orc_tri_intersect(_p(p0), _p(p1), _p(p2), _p(o), _p(d), 0.0, None) == 1


def test_oracle_closest_skips_emissive(orc):
    p0 = np.array([[0, 0, 0], [0, 1, 0]], np.float32)
    p1 = np.array([[1, 0, 0], [1, 1, 0]], np.float32)
    p2 = np.array([[0, 0, 1], [0, 1, 1]], np.float32)
    mat = np.array([1, 2], np.int32)
    emis = np.zeros(256, np.int32)
    t = np.zeros(1, np.float32)
    o, d = f3(0.2, 2, 0.2), f3(0, -1, 0)
    L = orc.lib()
    assert L.orc_trace_closest(2, _p(p0), _p(p1), _p(p2), _p(mat), _p(emis), _p(o), _p(d), 1e-4, _p(t)) == 1
    assert t[0] == pytest.approx(1.0)
    emis[2] = 1  # the nearer triangle is a light: the ray ends with no hit
    assert L.orc_trace_closest(2, _p(p0), _p(p1), _p(p2), _p(mat), _p(emis), _p(o), _p(d), 1e-4, _p(t)) == -1


def test_scene_triangles_face_inwards():
    ts = S.triangle_scene(S.desc(0, 8, 8, 0, 1, 3, 0.0, 1.0, False, 0.3, True, 0))
    n = np.cross(ts.p2 - ts.p0, ts.p1 - ts.p0)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    np.testing.assert_allclose(n[:2], [[0, 1, 0]] * 2, atol=1e-6)     # floor
    np.testing.assert_allclose(n[-2:], [[0, -1, 0]] * 2, atol=1e-6)   # light
    assert (ts.material[-2:] == len(ts.materials) - 1).all()
    assert ts.materials[-1, :3].any()  # emissive


# ------------------------------------------------------------------- GPU --
def _random_rays(ts, n, seed):
    rng = np.random.default_rng(seed)
    org = rng.uniform(0.02, 0.98, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org, d.astype(np.float32)


def _oracle_closest(orc, ts, org, dirs, det_eps=1e-4):
    emis = np.zeros(256, np.int32)
    for k, m in enumerate(ts.materials):
        emis[k] = int((m[:3] ** 2).sum() > 0)
    L = orc.lib()
    tri = np.zeros(len(org), np.int32)
    tt = np.zeros(len(org), np.float32)
    t = np.zeros(1, np.float32)
    for i in range(len(org)):
        tri[i] = L.orc_trace_closest(len(ts.material), _p(ts.p0), _p(ts.p1), _p(ts.p2), _p(ts.material), _p(emis),
                                     _p(org[i]), _p(dirs[i]), det_eps, _p(t))
        tt[i] = t[0]
    return tri, tt


@pytest.mark.gpu
@pytest.mark.parametrize("name,tris", [("cornell", 0), ("clutter", 20_000)])
def test_closest_hit_vs_brute_force(gpu_ctx, orc, name, tris):
    from antiradiancecuts_amd.runtime import TraceScene

    d = S.desc(3, 8, 8, 0, 1, 3, 0.0, 1.0, False, 0.3, tris == 0, tris, 1e-4, 1e-2)
    ts = S.triangle_scene(d)
    sc = TraceScene(gpu_ctx, ts)
    ntri, depth = sc.info()
    print(f"[trace] {name}: {ntri} triangles, tree depth {depth}", flush=True)
    org, dirs = _random_rays(ts, 3000, 11)
    t, tri = sc.intersect(org, dirs)
    gpu_ctx.sync()
    t, tri = t.cpu().numpy(), tri.cpu().numpy()
    otri, ot = _oracle_closest(orc, ts, org, dirs)
    same = tri == otri
    print(f"[trace] same triangle {same.mean():.5f}, hits {(otri >= 0).mean():.3f}", flush=True)
    assert same.mean() >= 0.999
    np.testing.assert_array_equal(t[same & (tri >= 0)], ot[same & (tri >= 0)])
    # a differing choice can only be a tie in distance (rays through shared edges)
    diff = ~same
    if diff.any():
        np.testing.assert_allclose(t[diff], ot[diff], rtol=1e-6)
    sc.close()


def _light_vector(ts):
    L = ts.light
    s, u = ts.light_axes
    return np.concatenate([L["center"], L["front"] / np.linalg.norm(L["front"]), s, u,
                           [0.5 * L["width"], 0.5 * L["height"]], L["radiance"], [L["material"]]]).astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("explicit", [1, 0])
def test_shooter_vs_oracle_replay(gpu_ctx, orc, explicit):
    from antiradiancecuts_amd.runtime import TraceScene, avpls_to_host

    d = S.desc(5, 8, 8, 0, 1, 3, 0.0, 1.0, False, 0.3, True, 2000, 1e-3, 1e-2)
    ts = S.triangle_scene(d)
    sc = TraceScene(gpu_ctx, ts)
    paths, seed = 600, 1234
    flags = 0 if explicit else ffi.ARC_SHOOT_NO_EXPLICIT_DIRECT
    g = avpls_to_host(sc.shoot(paths, seed, flags), gpu_ctx)
    n = len(g["bounce"])
    print(f"[trace] {paths} paths -> {n} indirect AVPLs", flush=True)
    lv = _light_vector(ts)
    counts = np.zeros(paths, np.int32)
    cap = 4 * n + 1000
    o = {k: np.zeros((cap, 3), np.float32) for k in ("pos", "nrm", "dir", "rad", "anti")}
    om, ob = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    total = orc.lib().orc_trace_shoot(len(ts.material), _p(ts.p0), _p(ts.p1), _p(ts.p2), _p(ts.material),
                                      _p(ts.materials), len(ts.materials), _p(lv), seed, 0, paths, explicit,
                                      _p(counts), _p(o["pos"]), _p(o["nrm"]), _p(o["dir"]), _p(o["rad"]),
                                      _p(o["anti"]), _p(om), _p(ob), cap)
    # The GPU writes each path's AVPLs together, in path order, and every
    # path starts with a bounce-1 AVPL; paths without a hit write nothing.
    starts = np.flatnonzero(g["bounce"] == 1)
    ostarts = np.concatenate([[0], np.cumsum(counts)])[:-1][counts > 0]
    ends_g = np.concatenate([starts[1:], [n]])
    ends_o = np.concatenate([ostarts[1:], [total]])
    agree = 0
    for k in range(min(len(starts), len(ostarts))):
        if ends_g[k] - starts[k] != ends_o[k] - ostarts[k]:
            continue  # diverged (an edge-grazing ray, last-ulp math differences)
        sl_g, sl_o = slice(starts[k], ends_g[k]), slice(ostarts[k], ends_o[k])
        ok = (np.allclose(g["pos"][sl_g], o["pos"][sl_o], rtol=1e-4, atol=1e-4) and
              np.array_equal(g["mat"][sl_g], om[sl_o]) and np.array_equal(g["bounce"][sl_g], ob[sl_o]) and
              np.allclose(g["rad"][sl_g], o["rad"][sl_o], rtol=1e-3, atol=1e-6) and
              np.allclose(g["anti"][sl_g], o["anti"][sl_o], rtol=1e-3, atol=1e-6))
        agree += ok
    first = 0  # every path's first AVPL: the light sample and the first direction
    for k in range(min(len(starts), len(ostarts))):
        first += np.allclose(g["pos"][starts[k]], o["pos"][ostarts[k]], rtol=1e-5, atol=1e-5) and \
            g["mat"][starts[k]] == om[ostarts[k]]
    first /= max(1, len(ostarts))
    print(f"[trace] first AVPLs identical {first:.4f}", flush=True)
    assert first >= 0.99
    # statistics of the whole population (paths diverge chaotically once an ulp tips a ray across an edge)
    assert abs(g["bounce"].mean() - ob[:total].mean()) < 0.15 * ob[:total].mean()
    assert abs(g["rad"].mean() - o["rad"][:total].mean()) < 0.2 * o["rad"][:total].mean()
    frac = agree / max(1, len(ostarts))
    print(f"[trace] paths with AVPLs: gpu {len(starts)}, oracle {len(ostarts)}; identical paths {frac:.4f}", flush=True)
    assert abs(len(starts) - len(ostarts)) <= max(2, len(ostarts) // 100)
    assert frac >= 0.6
    assert abs(n - total) <= max(5, total // 50)
    # invariants of every AVPL
    assert (g["bounce"] >= 1).all()
    np.testing.assert_allclose(np.linalg.norm(g["nrm"], axis=1), 1.0, atol=1e-5)
    assert (np.einsum("ij,ij->i", g["nrm"], g["dir"]) < 0).all()  # front-face hits
    if explicit:
        assert not g["anti"][g["bounce"] == 1].any()
    sc.close()


@pytest.mark.gpu
def test_shoot_then_gather(gpu_ctx):
    """AVPLs from the shooter feed the gather directly (device to device)."""
    from antiradiancecuts_amd.runtime import Gather, TraceScene, default_params, device_array

    sc0 = S.cornell_256(seed=1, width=64, height=48, num_avpls=16)
    d = S.desc(1, 64, 48, 16, 1, 3, 0.0, 1.0, False, 0.3, True, 0)
    ts = S.triangle_scene(d)
    tsc = TraceScene(gpu_ctx, ts)
    av = tsc.shoot(4000, 7)
    assert av.struct.numAvpls > 4000
    g = Gather(gpu_ctx, sc0.width, sc0.height, sc0.materials)
    g.set_gbuffer(device_array(sc0.pos4), device_array(sc0.nrm4), device_array(sc0.albedo4))
    g.run(av, sc0.cam, sc0.extent, default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND),
          ffi.ARC_TRUST_INPUT)
    gpu_ctx.sync()
    res = g.result[:, :3].cpu().numpy()
    assert np.isfinite(res).all() and np.abs(res).max() > 0
    g.close()
    tsc.close()


@pytest.mark.gpu
def test_gbuffer_vs_host_caster(gpu_ctx):
    from antiradiancecuts_amd.runtime import TraceScene

    d = S.desc(2, 160, 120, 0, 1, 3, 0.0, 1.0, False, 0.3, True, 3000, 1e-3, 1e-2)
    host = S.make_scene(d)
    ts = S.triangle_scene(d, light_quad=False)
    sc = TraceScene(gpu_ctx, ts)
    pos4, nrm4 = sc.gbuffer(host.cam, (0, 0, 1), (0, 1, 0), math.pi / 4, d.width, d.height)
    gpu_ctx.sync()
    pos4, nrm4 = pos4.cpu().numpy(), nrm4.cpu().numpy()
    same = (nrm4[:, 3] == host.nrm4[:, 3]) & (np.abs(nrm4[:, :3] - host.nrm4[:, :3]).max(1) < 1e-4)
    print(f"[trace] G-buffer pixels on the same surface: {same.mean():.5f}", flush=True)
    assert same.mean() >= 0.995
    np.testing.assert_allclose(pos4[same], host.pos4[same], atol=2e-5)
    sc.close()
