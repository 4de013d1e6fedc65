"""Host-side pieces that need no GPU: the schedule tuner's candidate set,
the benchmark's reading of the committed traffic profile, and the learned
cut's equal-error interpolation."""
import math

import pytest

from antiradiancecuts_amd import ffi


def params(**kw):
    p = ffi.arc_gather_params()
    ffi.lib().arc_gather_params_default(p)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_tuner_candidates_cover_the_schedule_knobs():
    from antiradiancecuts_amd import autotune

    base = params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, errorBound=0.07, cutEvalDepth=3)
    cands = autotune.candidates(base)
    assert {(c.cutPerTile, c.cutEstimateDepth) for c in cands} == {(t, d) for t in (0, 1) for d in (4, 6, 8)}
    # the accuracy knobs are the caller's
    assert all(math.isclose(c.errorBound, 0.07, rel_tol=1e-6) and c.cutEvalDepth == 3 for c in cands)
    assert base.cutEstimateDepth == 6  # copies, not the base itself
    assert len(autotune.candidates(params(useLightCuts=0))) == 1
    # the moment cut: one parameter set, three launch-shape knobs
    mom = params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS,
                 antiradianceMode=ffi.ARC_ANTI_CLUSTERED)
    assert len(autotune.candidates(mom)) == 1
    axes = dict(autotune.schedule_axes(mom))
    assert axes == {"cut_stack": (96, 64, 128), "estimate_pool": (0, 32), "anti_chunk": (1536, 768, 3072)}
    # each axis starts at the library's default
    d = autotune.Schedule()
    assert all(getattr(d, k) == v[0] for k, v in axes.items())
    assert autotune.schedule_axes(params(useLightCuts=0, antiradianceMode=ffi.ARC_ANTI_EXACT)) == []


def test_bench_reads_the_traffic_profile():
    import bench

    if not bench.TRAFFIC_PROFILE.exists():
        pytest.skip("no traffic profile committed")
    # the setting the committed traffic profile covers: the moment
    # estimator and cone antiradiance; other settings have no figure
    anti = bench.profiled_traffic("antiradiance", True, True, 0, True, True)
    cut = bench.profiled_traffic("radiance", True, True, 0, True, True)
    bvh = bench.profiled_traffic("bvh", True, True, 0)
    assert anti and cut and bvh
    assert cut > anti > bvh  # the cut's node reads dominate the frame's bytes
    assert bench.profiled_traffic("radiance", True, True, 1, False, False) is None
    assert bench.profiled_traffic("antiradiance", True, True, 0, True, False) is None


def test_equal_error_interpolates_in_log_log():
    from antiradiancecuts_amd.learned_cut import equal_error

    rows = [{"mode": "error-bound", "eps": 0.01, "rel_l2": 0.01, "cut_ms": 8.0},
            {"mode": "error-bound", "eps": 0.1, "rel_l2": 0.04, "cut_ms": 2.0},
            {"mode": "learned", "eps": 0.01, "rel_l2": 0.02, "cut_ms": 2.0},
            {"mode": "learned", "eps": 0.1, "rel_l2": 0.08, "cut_ms": 1.0}]
    out = equal_error(rows)
    # 0.02 is halfway between 0.01 and 0.04 in log: time sqrt(8 * 2) = 4 ms
    assert math.isclose(out[0]["error_bound_cut_ms_at_equal_error"], 4.0, rel_tol=1e-9)
    assert math.isclose(out[0]["speedup"], 2.0, rel_tol=1e-9)
    assert out[1]["speedup"] is None  # beyond the error-bound points
