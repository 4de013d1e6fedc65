"""Host-side FP64 checks of the light cut's node records against sums over
each node's leaf range (Karras ranges: node k covers sorted leaves
ranges[k, 0] .. ranges[k, 1]).

The leaves' own records (crec slots [0, n)) are the inputs: position r0.xyz,
unit normal r1.xyz, BRDF-weighted intensity I = r2.xyz, w = mean(I).  Each
node's sums come from numpy FP64 prefix sums; the GPU records must equal
them to FP32 rounding.
"""
import numpy as np


def leaf_inputs(crec: np.ndarray, n: int):
    x = crec[:n, 0, :3].astype(np.float64)
    nm = crec[:n, 1, :3].astype(np.float64)
    I = crec[:n, 2, :3].astype(np.float64)
    return x, nm, I


def range_sums(values: np.ndarray, ranges: np.ndarray) -> np.ndarray:
    """sum over each node's leaf range of the per-leaf rows of `values`
    (prefix sums in the values' precision)."""
    pre = np.concatenate([np.zeros((1,) + values.shape[1:], values.dtype), np.cumsum(values, axis=0)], axis=0)
    return pre[ranges[:, 1] + 1] - pre[ranges[:, 0]]


def check_centroid_records(crec: np.ndarray, ranges: np.ndarray, n: int):
    """The centroid records of the internal nodes (slots n + k): count, I,
    centroid sum(w x)/sum(w) and mean normal sum(w n)/sum(w)."""
    x, nm, I = leaf_inputs(crec, n)
    w = I.mean(1, keepdims=True)
    inner = crec[n:2 * n - 1]
    count = ranges[:, 1] - ranges[:, 0] + 1
    np.testing.assert_array_equal(inner[:, 3, 2].astype(np.int64), count)
    Is = range_sums(I, ranges)
    np.testing.assert_allclose(inner[:, 2, :3], Is, rtol=2e-5, atol=1e-6 * Is.max())
    np.testing.assert_allclose(inner[:, 0, 3], Is.max(1), rtol=2e-5, atol=1e-6 * Is.max())
    W = range_sums(w, ranges)[:, 0]
    live = W > 1e-6 * W.max()
    c = range_sums(w * x, ranges)[live] / W[live, None]
    np.testing.assert_allclose(inner[live, 0, :3], c, rtol=0, atol=2e-5 * (1 + np.abs(c).max()))
    m = range_sums(w * nm, ranges)[live] / W[live, None]
    np.testing.assert_allclose(inner[live, 1, :3], m, rtol=0, atol=5e-5)
    return int(live.sum())


def moment_sums(crec: np.ndarray, ranges: np.ndarray, n: int, ref=None):
    """Range sums of w, w x, w x x^T (x relative to `ref`, in extended
    precision so that small nodes' covariances keep their digits), I, I n and
    I n n^T per channel."""
    x, nm, I = leaf_inputs(crec, n)
    w = I.mean(1, keepdims=True)
    W = range_sums(w, ranges)[:, 0]
    xl = (x - (0 if ref is None else np.asarray(ref))).astype(np.longdouble)
    wl = w.astype(np.longdouble)
    Sx = range_sums(wl * xl, ranges)
    Sxx = range_sums((wl[:, :, None] * xl[:, :, None] * xl[:, None, :]).reshape(n, 9), ranges).reshape(-1, 3, 3)
    Is = range_sums(I, ranges)
    M = np.stack([range_sums(I[:, c:c + 1] * nm, ranges) for c in range(3)], 1)          # (k, c, 3)
    nn = (nm[:, :, None] * nm[:, None, :]).reshape(n, 9)
    T = np.stack([range_sums(I[:, c:c + 1] * nn, ranges).reshape(-1, 3, 3) for c in range(3)], 1)  # (k, c, 3, 3)
    return W, Sx, Sxx, Is, M, T


def check_moment_records(mom: np.ndarray, crec: np.ndarray, ranges: np.ndarray, n: int, nodes):
    """The moment estimator's records (13 float4 per internal node: (c, max
    I), (s, first, count, -), the four points' 12 floats, per channel
    (-M/8, 3I/128 + t_zz), then per channel t = 15T/128 as (t_xx - t_zz,
    t_yy - t_zz, 0, 2t_xy, 2t_xz, 2t_yz): the SH2 clamped cosine over 4, its
    zz term moved into the constant for a unit direction).  The
    positions are taken relative to the centre of the root's box (nodes[0]),
    as the library takes them."""
    ref = 0.5 * (nodes["bbMin"][0].astype(np.float64) + nodes["bbMax"][0].astype(np.float64))
    W, Sx, Sxx, Is, M, T = moment_sums(crec, ranges, n, ref)
    k = len(ranges)
    assert mom.shape == (k, 13, 4)
    first = mom[:, 1, 1].copy().view(np.int32)
    count = mom[:, 1, 2].copy().view(np.int32)
    np.testing.assert_array_equal(first, ranges[:, 0])
    np.testing.assert_array_equal(count, ranges[:, 1] - ranges[:, 0] + 1)
    scale = Is.max()
    np.testing.assert_allclose(mom[:, 0, 3], Is.max(1), rtol=1e-6, atol=1e-7 * scale)
    # per channel, over 4: 3/32 I_c + 15/32 T_zz, -1/2 M_c, 15/32 T_c with the
    # zz term taken off the diagonal (off-diagonals doubled)
    tf = mom[:, 8:13, :].reshape(k, 20)[:, :18].reshape(k, 3, 6)
    for c in range(3):
        Tc = 0.25 * 0.46875 * T[:, c]
        np.testing.assert_allclose(mom[:, 5 + c, 3], 0.25 * 0.09375 * Is[:, c] + Tc[:, 2, 2], rtol=1e-6,
                                   atol=1e-7 * scale)
        np.testing.assert_allclose(mom[:, 5 + c, :3], -0.125 * M[:, c], rtol=1e-6, atol=1e-7 * scale)
        want = np.stack([Tc[:, 0, 0] - Tc[:, 2, 2], Tc[:, 1, 1] - Tc[:, 2, 2], 0 * Tc[:, 2, 2],
                         2 * Tc[:, 0, 1], 2 * Tc[:, 0, 2], 2 * Tc[:, 1, 2]], 1)
        np.testing.assert_allclose(tf[:, c], want, rtol=1e-6, atol=1e-7 * scale)
    # position mean and covariance of the four points
    live = W > 1e-6 * W.max()
    Wl = W[live].astype(np.longdouble)
    cl = Sx[live] / Wl[:, None]
    S = (Sxx[live] / Wl[:, None, None] - cl[:, :, None] * cl[:, None, :]).astype(np.float64)
    c = (cl + np.asarray(ref, np.longdouble)).astype(np.float64)
    pts = mom[live, 2:5, :].reshape(-1, 4, 3).astype(np.float64)
    np.testing.assert_allclose(mom[live, 0, :3], c, rtol=0, atol=1e-6 * (1 + np.abs(c).max()))
    off = pts - c[:, None, :]
    cmax = 1 + np.abs(c).max()
    np.testing.assert_allclose(off.mean(1), 0, atol=1e-6 * cmax)
    cov = np.einsum("kqi,kqj->kij", off, off) / 4
    tr = np.trace(S, axis1=1, axis2=2)
    # the library's FP64 prefix differences carry about 2^-52 of the prefix
    # total, over W (noise); a rank-deficient covariance (2 or 3 AVPLs, a
    # flat node) turns noise e in its small directions into about sqrt(e tr)
    # in its factor
    pref = 64 * 2.2e-16 * float(np.abs(Sxx).max()) / W[live]
    tr0 = np.maximum(tr, 0)
    tol = 2e-3 * tr0 + 4e-7 * np.sqrt(tr0) * cmax + pref + 16 * np.sqrt(pref * tr0) + 1e-12
    assert np.all(np.abs(cov - S) <= tol[:, None, None] + 1e-6 * np.abs(S)), \
        float(np.max(np.abs(cov - S) - tol[:, None, None]))
    # the half-diagonal of the node's box
    e = nodes["bbMax"].astype(np.float64) - nodes["bbMin"].astype(np.float64)
    np.testing.assert_allclose(mom[:, 1, 0], 0.5 * np.sqrt((e * e).sum(1)), rtol=1e-6, atol=1e-12)
    return int(live.sum())
