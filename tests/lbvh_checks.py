"""Independent LBVH checks shared by the oracle tests and the GPU tests.

None of these restate the reference's bottom-up construction: the tree is
rebuilt top-down from the highest differing bit of the augmented key
(code << 32 | slot), and boxes / cluster data are checked against brute force
over each node's leaf range.
"""
from __future__ import annotations

import bisect

import numpy as np


def topdown_tree(codes: np.ndarray):
    """Radix tree over sorted codes by recursive splitting.

    Node numbering follows Karras 2012: the root is internal node 0; an
    internal node covering [first, last] split after slot s has its left
    child at index s (leaf s when s == first) and its right child at s + 1
    (leaf s + 1 when s + 1 == last).  Returns left, right, parents, ranges
    in the encoding of bvh.h:19-20 (leaf >= 0, internal k as -(k + 1)).
    """
    n = len(codes)
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]
    left = np.zeros(n - 1, np.int64)
    right = np.zeros(n - 1, np.int64)
    parents = np.full(2 * n - 1, -2, np.int64)
    ranges = np.zeros((n - 1, 2), np.int64)
    parents[n] = -1
    stack = [(0, 0, n - 1)]
    while stack:
        k, first, last = stack.pop()
        ranges[k] = (first, last)
        x = keys[first] ^ keys[last]
        bit = x.bit_length() - 1
        # keys in [first, last] share every bit above `bit`; those with the bit
        # clear come first: the split is the last of them
        hi_flag = [(keys[i] >> bit) & 1 for i in range(first, last + 1)]
        s = first + bisect.bisect_left(hi_flag, 1) - 1
        if s == first:
            left[k] = s
            parents[s] = k
        else:
            left[k] = -(s + 1)
            parents[n + s] = k
            stack.append((s, first, s))
        if s + 1 == last:
            right[k] = s + 1
            parents[s + 1] = k
        else:
            right[k] = -(s + 2)
            parents[n + s + 1] = k
            stack.append((s + 1, s + 1, last))
    return left, right, parents, ranges


def check_topology(codes, nodes, parents, ranges=None):
    left, right, par, rng = topdown_tree(codes)
    np.testing.assert_array_equal(nodes["left"], left)
    np.testing.assert_array_equal(nodes["right"], right)
    np.testing.assert_array_equal(parents, par)
    if ranges is not None:
        np.testing.assert_array_equal(np.asarray(ranges).reshape(-1, 2), rng)


def check_karras_ranges(ranges):
    """Each internal node k has k as one end of its range, ranges are proper."""
    r = np.asarray(ranges).reshape(-1, 2)
    k = np.arange(len(r))
    assert np.all((r[:, 0] == k) | (r[:, 1] == k))
    assert np.all(r[:, 0] < r[:, 1])
    assert r[0, 0] == 0 and r[0, 1] == len(r)


def check_boxes(nodes, ranges, sorted_pos):
    """Boxes equal min / max over the leaf range, bit for bit."""
    r = np.asarray(ranges).reshape(-1, 2)
    sp = np.asarray(sorted_pos, np.float32)
    # prefix-free: per node slice (fine for the sizes tested here)
    for k, (lo, hi) in enumerate(r):
        seg = sp[lo:hi + 1]
        np.testing.assert_array_equal(nodes["bbMin"][k], seg.min(0), err_msg=f"node {k}")
        np.testing.assert_array_equal(nodes["bbMax"][k], seg.max(0), err_msg=f"node {k}")


def check_cluster(t: dict, rad_leaf_sorted: np.ndarray):
    """AvplBvh node data: sizes are leaf counts, radiance sums match the leaf
    sum (float tolerance), each representative is a leaf of the node's range
    and the representative of one of its two children."""
    n = len(t["codes"])
    r = np.asarray(t["ranges"]).reshape(-1, 2)
    size = t["nd_size"]
    np.testing.assert_array_equal(size[:n], 1)
    np.testing.assert_array_equal(size[n:], r[:, 1] - r[:, 0] + 1)
    csum = np.concatenate([[np.zeros(3)], np.cumsum(rad_leaf_sorted.astype(np.float64), 0)])
    want = csum[r[:, 1] + 1] - csum[r[:, 0]]
    got = t["nd_rad"][n:].astype(np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-6)
    # representative = a leaf inside the range, and the representative of a child
    nodes = t["nodes"]
    rec = np.concatenate([t["nd_pos"], t["nd_nrm"], t["nd_dir"], t["nd_mat"][:, None].astype(np.float32)], 1)
    for k in range(n - 1):
        lo, hi = r[k]
        assert np.any((rec[lo:hi + 1] == rec[n + k]).all(1)), f"node {k}: representative outside its range"
        kids = [c if c >= 0 else n + (-(c + 1)) for c in (nodes["left"][k], nodes["right"][k])]
        assert any(np.array_equal(rec[c], rec[n + k]) for c in kids), f"node {k}"


def right_choice_stats(t: dict):
    """For each internal node: (rand, i2 / (i1 + i2), chose_right) with i the
    child radiance norms, to test that the right child is picked with
    probability i2 / (i1 + i2) (bvh.cu:732-736)."""
    n = len(t["codes"])
    nodes = t["nodes"]
    rows = []
    for k in range(n - 1):
        kids = [c if c >= 0 else n + (-(c + 1)) for c in (nodes["left"][k], nodes["right"][k])]
        i1 = float(np.linalg.norm(t["nd_rad"][kids[0]].astype(np.float64)))
        i2 = float(np.linalg.norm(t["nd_rad"][kids[1]].astype(np.float64)))
        same_l = np.array_equal(t["nd_pos"][kids[0]], t["nd_pos"][n + k]) and np.array_equal(
            t["nd_dir"][kids[0]], t["nd_dir"][n + k])
        same_r = np.array_equal(t["nd_pos"][kids[1]], t["nd_pos"][n + k]) and np.array_equal(
            t["nd_dir"][kids[1]], t["nd_dir"][n + k])
        if same_l == same_r:  # indistinguishable children
            continue
        rows.append((float(t["nd_rand"][n + k]) if "nd_rand" in t else np.nan,
                     i2 / (i1 + i2) if i1 + i2 > 0 else np.nan, same_r))
    return np.array(rows, dtype=np.float64)
