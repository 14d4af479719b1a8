"""Pruning / densification oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, and only as the checker of tt_densify_plan /
tt_densify_apply; the product path never does.

The statistics it consumes are the reference forward's: max_blending (atomic max of
alpha * T per triangle, tritracer/forward.cu:138-143), density_factor (set to 1 when a
triangle is blended, forward.cu:134) and scaling (largest angle between the ray origin ->
centre and ray origin -> corner directions, forward.cu:34-74, computeScaling). The
reference ships no consumer of them (its callers prune and densify in Python), so the
rules below are this project's specification, restated here in numpy:

  plan   keep[i] = not (max_blending[i] < min_blending) and not (opacity[i] < min_opacity)
         cand[i] = keep[i] and scaling given and (density absent or density[i] != 0)
                   and scaling[i] > split_scaling and (score absent or score[i] > split_score)
         (with action_in: keep = action_in in {1, 2}, cand = action_in == 2)
         cap = max_out if 0 < max_out <= 2^31 - 1 else 2^31 - 1
         allowed = min(cand.sum(), max(0, (cap - keep.sum()) // 3))
         split[i] = cand[i] and (number of candidates before i) < allowed
         count[i] = keep[i] + 3 split[i]; offset = exclusive cumsum of count, offset[P] = total
  apply  outputs of triangle i occupy rows offset[i] .. offset[i] + count[i] - 1, in order;
         children of (a, b, c) with ab = (a + b) * 0.5 (float32), ...:
         (a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)

parity: pinned by tests/test_oracle_densify.py (hand-worked cases, invariants of the
subdivision: child areas, centroids, orientation, tiling of the parent).
"""
from __future__ import annotations

import numpy as np

DROP, KEEP, SPLIT = 0, 1, 2
ROWS_TRIANGLE, ROWS_VERTEX = 0, 1
SPLIT_COPY, SPLIT_ZERO, SPLIT_SUBDIVIDE = 0, 1, 2
INT_MAX = 2 ** 31 - 1

# (u, w) corner pairs of each child's three points; the point is (p_u + p_w) / 2
CHILDREN = (((0, 0), (0, 1), (2, 0)), ((0, 1), (1, 1), (1, 2)), ((2, 0), (1, 2), (2, 2)), ((0, 1), (1, 2), (2, 0)))


def plan(max_blending=None, opacity=None, density=None, scaling=None, score=None, action_in=None,
         min_blending=0.0, min_opacity=0.0, split_scaling=np.inf, split_score=0.0, max_out=0):
    """(action uint8 [P], offset int32 [P + 1])"""
    f32 = np.float32
    if action_in is not None:
        a = np.asarray(action_in, np.uint8)
        keep = (a == KEEP) | (a == SPLIT)
        cand = a == SPLIT
    else:
        mb = np.asarray(max_blending, f32)
        op = np.asarray(opacity, f32)
        keep = ~(mb < f32(min_blending)) & ~(op < f32(min_opacity))
        if scaling is None:
            cand = np.zeros_like(keep)
        else:
            cand = keep & (np.asarray(scaling, f32) > f32(split_scaling))
            if density is not None:
                cand &= np.asarray(density, f32) != 0
            if score is not None:
                cand &= np.asarray(score, f32) > f32(split_score)
    cap = max_out if 0 < max_out <= INT_MAX else INT_MAX
    n_keep, n_cand = int(keep.sum()), int(cand.sum())
    allowed = min(n_cand, max(0, (cap - n_keep) // 3)) if cap > n_keep else 0
    rank = np.cumsum(cand) - cand  # candidates before i
    split = cand & (rank < allowed)
    count = keep.astype(np.int64) + 3 * split
    offset = np.zeros(len(count) + 1, np.int64)
    np.cumsum(count, out=offset[1:])
    action = np.where(split, SPLIT, np.where(keep, KEEP, DROP)).astype(np.uint8)
    return action, offset.astype(np.int32)


def apply(offset, cumsum, arrays):
    """arrays: list of (src float32 ndarray, width, rows, on_split). Returns the output arrays
    ([P_out * width] float32 each), npts_out and cumsum_out."""
    offset = np.asarray(offset, np.int64)
    count = np.diff(offset)
    P_out = int(offset[-1])
    parent = np.repeat(np.arange(len(count)), count)
    child = np.arange(P_out) - offset[:-1][parent]
    split = count[parent] == 4
    outs = []
    for src, width, rows, on_split in arrays:
        src = np.asarray(src, np.float32).reshape(-1)
        if rows == ROWS_VERTEX:
            base = 3 * np.asarray(cumsum, np.int64)[parent]
            pts = src[base[:, None, None] + 3 * np.arange(3)[None, :, None] + np.arange(3)[None, None, :]]  # [n,3,3]
            out = pts.copy()
            if on_split == SPLIT_SUBDIVIDE:
                for k, corners in enumerate(CHILDREN):
                    sel = split & (child == k)
                    for v, (u, w) in enumerate(corners):
                        out[sel, v] = pts[sel, u] if u == w else (pts[sel, u] + pts[sel, w]) * np.float32(0.5)
        else:
            out = src.reshape(-1, width)[parent].copy()
        out = out.reshape(P_out, width)
        if on_split == SPLIT_ZERO:
            out[split] = 0.0
        outs.append(out.reshape(-1))
    return outs, np.full(P_out, 3, np.int32), (3 * np.arange(P_out)).astype(np.int32)
