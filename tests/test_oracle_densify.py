"""Pins of the pruning / densification oracle (oracle/densify.py), CPU only:
  - plan: hand-worked cases (thresholds, NaN, density and score gates, the cap, caller
    actions) and a sequential greedy restatement of the cap rule on random inputs;
  - apply: midpoint subdivision checked by geometry (each child has a quarter of the
    parent's area, the children's centroids average to the parent's, orientation is kept,
    the children tile the parent: interior sample points fall in exactly one child), copy /
    zero policies, identity when everything is kept, gathering through a shuffled cumsum.
"""
import numpy as np

import densify as D


def test_plan_hand_worked_thresholds():
    mb = np.float32([0.5, 0.0, 0.2, 0.01, np.nan, 0.3, 0.3, 0.3])
    op = np.float32([0.9, 0.9, 0.05, 0.9, 0.9, 0.9, 0.9, 0.9])
    sc = np.float32([0.2, 0.2, 0.2, 0.2, 0.2, 0.01, 0.2, 0.2])
    dn = np.float32([1, 1, 1, 1, 1, 1, 0, 1])
    # drop: 1 (mb 0 < 0.02), 2 (opacity 0.05 < 0.1), 3 (mb 0.01 < 0.02); 4 kept (NaN is not below)
    # split candidates: kept, blended, scaling > 0.1: 0, 4, 7 (5 too small, 6 not blended)
    a, off = D.plan(mb, op, dn, sc, min_blending=0.02, min_opacity=0.1, split_scaling=0.1)
    assert a.tolist() == [2, 0, 0, 0, 2, 1, 1, 2]
    assert off.tolist() == [0, 4, 4, 4, 4, 8, 9, 10, 14]
    # score gate: only 7 passes
    score = np.float32([0.1, 0, 0, 0, 0.1, 0, 0, 0.9])
    a, off = D.plan(mb, op, dn, sc, score=score, min_blending=0.02, min_opacity=0.1, split_scaling=0.1,
                    split_score=0.5)
    assert a.tolist() == [1, 0, 0, 0, 1, 1, 1, 2] and off[-1] == 8
    # no density array: 6 becomes a candidate as well
    a, _ = D.plan(mb, op, None, sc, min_blending=0.02, min_opacity=0.1, split_scaling=0.1)
    assert a.tolist() == [2, 0, 0, 0, 2, 1, 2, 2]


def test_plan_hand_worked_cap():
    mb = np.ones(6, np.float32)
    op = np.ones(6, np.float32)
    sc = np.float32([1, 0, 1, 1, 0, 1])  # candidates 0, 2, 3, 5; 6 kept
    for cap, splits in ((6, 0), (8, 0), (9, 1), (11, 1), (12, 2), (17, 3), (18, 4), (100, 4), (3, 0), (0, 4)):
        a, off = D.plan(mb, op, None, sc, split_scaling=0.5, max_out=cap)
        granted = [i for i in (0, 2, 3, 5)][:splits]
        assert [i for i in range(6) if a[i] == 2] == granted, cap
        assert off[-1] == 6 + 3 * splits


def test_plan_caller_actions():
    act = np.uint8([2, 0, 1, 2, 7, 2])
    a, off = D.plan(action_in=act, max_out=9)  # keeps 0, 2, 3, 5 (7 is not an action: dropped); one split fits
    assert a.tolist() == [2, 0, 1, 1, 0, 1] and off.tolist() == [0, 4, 4, 5, 6, 6, 7]


def _greedy(keep, cand, cap):
    budget = cap - int(keep.sum())
    split = np.zeros_like(cand)
    for i in range(len(cand)):
        if cand[i] and budget >= 3:
            split[i] = True
            budget -= 3
    return split


def test_plan_cap_matches_sequential_greedy():
    rng = np.random.default_rng(5)
    for trial in range(40):
        P = int(rng.integers(0, 300))
        mb = rng.random(P).astype(np.float32)
        op = rng.random(P).astype(np.float32)
        sc = rng.random(P).astype(np.float32)
        cap = int(rng.integers(0, 4 * P + 2))
        a, off = D.plan(mb, op, None, sc, min_blending=0.2, min_opacity=0.1, split_scaling=0.6, max_out=cap)
        keep = (mb >= 0.2) & (op >= 0.1)
        cand = keep & (sc > 0.6)
        split = _greedy(keep, cand, cap if cap > 0 else 2 ** 31 - 1)
        assert np.array_equal(a == 2, split), trial
        assert np.array_equal(a >= 1, keep)
        assert np.array_equal(np.diff(off), keep.astype(int) + 3 * split)


def _tri_scene(P, rng, shuffle=False):
    verts = rng.standard_normal((3 * P, 3)).astype(np.float32)
    cumsum = (3 * np.arange(P)).astype(np.int32)
    if shuffle:  # triangles stored in a permuted order of vertex triples
        perm = rng.permutation(P)
        cumsum = (3 * perm).astype(np.int32)
    return verts, cumsum


def _area_normal(t):
    n = np.cross(t[..., 1, :] - t[..., 0, :], t[..., 2, :] - t[..., 0, :])
    return 0.5 * np.linalg.norm(n, axis=-1), n


def test_subdivision_geometry():
    rng = np.random.default_rng(1)
    P = 200
    verts, cumsum = _tri_scene(P, rng)
    act = np.full(P, D.SPLIT, np.uint8)
    _, off = D.plan(action_in=act)
    (out,), _, _ = D.apply(off, cumsum, [(verts, 9, D.ROWS_VERTEX, D.SPLIT_SUBDIVIDE)])
    kids = out.reshape(P, 4, 3, 3).astype(np.float64)
    par = verts.reshape(P, 3, 3).astype(np.float64)
    pa, pn = _area_normal(par)
    ka, kn = _area_normal(kids)
    np.testing.assert_allclose(ka, np.broadcast_to(pa[:, None] / 4, ka.shape), rtol=1e-4)
    assert np.all(np.einsum("pkc,pc->pk", kn, pn) > 0)  # orientation kept
    np.testing.assert_allclose(kids.mean(axis=2).mean(axis=1), par.mean(axis=1), atol=1e-5)
    # the children tile the parent: interior sample points (barycentric weights >= 0.02 away
    # from the parent's and children's edges) lie in exactly one child
    for p in range(20):
        w = rng.dirichlet((1, 1, 1), 200)
        w = w[(np.abs(w - 0.5) > 0.02).all(1) & (w > 0.02).all(1)]
        x = w @ par[p]
        inside = np.zeros((len(x), 4), bool)
        for k in range(4):
            a, b, c = kids[p, k]
            m = np.stack([b - a, c - a], 1)  # 3x2
            uv = np.linalg.lstsq(m, (x - a).T, rcond=None)[0].T
            inside[:, k] = (uv >= -1e-9).all(1) & (uv.sum(1) <= 1 + 1e-9)
        assert np.all(inside.sum(1) == 1)


def test_apply_policies_identity_and_shuffled_cumsum():
    rng = np.random.default_rng(2)
    P = 64
    verts, cumsum = _tri_scene(P, rng, shuffle=True)
    sh = rng.standard_normal((P, 48)).astype(np.float32)
    mom = rng.standard_normal((3 * P, 3)).astype(np.float32)
    arrays = [(verts, 9, D.ROWS_VERTEX, D.SPLIT_SUBDIVIDE), (sh, 48, D.ROWS_TRIANGLE, D.SPLIT_COPY),
              (sh, 48, D.ROWS_TRIANGLE, D.SPLIT_ZERO), (mom, 9, D.ROWS_VERTEX, D.SPLIT_ZERO)]
    # everything kept: a gather of each triangle's rows into index order
    _, off = D.plan(action_in=np.ones(P, np.uint8))
    (v, s1, s2, m), npts, cs = D.apply(off, cumsum, arrays)
    gathered = verts.reshape(P, 3, 3)[cumsum // 3].reshape(-1)
    assert np.array_equal(v, gathered) and np.array_equal(s1, sh.reshape(-1)) and np.array_equal(s2, sh.reshape(-1))
    assert np.array_equal(m, mom.reshape(P, 3, 3)[cumsum // 3].reshape(-1))
    assert np.all(npts == 3) and np.array_equal(cs, 3 * np.arange(P))
    # mixed actions: drop / keep / split
    act = rng.integers(0, 3, P).astype(np.uint8)
    _, off = D.plan(action_in=act)
    (v, s1, s2, m), _, _ = D.apply(off, cumsum, arrays)
    src = np.nonzero(act > 0)[0]
    rows = np.repeat(src, np.where(act[src] == 2, 4, 1))
    splitrow = np.repeat(act[src] == 2, np.where(act[src] == 2, 4, 1))
    assert np.array_equal(s1.reshape(-1, 48), sh[rows])
    assert np.all(s2.reshape(-1, 48)[splitrow] == 0) and np.array_equal(s2.reshape(-1, 48)[~splitrow], sh[rows][~splitrow])
    assert np.all(m.reshape(-1, 9)[splitrow] == 0)
    kept_pts = verts.reshape(P, 9)[cumsum[rows] // 3]
    assert np.array_equal(v.reshape(-1, 9)[~splitrow], kept_pts[~splitrow])
    # a split parent's children hold its corners (once each) and its edge midpoints (three times each)
    i = int(np.nonzero(act == 2)[0][0])
    o = int(off[i])
    kids = v.reshape(-1, 3, 3)[o:o + 4].reshape(-1, 3)
    par = verts.reshape(P, 3, 3)[cumsum[i] // 3]
    for c in range(3):
        assert (kids == par[c]).all(1).sum() == 1
    for u, w in ((0, 1), (1, 2), (2, 0)):
        assert (kids == (par[u] + par[w]) * np.float32(0.5)).all(1).sum() == 3


def test_empty_inputs():
    a, off = D.plan(np.zeros(0, np.float32), np.zeros(0, np.float32))
    assert a.size == 0 and off.tolist() == [0]
    (v,), npts, cs = D.apply(off, np.zeros(0, np.int32), [(np.zeros(0, np.float32), 9, D.ROWS_VERTEX,
                                                           D.SPLIT_SUBDIVIDE)])
    assert v.size == 0 and npts.size == 0
