"""The merge oracle (oracle/lm_merge_oracle.c, a restatement of
mergeDuplicateOrientations, src/LaueMatchingHeaders.h:949-1065) pinned by
scipy's rotation algebra, known answers and brute force; then the library's
host merge (lm_merge_candidates) compared with it.

What pins what:
  * quaternion of a matrix: scipy, in all four branches of the conversion;
  * misorientation: scipy's one-sided minimum over the point group (cubic
    and a hexagonal group built here from its axes), known angles;
  * the float32 matrix and the `<=` test: pairs whose float32 misorientation
    lies below / above the float64 value, with MaxAngle set exactly at it;
  * the strict `>` best-score rule: tied and strictly larger scores;
  * greedy list-order clustering: well separated grains (brute force);
  * the 50000 cap: numpy's stable descending argsort (glibc's merge-sort
    qsort keeps ties in list order) followed by the uncapped merge.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_bind as ob
from lauematching_amd import api


def _wxyz(r):
    x, y, z, w = r.as_quat()
    q = np.array([w, x, y, z])
    return q if w >= 0 else -q


def cubic_ops():
    gens = [Rotation.from_euler("x", 90, degrees=True), Rotation.from_euler("y", 90, degrees=True)]
    ops = [Rotation.identity()]
    grown = True
    while grown:
        grown = False
        for a in list(ops):
            for g in gens:
                c = a * g
                if all((c * o.inv()).magnitude() > 1e-6 for o in ops):
                    ops.append(c)
                    grown = True
    assert len(ops) == 24
    return ops


def hexagonal_ops():
    """622: six turns about z by multiples of 60 degrees and six half turns
    about axes in the xy plane at multiples of 30 degrees."""
    ops = [Rotation.from_euler("z", 60 * k, degrees=True) for k in range(6)]
    for k in range(6):
        a = math.radians(30 * k)
        ops.append(Rotation.from_rotvec(np.pi * np.array([math.cos(a), math.sin(a), 0.0])))
    return ops


def test_quat_from_matrix_all_branches_vs_scipy():
    rots = list(Rotation.random(400, random_state=31))
    # trace <= 0 with each diagonal entry the largest: half turns about x, y, z
    # (and near them), which take the three non-trace branches
    for axis in np.eye(3):
        for d in (0.0, 0.05, -0.07):
            v = axis + d * np.array([0.3, -0.2, 0.1])
            rots.append(Rotation.from_rotvec(np.pi * 0.97 * v / np.linalg.norm(v)))
    branches = set()
    for r in rots:
        m = r.as_matrix()
        tr = np.trace(m)
        branches.add("trace" if tr > 0 else ("x" if m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]
                                             else ("y" if m[1, 1] > m[2, 2] else "z")))
        q = ob.quat_from_matrix(m.ravel())
        np.testing.assert_allclose(q, _wxyz(r), atol=1e-12)
        assert q[0] >= 0
    assert branches == {"trace", "x", "y", "z"}


@pytest.mark.parametrize("sg,ops", [(225, cubic_ops), (194, hexagonal_ops)])
def test_misorientation_one_sided_minimum_vs_scipy(sg, ops):
    """min over g of angle(a^-1 b g), the reduction applied on the right
    (src/LaueMatchingHeaders.h:222-260); the 5-decimal tables allow ~5e-3
    degrees."""
    G = ops()
    A = Rotation.random(150, random_state=sg)
    B = Rotation.random(150, random_state=sg + 1)
    for a, b in zip(A, B):
        want = math.degrees(min((a.inv() * b * g).magnitude() for g in G))
        got = ob.misorientation(_wxyz(a), _wxyz(b), sg)
        assert abs(got - want) < 5e-3, (got, want)


def test_misorientation_known_answers():
    ident = np.array([1.0, 0, 0, 0])
    for ang in (0.0, 0.7, 12.0, 40.0):
        q = _wxyz(Rotation.from_euler("z", ang, degrees=True))
        assert abs(ob.misorientation(ident, q, 225) - ang) < 2e-3
    for op in cubic_ops():
        assert ob.misorientation(ident, _wxyz(op), 225) < 2e-3
    # triclinic: no symmetry, the plain rotation angle
    r = Rotation.from_rotvec([0.3, -1.1, 0.4])
    assert abs(ob.misorientation(ident, _wxyz(r), 1) - math.degrees(r.magnitude())) < 1e-9


def _pair_with_float_rounding(direction, sg=225, seed=0):
    """Two orientations whose float64 misorientation m rounds to float32
    below m (direction < 0) or above it (> 0)."""
    rng = np.random.default_rng(seed)
    base = Rotation.random(1, random_state=seed)[0]
    while True:
        ax = rng.normal(size=3)
        d = Rotation.from_rotvec(ax / np.linalg.norm(ax) * math.radians(rng.uniform(0.5, 3.0)))
        other = base * d
        m = ob.misorientation(_wxyz(base), _wxyz(other), sg)
        f = float(np.float32(m))
        if (f < m) if direction < 0 else (f > m):
            return np.stack([base.as_matrix().ravel(), other.as_matrix().ravel()]), m, f


def test_float32_matrix_and_le_boundary():
    # float32(m) < m: MaxAngle = float32(m) merges (float test), a double
    # test m <= MaxAngle would not
    o, m, f = _pair_with_float_rounding(-1)
    rows, sc = np.array([0, 1], np.uint64), np.array([1.0, 2.0])
    assert list(ob.merge(o, rows, sc, f)[2]) == [2]
    assert list(ob.merge(o, rows, sc, np.nextafter(f, -np.inf))[2]) == [1, 1]
    # float32(m) > m: MaxAngle = m does not merge (a double test would)
    o, m, f = _pair_with_float_rounding(+1, seed=5)
    assert list(ob.merge(o, rows, sc, m)[2]) == [1, 1]
    assert list(ob.merge(o, rows, sc, f)[2]) == [2]


def test_best_score_strictly_greater():
    r = Rotation.random(1, random_state=2)[0]
    near = [r * Rotation.from_rotvec([0.002 * k, 0, 0]) for k in range(4)]
    o = np.stack([x.as_matrix().ravel() for x in near])
    rows = np.arange(4, dtype=np.uint64)
    b, s, c = ob.merge(o, rows, np.array([5.0, 5.0, 5.0, 5.0]), 2.0)
    assert (list(b), list(s), list(c)) == ([0], [5.0], [4])     # ties keep the head
    b, s, c = ob.merge(o, rows, np.array([5.0, 7.0, 7.0, 6.0]), 2.0)
    assert (list(b), list(s), list(c)) == ([1], [7.0], [4])     # first strictly larger
    b, s, c = ob.merge(o, rows[::-1].copy(), np.array([5.0, 7.0, 7.0, 6.0]), 2.0)
    assert list(b) == [2]


def _grains(n_grains, per, spread_deg, seed, sep_check=True):
    rng = np.random.default_rng(seed)
    centres = Rotation.random(n_grains, random_state=seed)
    ops = cubic_ops()
    mats, labels = [], []
    for c in range(n_grains):
        for _ in range(per if per else int(rng.integers(1, 6))):
            ax = rng.normal(size=3)
            d = Rotation.from_rotvec(ax / np.linalg.norm(ax) *
                                     math.radians(rng.uniform(0, spread_deg)))
            op = ops[int(rng.integers(0, 24))]
            mats.append((centres[c] * d * op).as_matrix().ravel())
            labels.append(c)
    mats, labels = np.array(mats), np.array(labels)
    if sep_check:  # grains far apart: every cross-grain pair well above 2 degrees
        q = [_wxyz(c) for c in centres]
        for i in range(n_grains):
            for j in range(i):
                assert ob.misorientation(q[i], q[j], 225) > 4 * spread_deg + 3.0
    return mats, labels


def test_greedy_clustering_vs_bruteforce():
    mats, labels = _grains(14, 0, 0.4, seed=8)
    rng = np.random.default_rng(9)
    perm = rng.permutation(len(mats))
    mats, labels = mats[perm], labels[perm]
    rows = np.arange(len(mats), dtype=np.uint64)
    scores = rng.uniform(100, 1000, len(rows))
    b, s, c = ob.merge(mats, rows, scores, 2.0)
    seen, exp = set(), []
    for i in range(len(rows)):
        if labels[i] in seen:
            continue
        seen.add(labels[i])
        mem = [j for j in range(len(rows)) if labels[j] == labels[i]]
        top = max(mem, key=lambda j: (scores[j], -j))
        exp.append((top, scores[top], len(mem)))
    assert [(int(x), y, int(z)) for x, y, z in zip(b, s, c)] == exp


def test_cap_keeps_50000_best_stable():
    """60000 candidates (40 grains) with integer scores, many tied at the
    cut: the merge equals the uncapped merge of numpy's stable descending
    selection."""
    mats, _ = _grains(40, 1500, 0.5, seed=17, sep_check=False)
    rng = np.random.default_rng(18)
    rows = rng.permutation(len(mats)).astype(np.uint64)
    scores = rng.integers(100, 200, len(rows)).astype(np.float64)
    assert len(rows) == 60000
    keep = np.argsort(-scores, kind="stable")[:50000]
    want = ob.merge(mats, rows[keep], scores[keep], 2.0)
    got = ob.merge(mats, rows, scores, 2.0)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    # the cut falls inside a run of tied scores, so the order of ties matters
    cut = scores[keep[-1]]
    assert np.count_nonzero(scores == cut) > np.count_nonzero(scores[keep] == cut)


def merge_cases():
    """Candidate sets with many pairs decided close to MaxAngle, tied scores,
    for three point groups (the sets of the device-merge tests)."""
    from test_gpu_merge import _candidates
    return [(sg, ma, _candidates(ng, pg, ma, seed=sg + ng))
            for sg, ma, ng, pg in [(225, 2.0, 150, 12), (194, 1.0, 100, 20), (139, 3.0, 60, 30)]]


@pytest.mark.parametrize("case", range(3))
def test_host_merge_equals_oracle(case):
    sg, ma, (orients, rows, scores) = merge_cases()[case]
    want = ob.merge(orients, rows, scores, ma, sg)
    got = api.merge_candidates(orients, rows, scores, ma, sg, threads=8)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert len(want[0]) < len(rows)


def test_host_merge_boundary_pairs_equal_oracle():
    rows, sc = np.array([0, 1], np.uint64), np.array([1.0, 2.0])
    for direction, seed in ((-1, 0), (+1, 5)):
        o, m, f = _pair_with_float_rounding(direction, seed=seed)
        for ma in (m, f, np.nextafter(f, -np.inf), np.nextafter(f, np.inf)):
            want = ob.merge(o, rows, sc, ma)
            got = api.merge_candidates(o, rows, sc, ma, 225, threads=1)
            for g, w in zip(got, want):
                np.testing.assert_array_equal(g, w)
