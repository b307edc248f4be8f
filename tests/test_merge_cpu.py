"""Merge of duplicate candidates (lm_merge_candidates, CPU code in the
library): quaternions and misorientations against scipy's rotation algebra,
clustering against a brute-force greedy pass on well separated clusters."""
import ctypes
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from lauematching_amd import _lib, synth

L = _lib.lib()
L.lm_quat_from_matrix.restype = None
L.lm_quat_from_matrix.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
L.lm_misorientation_deg.restype = ctypes.c_double
L.lm_misorientation_deg.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
L.lm_merge_candidates.restype = ctypes.c_int
L.lm_merge_candidates.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                  ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.POINTER(ctypes.c_int32)]


def quat(m):
    q = np.zeros(4)
    L.lm_quat_from_matrix(np.ascontiguousarray(m, dtype=np.float64).ctypes.data, q.ctypes.data)
    return q


def miso(q1, q2, sg=225):
    return L.lm_misorientation_deg(np.ascontiguousarray(q1).ctypes.data,
                                   np.ascontiguousarray(q2).ctypes.data, sg)


def cubic_ops():
    """The 24 proper rotations of the cube, generated from 90-degree turns."""
    gens = [Rotation.from_euler("x", 90, degrees=True), Rotation.from_euler("y", 90, degrees=True)]
    ops = [Rotation.identity()]
    changed = True
    while changed:
        changed = False
        for a in list(ops):
            for g in gens:
                c = a * g
                if all((c * o.inv()).magnitude() > 1e-6 for o in ops):
                    ops.append(c)
                    changed = True
    assert len(ops) == 24
    return ops


def test_quat_from_matrix_matches_scipy():
    R = Rotation.random(500, random_state=3)
    for r in R:
        q = quat(r.as_matrix())
        x, y, z, w = r.as_quat()
        ref = np.array([w, x, y, z]) * (1 if w >= 0 else -1)
        np.testing.assert_allclose(q, ref, atol=1e-12)
        assert q[0] >= 0


def test_misorientation_known_answers():
    ident = quat(np.eye(3))
    for ang in (0.0, 1.0, 10.0, 30.0, 44.0):
        q = quat(Rotation.from_euler("z", ang, degrees=True).as_matrix())
        assert abs(miso(ident, q) - ang) < 2e-3
    # a cube symmetry operation is no misorientation
    for op in cubic_ops():
        assert miso(ident, quat(op.as_matrix())) < 2e-3


def test_misorientation_random_vs_scipy():
    """The reference reduces a^-1 b by symmetry applied on the right only
    (src/LaueMatchingHeaders.h:224-260): min over g of angle(a^-1 b g).  That
    can exceed the two-sided crystal disorientation, which is checked as a
    lower bound."""
    ops = cubic_ops()
    R1 = Rotation.random(200, random_state=5)
    R2 = Rotation.random(200, random_state=6)
    for a, b in zip(R1, R2):
        one_sided = min((a.inv() * b * g).magnitude() for g in ops)
        two_sided = min((g2 * a.inv() * b * g).magnitude() for g in ops for g2 in ops)
        got = miso(quat(a.as_matrix()), quat(b.as_matrix()))
        assert abs(got - math.degrees(one_sided)) < 5e-3
        assert got >= math.degrees(two_sided) - 5e-3


def _merge(orients, rows, scores, max_angle, sg=225):
    nr = len(rows)
    best = np.zeros(max(nr, 1), np.uint64)
    bs = np.zeros(max(nr, 1))
    cs = np.zeros(max(nr, 1), np.int32)
    k = ctypes.c_int32()
    o = np.ascontiguousarray(orients, dtype=np.float64)
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    s = np.ascontiguousarray(scores, dtype=np.float64)
    _lib.check(L.lm_merge_candidates(o.ctypes.data, o.shape[0], r.ctypes.data, s.ctypes.data, nr,
                                     max_angle, sg, 4, best.ctypes.data, bs.ctypes.data,
                                     cs.ctypes.data, ctypes.byref(k)))
    return best[:k.value], bs[:k.value], cs[:k.value]


def test_merge_clusters_like_greedy_bruteforce():
    rng = np.random.default_rng(8)
    centers = Rotation.random(12, random_state=9)
    mats, labels = [], []
    for c in range(12):
        for _ in range(rng.integers(1, 6)):
            d = Rotation.from_rotvec(rng.normal(size=3) * math.radians(0.3))
            # add a random cube symmetry so members differ by a symmetry op too
            op = cubic_ops()[rng.integers(0, 24)]
            mats.append((d * centers[c] * op).as_matrix().ravel())
            labels.append(c)
    perm = rng.permutation(len(mats))
    orients = np.array(mats)[perm]
    labels = np.array(labels)[perm]
    # candidates: every DB row, in DB order, with random scores
    rows = np.arange(len(orients), dtype=np.uint64)
    scores = rng.uniform(100, 1000, len(rows))
    best, bs, cs = _merge(orients, rows, scores, 2.0)
    # greedy reference: the first unassigned candidate opens a cluster that
    # takes every later unassigned member within 2 degrees (here: same label)
    seen, exp = set(), []
    for i in range(len(rows)):
        if labels[i] in seen:
            continue
        seen.add(labels[i])
        mem = [j for j in range(len(rows)) if labels[j] == labels[i]]
        top = max(mem, key=lambda j: (scores[j], -j))
        exp.append((top, scores[top], len(mem)))
    assert len(best) == len(exp)
    for (b, s, c), (eb, es, ec) in zip(zip(best, bs, cs), exp):
        assert (int(b), s, int(c)) == (eb, es, ec)


def test_merge_empty_and_bad_row():
    best, bs, cs = _merge(np.eye(3).reshape(1, 9), [], [], 2.0)
    assert len(best) == 0
    with pytest.raises(_lib.LaueMatchError):
        _merge(np.eye(3).reshape(1, 9), [5], [1.0], 2.0)
