"""Merge of duplicate candidates with the pair tests on the device
(lm_merge_candidates_dev): identical clusters, best rows, scores and sizes to
the merge oracle (oracle/lm_merge_oracle.c, pinned in test_merge_oracle.py)
and to the host merge (lm_merge_candidates), on clustered candidate sets with
many pairs within 0.01 degree of MaxAngle, for cubic, hexagonal and
tetragonal point groups; pairs whose float32 misorientation equals MaxAngle;
and more than 50000 candidates with tied scores (the reference's cap)."""
import math
import sys

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_bind as ob
from lauematching_amd import _lib, api

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return api.Context(0)


def _candidates(n_grains, per_grain, max_angle, seed):
    """Grains with members scattered around them: half well inside
    max_angle, half at max_angle +- 0.01 degree from the grain centre (so the
    pass test is decided close to its boundary), plus unrelated rows; rows
    and scores in a shuffled list order with tied scores."""
    rng = np.random.default_rng(seed)
    centres = Rotation.random(n_grains, random_state=seed)
    mats = []
    for c in range(n_grains):
        mats.append(centres[c].as_matrix().ravel())
        for j in range(per_grain):
            axis = rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            ang = (rng.uniform(0, 0.8 * max_angle) if j % 2 else
                   max_angle + rng.uniform(-0.01, 0.01))
            d = Rotation.from_rotvec(axis * math.radians(ang))
            mats.append((centres[c] * d).as_matrix().ravel())
    mats += [r.as_matrix().ravel() for r in Rotation.random(n_grains * 2, random_state=seed + 1)]
    orients = np.array(mats)
    rows = rng.permutation(len(orients)).astype(np.uint64)
    scores = np.round(rng.uniform(100, 1000, len(rows)), 0)  # many ties
    return orients, rows, scores


@pytest.mark.parametrize("sg,max_angle,n_grains,per_grain", [(225, 2.0, 300, 12),
                                                             (194, 1.0, 200, 20),
                                                             (139, 3.0, 100, 30),
                                                             (225, 0.5, 60, 80)])
def test_device_merge_equals_host_merge(ctx, sg, max_angle, n_grains, per_grain):
    orients, rows, scores = _candidates(n_grains, per_grain, max_angle, seed=sg + n_grains)
    _say(f"merge sg={sg} max_angle={max_angle} n={len(rows)}")
    hb, hs, hc = api.merge_candidates(orients, rows, scores, max_angle, sg, threads=8)
    db, ds, dc = api.merge_candidates(orients, rows, scores, max_angle, sg, ctx=ctx)
    ob_, os_, oc_ = ob.merge(orients, rows, scores, max_angle, sg)
    np.testing.assert_array_equal(db, ob_)
    np.testing.assert_array_equal(ds, os_)
    np.testing.assert_array_equal(dc, oc_)
    assert len(hb) == len(db)
    np.testing.assert_array_equal(db, hb)
    np.testing.assert_array_equal(ds, hs)
    np.testing.assert_array_equal(dc, hc)
    assert hc.sum() == len(rows)
    assert len(hb) < len(rows)  # clusters formed


def test_device_merge_edge_cases(ctx):
    o = Rotation.random(5, random_state=1).as_matrix().reshape(-1, 9)
    b, s, c = api.merge_candidates(o, [], [], 2.0, 225, ctx=ctx)
    assert len(b) == 0
    b, s, c = api.merge_candidates(o, [3], [7.0], 2.0, 225, ctx=ctx)
    assert list(b) == [3] and list(s) == [7.0] and list(c) == [1]
    # negative angle: nothing merges; 180 degrees: everything merges
    rows = np.arange(5, dtype=np.uint64)
    sc = np.array([5.0, 9.0, 1.0, 9.0, 3.0])
    b, s, c = api.merge_candidates(o, rows, sc, -1.0, 225, ctx=ctx)
    assert list(c) == [1] * 5
    b, s, c = api.merge_candidates(o, rows, sc, 180.0, 225, ctx=ctx)
    assert list(b) == [1] and list(s) == [9.0] and list(c) == [5]
    with pytest.raises(_lib.LaueMatchError):
        api.merge_candidates(o, [7], [1.0], 2.0, 225, ctx=ctx)  # row outside the DB


def test_device_merge_boundary_pairs_equal_oracle(ctx):
    """MaxAngle exactly at a pair's float32 misorientation, one float64 step
    either side of it, and at the float64 value."""
    from test_merge_oracle import _pair_with_float_rounding

    rows, sc = np.array([0, 1], np.uint64), np.array([1.0, 2.0])
    for direction, seed in ((-1, 0), (+1, 5)):
        o, m, f = _pair_with_float_rounding(direction, seed=seed)
        for ma in (m, f, np.nextafter(f, -np.inf), np.nextafter(f, np.inf)):
            _say(f"boundary pair direction={direction} max_angle={ma!r}")
            want = ob.merge(o, rows, sc, ma)
            got = api.merge_candidates(o, rows, sc, ma, 225, ctx=ctx)
            for g, w in zip(got, want):
                np.testing.assert_array_equal(g, w)


def test_device_and_host_merge_at_the_cap_equal_oracle(ctx):
    """60000 candidates of 40 grains, integer scores tied across the cut at
    50000: the device merge, the host merge and the oracle agree."""
    from test_merge_oracle import _grains

    mats, _ = _grains(40, 1500, 0.5, seed=17, sep_check=False)
    rng = np.random.default_rng(18)
    rows = rng.permutation(len(mats)).astype(np.uint64)
    scores = rng.integers(100, 200, len(rows)).astype(np.float64)
    _say(f"merge at the cap: {len(rows)} candidates")
    want = ob.merge(mats, rows, scores, 2.0, 225, threads=16)
    got = api.merge_candidates(mats, rows, scores, 2.0, 225, ctx=ctx)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    host = api.merge_candidates(mats, rows, scores, 2.0, 225, threads=16)
    for g, w in zip(host, want):
        np.testing.assert_array_equal(g, w)
