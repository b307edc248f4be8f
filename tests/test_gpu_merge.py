"""Merge of duplicate candidates with the pair tests on the device
(lm_merge_candidates_dev): identical clusters, best rows, scores and sizes to
the host merge (lm_merge_candidates, pinned in test_merge_cpu.py), on
clustered candidate sets with many pairs within 0.01 degree of MaxAngle, for
cubic, hexagonal and tetragonal point groups."""
import math
import sys

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

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
