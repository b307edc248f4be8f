"""The preprocessing oracle (oracle/lm_prep_oracle.py) against the
reference's own threshold goldens (tests/golden/ref_threshold_*.json, copied
from the reference's tests/golden): for four seeded frames -- bright spots,
faint spots, faint spots with saturated hot pixels, an empty frame -- the
threshold value (to the 6 decimals the goldens keep) and the number of pixels
above it, for the adaptive noise floor, the 99.5th percentile and a fixed
240.  The frames follow the recipe of the reference's
tests/test_char_threshold.py:15-37 (numpy's legacy RandomState stream, so
the same seeds give the same pixels)."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import lm_prep_oracle as po  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")

FRAMES = {
    "bright": dict(spot_amp=400.0, seed=1),
    "faint": dict(spot_amp=80.0, seed=2),
    "faint_with_hot": dict(spot_amp=150.0, seed=3, n_hot=33),
    "empty": dict(spot_amp=0.0, n_spots=0, noise=0.0, seed=4),
}

METHODS = {
    "adaptive": dict(method="adaptive", fixed_value=0.0, percentile=90.0),
    "percentile": dict(method="percentile", fixed_value=0.0, percentile=99.5),
    "fixed": dict(method="fixed", fixed_value=240.0, percentile=90.0),
}


def frame(spot_amp, n_spots=40, noise=8.0, size=256, seed=0, n_hot=0, hot_val=65535.0,
          hot_seed=99):
    """|N(0, noise)| background, n_spots 3x3 squares of height spot_amp at
    seeded positions, optional saturated hot pixels from a second stream."""
    rng = np.random.RandomState(seed)
    img = np.abs(rng.normal(0, noise, (size, size)))
    for _ in range(n_spots):
        y, x = rng.randint(8, size - 8, 2)
        img[y - 1:y + 2, x - 1:x + 2] += spot_amp
    if n_hot:
        hot = np.random.RandomState(hot_seed)
        for _ in range(n_hot):
            y, x = hot.randint(0, size, 2)
            img[y, x] = hot_val
    return img


def golden(method):
    with open(os.path.join(GOLDEN, f"ref_threshold_{method}.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("method", list(METHODS))
def test_oracle_threshold_matches_reference_goldens(method):
    want = golden(method)
    kw = METHODS[method]
    for name, fkw in FRAMES.items():
        img = frame(**fkw)
        t = po.threshold_value(img, kw["method"], kw["fixed_value"], kw["percentile"])
        out = img.copy()
        out[img <= t] = 0
        assert round(float(t), 6) == want[name]["thresh"], (method, name)
        assert int((out > 0).sum()) == want[name]["nonzero"], (method, name)
