"""Image preprocessing on the device (lm_preprocess) against the CPU oracle
(oracle/lm_prep_oracle.py, pinned in test_prep_oracle.py): threshold value,
component counts and centres-derived sigma exact; the float64 blur within
1e-12 relative (the Gaussian weights come from the host's exp, numpy's may
differ in the last bit); the uint8 image equal to quantizing the device's own
blurred image on the host, and to the oracle's except at most one level at
rounding boundaries.  The uint8 image left in device memory scores exactly
like the host copy."""
import os
import sys

import numpy as np
import pytest

from lauematching_amd import api, synth

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import lm_prep_oracle as po  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return api.Context(0)


def _raw(H, W, nspots, seed):
    """Poisson background plus Gaussian spots of random size and height, and
    a noisy background estimate; fractional values, so the uint16 cast
    truncates."""
    rng = np.random.default_rng(seed)
    img = rng.poisson(50.0, (H, W)).astype(np.float64)
    yy, xx = np.mgrid[0:H, 0:W]
    for _ in range(nspots):
        yc, xc = rng.uniform(0, H), rng.uniform(0, W)
        a, s = rng.uniform(300, 4000), rng.uniform(0.7, 2.5)
        y0, y1 = max(0, int(yc) - 10), min(H, int(yc) + 11)
        x0, x1 = max(0, int(xc) - 10), min(W, int(xc) + 11)
        img[y0:y1, x0:x1] += a * np.exp(-((yy[y0:y1, x0:x1] - yc) ** 2 +
                                          (xx[y0:y1, x0:x1] - xc) ** 2) / (2 * s * s))
    bg = 45.0 + rng.normal(0, 2.0, (H, W))
    return img, bg


CASES = [
    # H, W, spots, method, value, pct, min_area, px, distance, spacing, background
    (384, 512, 150, "adaptive", 0.0, 0.0, 10, 0.2, 513.0, 0.4, True),
    (384, 512, 150, "percentile", 0.0, 99.5, 10, 0.2, 513.0, 0.4, True),
    (384, 512, 150, "fixed", 120.0, 0.0, 5, 0.2, 1e6, 0.4, False),
    (300, 301, 80, "otsu", 0.0, 0.0, 3, 0.0, 0.0, 0.4, True),
    (2048, 2048, 3000, "adaptive", 0.0, 0.0, 10, 0.2, 513.0, 0.4, True),
    (200, 160, 0, "fixed", 1e9, 0.0, 10, 0.2, 513.0, 0.4, False),      # nothing passes
    (200, 160, 30, "fixed", 60.0, 0.0, 10000, 0.2, 513.0, 0.4, False),  # nothing kept
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[3]}-{c[0]}x{c[1]}-{i}" for i, c in
                                             enumerate(CASES)])
def test_preprocess_matches_oracle(ctx, case):
    H, W, n, method, val, pct, min_area, px, dist, spacing, use_bg = case
    raw, bg = _raw(H, W, n, seed=H + n)
    bg = bg if use_bg else None
    _say(f"preprocess {method} {H}x{W} spots={n}")
    u8, scale, info, blurred = api.preprocess(ctx, raw, bg, method, val, pct, min_area, px, dist,
                                              spacing, want_blurred=True)
    want, winfo = po.preprocess(raw, bg, method, val, pct, min_area, px, dist, spacing)
    assert info["threshold"] == winfo["threshold"]
    assert info["n_components"] == winfo["n_components"]
    assert info["n_kept"] == winfo["n_kept"]
    assert abs(info["sigma"] - winfo["sigma"]) <= 1e-12 * max(1.0, winfo["sigma"])
    np.testing.assert_allclose(blurred, want, rtol=1e-12, atol=1e-12 * max(1.0, want.max()))
    hu8, hscale = api.quantize_u8(blurred)
    np.testing.assert_array_equal(u8.ravel(), hu8)
    assert scale == hscale
    ou8, oscale = api.quantize_u8(want)
    diff = np.abs(u8.ravel().astype(int) - ou8.astype(int))
    assert diff.max() <= 1 and np.count_nonzero(diff) <= max(2, u8.size // 100000)
    if winfo["n_kept"] == 0:
        assert not u8.any() and scale == 1.0


def test_preprocessed_image_scores_from_device(ctx):
    import torch

    p = synth.baseline_params(MaxNrLaueSpots=64, NrPxX=512, NrPxY=384, PxX=0.0008, PxY=0.0008)
    hk = synth.fcc_hkls(6, "box")
    O = synth.random_orientations(20000, seed=17)
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), 64)
    db.forward()
    raw, bg = _raw(384, 512, 120, seed=99)
    dev = torch.zeros(384 * 512, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    u8, scale, info = api.preprocess(ctx, raw, bg, d_u8=dev.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), u8.ravel())
    a = db.score_dev(dev.data_ptr(), scale, 3, 200.0 * scale)
    b = db.score(u8.ravel(), scale, 3, 200.0 * scale)
    assert a.nr_results == b.nr_results
    np.testing.assert_array_equal(a.ids, b.ids)
    np.testing.assert_array_equal(a.score, b.score)
    db.close()


@pytest.mark.parametrize("method", ["adaptive", "percentile", "fixed"])
def test_preprocess_threshold_matches_reference_goldens(ctx, method):
    """The device threshold and the count of pixels it keeps, on the seeded
    frames of the reference's threshold characterization, equal the
    reference's goldens (and the oracle's threshold bit for bit)."""
    from test_prep_goldens import FRAMES, METHODS, frame, golden

    want = golden(method)
    kw = METHODS[method]
    for name, fkw in FRAMES.items():
        img = frame(**fkw)
        _say(f"preprocess golden frame {name} {method}")
        u8, scale, info = api.preprocess(ctx, img, None, kw["method"], kw["fixed_value"],
                                         kw["percentile"], 10, 0.2, 513.0, 0.4)
        t = po.threshold_value(img, kw["method"], kw["fixed_value"], kw["percentile"])
        assert info["threshold"] == t, (name, info["threshold"], t)
        assert round(info["threshold"], 6) == want[name]["thresh"], name
        assert info["n_above"] == want[name]["nonzero"], name
