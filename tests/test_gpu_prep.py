"""Image preprocessing on the device (lm_preprocess) against the CPU oracle
(oracle/lm_prep_oracle.py, pinned in test_prep_oracle.py): threshold value,
component counts and centres-derived sigma exact; the float64 blur within
1e-12 relative (the Gaussian weights come from libm's exp; numpy's float64
exp on AVX-512 hosts differs from it in the last bit for about 5% of the
weight arguments); the uint8 image equal to quantizing the device's own
blurred image on the host, and to the oracle's except at most one level at
rounding boundaries.  The uint8 image left in device memory scores exactly
like the host copy.  The median background (lm_median_background) equals
scipy.ndimage.median_filter, the reference's filter without diplib, pixel for
pixel, and on full detector frames the C running-median oracle (pinned
against scipy in test_prep_oracle.py)."""
import os
import sys

import numpy as np
import pytest

from lauematching_amd import _lib, api, synth

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import lm_prep_oracle as po  # noqa: E402
import oracle_bind as ob  # noqa: E402

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


def _frame(kind, H, W, seed):
    rng = np.random.default_rng(seed)
    if kind == "poisson":    # detector counts: few distinct values
        return rng.poisson(80.0, (H, W)).astype(np.float64)
    if kind == "wide":       # integers over most of the 16-bit range
        return rng.integers(0, 60000, (H, W)).astype(np.float64)
    if kind == "float":      # every value distinct: rank bits beyond 16
        return rng.normal(100.0, 25.0, (H, W))
    if kind == "signed":     # negatives and both zeros
        img = rng.integers(-3, 4, (H, W)).astype(np.float64)
        img[img == 0] = np.where(rng.random(np.count_nonzero(img == 0)) < 0.5, 0.0, -0.0)
        return img
    raise ValueError(kind)


MEDIAN_CASES = [
    # kind, H, W, filter size, passes
    ("poisson", 200, 160, 31, 1),
    ("wide", 256, 256, 25, 1),
    ("float", 300, 400, 21, 1),     # 120000 distinct values: low rank bits resolved
    ("float", 2048, 2048, 5, 1),    # 2^22 distinct values
    ("float", 4096, 4096, 3, 1),    # 2^24 pixels, the largest frame
    ("poisson", 40, 24, 101, 2),    # window wider than the image
    ("wide", 64, 48, 255, 1),       # the largest window, many reflections
    ("poisson", 97, 130, 10, 3),    # even size, three passes
    ("signed", 64, 80, 7, 1),
    ("wide", 50, 70, 1, 1),         # identity
    ("poisson", 50, 70, 0, 1),      # zeros (preprocess.py:53-54)
    ("poisson", 50, 70, 9, 0),
]


@pytest.mark.parametrize("case", MEDIAN_CASES,
                         ids=[f"{c[0]}-{c[1]}x{c[2]}-s{c[3]}-p{c[4]}" for c in MEDIAN_CASES])
def test_median_background_matches_oracle(ctx, case):
    kind, H, W, size, passes = case
    img = _frame(kind, H, W, seed=H * W + size)
    _say(f"median background {kind} {H}x{W} size={size} passes={passes}")
    bg, ms = api.median_background(ctx, img, size, passes)
    want = po.median_background(img, size, passes)
    np.testing.assert_array_equal(bg, want)


def test_median_background_limits(ctx):
    """Windows over 255 pixels and frames over 2^24 pixels are refused with an
    error (their counts or ranks would not fit), before any device work."""
    img = _frame("poisson", 32, 32, seed=1)
    with pytest.raises(_lib.LaueMatchError):
        api.median_background(ctx, img, 256, 1)
    with pytest.raises(_lib.LaueMatchError):
        api.median_background(ctx, np.zeros((1, (1 << 24) + 1)), 3, 1)
    # the pipeline refuses the same
    with pytest.raises(_lib.LaueMatchError):
        api.preprocess(ctx, img, None, "fixed", 10.0, filter_radius=300)


def test_median_background_full_frame_sampled(ctx):
    """A 2048 x 2048 detector frame with the reference's default filter (101,
    one pass): every pixel is an element of the frame, and a sample of pixels
    (all four edges and corners among them) equals the sort of its window."""
    H = W = 2048
    size = 101
    img, _ = _raw(H, W, 3000, seed=5)
    img = np.floor(img)
    _say(f"median background {H}x{W} size={size}")
    bg, ms = api.median_background(ctx, img, size, 1)
    _say(f"median background {H}x{W} size={size}: {ms:.2f} ms")
    assert np.isin(bg, img).all()
    rng = np.random.default_rng(6)
    ys = np.concatenate([rng.integers(0, H, 300), [0, 0, H - 1, H - 1], rng.integers(0, 60, 40),
                         rng.integers(H - 60, H, 40), rng.integers(0, H, 80)])
    xs = np.concatenate([rng.integers(0, W, 300), [0, W - 1, 0, W - 1], rng.integers(0, W, 40),
                         rng.integers(0, W, 40), rng.integers(0, 60, 40),
                         rng.integers(W - 60, W, 40)])
    lo = size // 2
    pad = np.pad(img, ((lo, size - 1 - lo), (lo, size - 1 - lo)), mode="symmetric")
    k = size * size // 2
    for y, x in zip(ys, xs):
        win = pad[y:y + size, x:x + size].ravel()
        assert bg[y, x] == np.partition(win, k)[k], (y, x)


@pytest.mark.parametrize("kind,n,size,passes", [("counts", 2048, 101, 1), ("float", 2048, 101, 1),
                                                ("float", 4096, 31, 2)])
def test_median_background_full_frame_equals_oracle(ctx, kind, n, size, passes):
    """Detector frames at the reference's default filter (101, one pass) and
    the largest frame (2^24 pixels) twice at 31: every pixel equal to the C
    running-median oracle's.  Integer counts (a few thousand distinct values:
    one wavefront per column) and float64 (every value distinct: 22- and
    24-bit ranks, the coarse walk with its rank band)."""
    img, _ = _raw(n, n, 3000 * (n // 2048) ** 2, seed=7 + n + size)
    if kind == "counts":
        img = np.floor(img)
    else:
        img += np.random.default_rng(n).random(img.shape)
    _say(f"median background {kind} {n}x{n} size={size} passes={passes}")
    bg, ms = api.median_background(ctx, img, size, passes)
    _say(f"median background {kind} {n}x{n}: {ms:.2f} ms; oracle")
    want = ob.median_background(img, size, passes)
    np.testing.assert_array_equal(bg, want)


def test_preprocess_computes_the_median_background(ctx):
    """No background given: the pipeline subtracts the median background of
    the raw frame (preprocess.py:255-266) and then matches the oracle."""
    raw, _ = _raw(256, 320, 60, seed=21)
    _say("preprocess with a median background 256x320")
    u8, scale, info, blurred = api.preprocess(ctx, raw, None, "adaptive", want_blurred=True,
                                              filter_radius=31, median_passes=2)
    want, winfo = po.preprocess(raw, None, "adaptive", filter_radius=31, median_passes=2)
    assert info["threshold"] == winfo["threshold"]
    assert info["n_components"] == winfo["n_components"]
    assert info["n_kept"] == winfo["n_kept"]
    np.testing.assert_allclose(blurred, want, rtol=1e-12, atol=1e-12 * max(1.0, want.max()))


def test_working_buffers_count_against_the_budget():
    """The buffers a context keeps for preprocessing and the median background
    count against its HBM budget: a frame whose buffers do not fit fails with
    LM_ERR_BUDGET, a smaller one runs and leaves them in use for the next."""
    small = api.Context(0, hbm_budget=64 << 20)
    try:
        big, _ = _raw(2048, 2048, 100, seed=3)
        _say("preprocess over a 64 MiB budget")
        with pytest.raises(_lib.LaueMatchError) as err:
            api.preprocess(small, big, None, "fixed", 100.0)
        assert err.value.code == -4  # LM_ERR_BUDGET
        with pytest.raises(_lib.LaueMatchError) as err:
            api.median_background(small, big, 31, 1)
        assert err.value.code == -4
        raw, bg = _raw(256, 256, 20, seed=4)
        u8, scale, info, blurred = api.preprocess(small, raw, bg, "adaptive", want_blurred=True)
        want, winfo = po.preprocess(raw, bg, "adaptive")
        assert info["threshold"] == winfo["threshold"]
        assert small.bytes_in_use() > 0
        np.testing.assert_array_equal(api.median_background(small, raw, 31, 1)[0],
                                      po.median_background(raw, 31, 1))
    finally:
        small.close()
