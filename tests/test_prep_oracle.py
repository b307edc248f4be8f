"""Pins of the preprocessing oracle (oracle/lm_prep_oracle.py) against
brute force and closed forms, so the device pipeline is checked against a
restatement that is itself checked: 8-connectivity against a flood fill,
centres of mass against explicit sums, the nearest-centre distance against
all pairs, the blur's edge mode against an explicit padded correlation, the
noise floor against a hand-computed case, the median background against a
brute-force sort of every window and the properties an order statistic has,
and the C running-median oracle against scipy's median_filter."""
import os
import sys

import numpy as np
import scipy.ndimage as ndi

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import lm_prep_oracle as po  # noqa: E402
import oracle_bind as ob  # noqa: E402


def _flood(mask):
    """Brute-force 8-connected components: sets of pixels."""
    H, W = mask.shape
    seen = np.zeros_like(mask, bool)
    comps = []
    for y in range(H):
        for x in range(W):
            if mask[y, x] and not seen[y, x]:
                stack, comp = [(y, x)], set()
                seen[y, x] = True
                while stack:
                    cy, cx = stack.pop()
                    comp.add((cy, cx))
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            ny, nx = cy + dy, cx + dx
                            if 0 <= ny < H and 0 <= nx < W and mask[ny, nx] and not seen[ny, nx]:
                                seen[ny, nx] = True
                                stack.append((ny, nx))
                comps.append(frozenset(comp))
    return set(comps)


def test_components_are_8_connected():
    rng = np.random.default_rng(4)
    for trial in range(20):
        u16 = (rng.random((12, 15)) < 0.3).astype(np.uint16) * rng.integers(1, 9, (12, 15))
        labels, n = po.components(u16.astype(np.uint16))
        got = {frozenset(zip(*np.nonzero(labels == k))) for k in range(1, n + 1)}
        assert got == _flood(u16 > 0)
    # diagonal neighbours join, as with 8-connectivity
    u = np.zeros((4, 4), np.uint16)
    u[0, 0] = u[1, 1] = u[2, 2] = 5
    assert po.components(u)[1] == 1


def test_centres_and_area_filter():
    u = np.zeros((10, 12), np.uint16)
    u[2, 3], u[2, 4], u[3, 4] = 2, 6, 4          # area 3
    u[7, 8:11] = [1, 1, 1]                      # area 3
    u[0, 11] = 9                                # area 1, removed
    labels, n = po.components(u)
    filt, centres = po.filter_small(u, labels, n, 2)
    assert filt[0, 11] == 0 and filt[2, 4] == 6
    want = sorted([((3 * 2 + 4 * 6 + 4 * 4) / 12, (2 * 2 + 2 * 6 + 3 * 4) / 12), (9.0, 7.0)])
    assert sorted(centres) == want


def test_sigma_from_nearest_centres():
    rng = np.random.default_rng(5)
    pts = [tuple(p) for p in rng.uniform(0, 500, (40, 2))]
    brute = min(np.hypot(a[0] - b[0], a[1] - b[1]) for i, a in enumerate(pts) for b in pts[i + 1:])
    # delta large: sigma = brute / 4 (>= 1)
    assert po.gaussian_sigma(pts, 0.2, 1e6, 0.4) == max(0.25 * brute, 1.0)
    # delta small: sigma = 0.25 * delta, delta = distance tan(spacing) / px
    d = 513.0 * np.tan(np.radians(0.01)) / 0.2
    assert abs(po.gaussian_sigma(pts, 0.2, 513.0, 0.01) - max(0.25 * min(brute, d), 1.0)) < 1e-12
    assert po.gaussian_sigma(pts[:1], 0.2, 513.0, 0.4) == 3.0


def test_blur_reflects_at_the_edges():
    rng = np.random.default_rng(6)
    img = rng.random((9, 13))
    sigma = 1.3
    r = int(4.0 * sigma + 0.5)
    x = np.arange(-r, r + 1)
    w = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    w /= w.sum()
    pad = np.pad(img, r, mode="symmetric")   # d c b a | a b c d | d c b a
    ref = np.zeros_like(img)
    for dy in range(2 * r + 1):
        for dx in range(2 * r + 1):
            ref += w[dy] * w[dx] * pad[dy:dy + img.shape[0], dx:dx + img.shape[1]]
    np.testing.assert_allclose(ndi.gaussian_filter(img, sigma), ref, rtol=1e-12, atol=1e-14)


def test_noise_floor_known_answer():
    img = np.zeros((3, 4))
    img.ravel()[:5] = [1.0, 2.0, 3.0, 10.0, 4.0]
    # nonzero median 3, |v - 3| = 2 1 0 7 1 -> median 1; 3 + 4 * 1.4826 * 1
    assert po.noise_floor(img) == 3.0 + 4.0 * (1.4826 * 1.0)
    assert po.noise_floor(np.zeros((2, 2))) == 1.0
    assert po.threshold_value(img, "otsu", 0, 0) == po.noise_floor(img)


def test_empty_and_all_small():
    raw = np.zeros((20, 30))
    out, info = po.preprocess(raw, method="fixed", fixed_value=5.0)
    assert not out.any() and info["n_components"] == 0
    raw[3, 3] = 100.0
    out, info = po.preprocess(raw, method="fixed", fixed_value=5.0, min_area=2)
    assert not out.any() and info["n_components"] == 1 and info["n_kept"] == 0


# (H, W, size): odd and even windows, windows wider than the image (repeated
# reflection), a single pixel
MEDIAN_SHAPES = [(7, 5, 3), (9, 6, 4), (5, 4, 11), (3, 2, 12), (1, 1, 5), (13, 17, 6),
                 (20, 11, 25), (16, 16, 1)]


def test_median_background_equals_window_sort():
    rng = np.random.default_rng(3)
    for H, W, size in MEDIAN_SHAPES:
        for img in (rng.integers(0, 9, (H, W)).astype(np.float64),  # many ties
                    rng.normal(0.0, 1.0, (H, W))):                   # negatives, no ties
            want = po.median_brute(img, size)
            np.testing.assert_array_equal(po.median_background(img, size, 1), want)
            twice = po.median_brute(want, size)
            np.testing.assert_array_equal(po.median_background(img, size, 2), twice)


def test_median_background_order_statistic_properties():
    rng = np.random.default_rng(4)
    img = rng.normal(100.0, 30.0, (24, 31))
    bg = po.median_background(img, 7, 1)
    # an element of its window, never an average of two
    assert np.isin(bg, img).all()
    # commutes with strictly increasing maps
    np.testing.assert_array_equal(po.median_background(np.exp(img / 50.0), 7, 1),
                                  np.exp(bg / 50.0))
    np.testing.assert_array_equal(po.median_background(-img, 7, 1), -bg)
    # constant frames, size 1, and the zero cases of preprocess.py:53-54
    np.testing.assert_array_equal(po.median_background(np.full((9, 8), 2.5), 5, 3), 2.5)
    np.testing.assert_array_equal(po.median_background(img, 1, 4), img)
    assert not po.median_background(img, 0, 1).any()
    assert not po.median_background(img, 5, 0).any()
    # a step edge survives, an isolated spike does not
    step = np.zeros((15, 15))
    step[:, 8:] = 10.0
    step[3, 3] = 1e6
    np.testing.assert_array_equal(po.median_background(step, 3, 1), np.where(
        np.arange(15)[None, :] >= 8, 10.0, 0.0) * np.ones((15, 1)))


def _median_frames(rng, H, W):
    """Frames with many ties, Poisson counts, distinct negatives, and signed
    integers holding both zeros."""
    signed = np.round(rng.normal(0.0, 1.5, (H, W)))
    z = signed == 0
    signed[z] = np.where(rng.random(int(z.sum())) < 0.5, 0.0, -0.0)
    return (rng.integers(0, 9, (H, W)).astype(np.float64),
            rng.poisson(50.0, (H, W)).astype(np.float64),
            rng.normal(0.0, 1.0, (H, W)), signed)


def test_median_c_oracle_equals_scipy():
    """lmo_median_ranks (the running-histogram oracle the device median is
    checked against on full frames) equals scipy.ndimage.median_filter, the
    reference's own call, pass by pass: the window-sort shapes, windows wider
    than the frame up to 255 pixels, even sizes, several passes, and a frame
    of 75000 distinct values (the coarse counts move across many bins).
    scipy mirrors with period 2n only while size < 8n on each axis (from
    size 8n on, its line buffers are read past the extension: on a 2 x 50
    frame, size 16 already differs); windows that wide are checked against
    the sort of every periodically mirrored window instead."""
    rng = np.random.default_rng(21)
    cases = [(H, W, s, 1) for H, W, s in MEDIAN_SHAPES] + [
        (60, 45, 31, 1), (33, 70, 101, 2), (40, 32, 255, 1), (64, 96, 10, 3), (250, 300, 21, 1)]
    for H, W, size, passes in cases:
        assert size < 8 * min(H, W)
        for img in _median_frames(rng, H, W):
            want = po.median_background(img, size, passes)
            np.testing.assert_array_equal(ob.median_background(img, size, passes), want,
                                          err_msg=f"{H}x{W} size {size} passes {passes}")
    for H, W, size in ((2, 2, 17), (3, 5, 31), (2, 50, 16), (30, 7, 101)):
        for img in _median_frames(rng, H, W):
            np.testing.assert_array_equal(ob.median_background(img, size, 1),
                                          po.median_brute(img, size),
                                          err_msg=f"{H}x{W} size {size}")
    assert not ob.median_background(np.ones((5, 5)), 0, 1).any()
    assert not ob.median_background(np.ones((5, 5)), 3, 0).any()


def test_median_c_oracle_thread_count_invariant():
    """Rows are independent: one thread and eight give the same frame."""
    rng = np.random.default_rng(22)
    img = rng.normal(0.0, 1.0, (97, 131))
    vals, inv = np.unique(img, return_inverse=True)
    rk = np.ascontiguousarray(inv.reshape(img.shape), dtype=np.uint32)
    outs = []
    for threads in ("1", "8"):
        code = ("import sys, numpy as np; sys.path.insert(0, %r); import oracle_bind as ob; "
                "rk = np.load(sys.argv[1]); out = np.empty_like(rk); "
                "assert ob.lib().lmo_median_ranks(ob._p(rk), rk.shape[1], rk.shape[0], 23, "
                "%d, ob._p(out)) == 0; np.save(sys.argv[2], out)"
                % (os.path.dirname(os.path.abspath(__file__)), vals.size))
        import subprocess
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            np.save(os.path.join(d, "rk.npy"), rk)
            subprocess.run([sys.executable, "-c", code, os.path.join(d, "rk.npy"),
                            os.path.join(d, "out.npy")], check=True,
                           env={**os.environ, "OMP_NUM_THREADS": threads})
            outs.append(np.load(os.path.join(d, "out.npy")))
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(vals[outs[0]], po.median_background(img, 23, 1))
