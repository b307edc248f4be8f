"""CPU oracle for the image preprocessing before the compare.

TEST INFRASTRUCTURE ONLY: the checker for the library's device pipeline
(lm_preprocess), never the product; only tests/ load it.  It restates
preprocess_image (laue_index/preprocess.py:221-337) with the reference's own
numerical libraries (numpy, scipy.ndimage, scipy.spatial), function by
function.  OpenCV is not installed here: connectedComponentsWithStats with
8-connectivity (preprocess.py:136-155) is restated with scipy.ndimage.label and
a 3x3 structure, which yields the same components (labels may be numbered
differently; nothing downstream depends on the numbering).  scikit-image is not
installed either, so "otsu" falls back to the adaptive noise floor exactly as
the reference does without it (thresholds.py:92-110).  diplib is not
installed either, so the background median filter is the reference's scipy
branch (compute_background, preprocess.py:43-64), pinned on small frames by a
brute-force window sort (median_brute).
"""
from __future__ import annotations

import numpy as np
import scipy.ndimage as ndi
from scipy.spatial import cKDTree

METHODS = ("adaptive", "percentile", "fixed", "otsu")


def median_background(image, filter_radius=101, median_passes=1):
    """compute_background without diplib (preprocess.py:43-64): zeros when
    filter_radius <= 0 or median_passes <= 0 (:53-54), else
    scipy.ndimage.median_filter(bg, size=filter_radius) once per pass
    (:61-65)."""
    if filter_radius <= 0 or median_passes <= 0:
        return np.zeros_like(image, dtype=np.float64)
    bg = image.astype(np.float64)
    for _ in range(median_passes):
        bg = ndi.median_filter(bg, size=filter_radius)
    return bg


def median_brute(image, size):
    """One median pass by sorting every window: a size x size window whose
    first row and column lie size // 2 before the pixel, the image extended by
    numpy's 'symmetric' padding (repeated for windows wider than the image),
    the element of rank size*size // 2.  For small frames only."""
    lo = size // 2
    a = np.asarray(image, dtype=np.float64)
    p = np.pad(a, ((lo, size - 1 - lo), (lo, size - 1 - lo)), mode="symmetric")
    win = np.lib.stride_tricks.sliding_window_view(p, (size, size))
    k = (size * size) // 2
    return np.sort(win.reshape(a.shape[0], a.shape[1], -1), axis=-1)[:, :, k]


def noise_floor(image, k=4.0):
    """median(nz) + k * 1.4826 * MAD(nz), at least 1 (thresholds.py:36-47)."""
    nz = image[image > 0]
    if nz.size:
        med = float(np.median(nz))
        mad = 1.4826 * float(np.median(np.abs(nz - med)))
        return max(med + k * mad, 1.0)
    return 1.0


def threshold_value(image, method, fixed_value, percentile):
    """String dispatch of thresholds.py:92-110 (otsu -> adaptive without
    scikit-image)."""
    if method == "percentile":
        return float(np.percentile(image.ravel(), percentile))
    if method == "fixed":
        return float(fixed_value)
    return noise_floor(image, k=4.0)


def components(u16):
    """8-connected components of the nonzero pixels (preprocess.py:136-155)."""
    labels, n = ndi.label(u16 > 0, structure=np.ones((3, 3), dtype=int))
    return labels, n


def filter_small(u16, labels, n, min_area):
    """Zero the components with fewer than min_area pixels; centres of mass of
    the others on their bounding boxes (preprocess.py:158-195)."""
    filt = u16.copy()
    centres = []
    objs = ndi.find_objects(labels)
    areas = np.bincount(labels.ravel(), minlength=n + 1)
    for lbl in range(1, n + 1):
        sl = objs[lbl - 1]
        if areas[lbl] >= min_area:
            y0, x0 = sl[0].start, sl[1].start
            mask = labels[sl] == lbl
            region = u16[sl] * mask
            com = ndi.center_of_mass(region)
            centres.append((com[1] + x0, com[0] + y0))
        else:
            filt[labels == lbl] = 0
    return filt, centres


def gaussian_sigma(centres, pixel_size, distance, orient_spacing):
    """sigma from the nearest pair of centres (preprocess.py:198-218)."""
    if not centres or len(centres) < 2:
        return 3.0
    coords = np.array(centres)
    dists, _ = cKDTree(coords).query(coords, k=2)
    min_px_dist = float(dists[:, 1].min())
    if pixel_size > 0 and distance > 0:
        delta = (distance * np.tan(np.radians(orient_spacing))) / pixel_size
    else:
        delta = min_px_dist
    sigma = 0.25 * min(min_px_dist, delta) if delta > 0 else 0.25 * min_px_dist
    return max(sigma, 1.0)


def preprocess(raw, background=None, method="adaptive", fixed_value=0.0, percentile=90.0,
               min_area=10, pixel_size=0.2, distance=0.513, orient_spacing=0.4,
               filter_radius=0, median_passes=1):
    """Steps 1, 3-6 of preprocess_image (preprocess.py:255-337), enhancement
    off (its default); without a background it is computed from the raw
    image (no background file: :255-266).  Returns the float64 blurred image
    and a dict with the threshold, sigma and component counts."""
    bg = (median_background(raw, filter_radius, median_passes) if background is None
          else background)
    bg_sub = np.maximum(raw - bg, 0.0)
    t = threshold_value(bg_sub, method, fixed_value, percentile)
    thr = bg_sub.copy()
    thr[bg_sub <= t] = 0
    u16 = thr.astype(np.uint16)
    labels, n = components(u16)
    info = {"threshold": t, "n_components": int(n), "n_kept": 0, "sigma": 0.0}
    if n == 0:
        return np.zeros_like(raw, dtype=np.float64), info
    filt, centres = filter_small(u16, labels, n, min_area)
    info["n_kept"] = len(centres)
    if not centres:
        return np.zeros_like(raw, dtype=np.float64), info
    sigma = gaussian_sigma(centres, pixel_size, distance, orient_spacing)
    info["sigma"] = sigma
    return ndi.gaussian_filter(filt.astype(np.float64), sigma), info
