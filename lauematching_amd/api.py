"""Python host interface over the C ABI (include/lauematch.h).

Mirrors the reference hot path's data flow (``src/LaueMatchingCPU.c``
main(): parameters -> orientation DB + HKLs -> forward table -> quantized
image -> compare -> ascending hit list) with objects that own the device
state:

>>> ctx = Context()
>>> db = OrientationDB(ctx, orients, hkls, Geometry.from_params(p), p.MaxNrLaueSpots)
>>> db.forward()                                   # DoFwd=1 table build
>>> u8, scale = quantize_u8(image_f64)
>>> hits = db.score(u8, scale, p.MinNrSpots, p.MinIntensity)
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (LM_FWD_FP32_GUARD, LM_FWD_FP64, LmFwdStats, LmGeom, LmHits, check, lib)

__all__ = ["Context", "Geometry", "OrientationDB", "Hits", "quantize_u8", "read_params",
           "LM_FWD_FP64", "LM_FWD_FP32_GUARD"]


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """One HIP device plus an HBM budget (0 = 90% of free memory)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = ctypes.c_void_p()
        check(lib().lm_ctx_create(device, hbm_budget, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def bytes_in_use(self) -> int:
        return int(lib().lm_ctx_bytes_in_use(self._h))

    def close(self):
        if self._h:
            lib().lm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Geometry:
    """Detector + crystal geometry (reference parameter-file units)."""

    def __init__(self, lattice, space_group, P, R, pxX, pxY, nrPxX, nrPxY, Elo, Ehi):
        g = LmGeom()
        lat = (ctypes.c_double * 6)(*lattice)
        PP = (ctypes.c_double * 3)(*P)
        RR = (ctypes.c_double * 3)(*R)
        check(lib().lm_geom_init(lat, int(space_group), PP, RR, float(pxX), float(pxY),
                                 int(nrPxX), int(nrPxY), float(Elo), float(Ehi), ctypes.byref(g)))
        self.c = g
        self.nrPxX, self.nrPxY = int(nrPxX), int(nrPxY)

    @classmethod
    def from_params(cls, p) -> "Geometry":
        return cls(p.LatticeParameter, p.SpaceGroup, p.P_Array, p.R_Array, p.PxX, p.PxY,
                   p.NrPxX, p.NrPxY, p.Elo, p.Ehi)

    @property
    def recip(self) -> np.ndarray:
        return np.array([[self.c.recip[i][j] for j in range(3)] for i in range(3)])

    @property
    def rotT(self) -> np.ndarray:
        return np.array([[self.c.rotT[i][j] for j in range(3)] for i in range(3)])


@dataclass
class Hits:
    """Ascending hit list of one image (reference src/LaueMatchingCPU.c:772-788)."""

    nr_results: int
    ids: np.ndarray
    nspots: np.ndarray
    totint: np.ndarray
    score: np.ndarray
    kernel_ms: float
    walk_ms: float = 0.0

    @classmethod
    def _take(cls, h: LmHits) -> "Hits":
        k = int(h.count)
        if k:
            def grab(ptr, dtype):  # one byte copy out of the C arrays (writable result)
                a = np.empty(k, dtype)
                ctypes.memmove(a.ctypes.data, ptr, k * a.itemsize)
                return a

            ids = grab(h.ids, np.uint64)
            ns = grab(h.nspots, np.int32)
            tot = grab(h.totint, np.float64)
            sc = grab(h.score, np.float64)
        else:
            ids = np.zeros(0, np.uint64)
            ns = np.zeros(0, np.int32)
            tot = np.zeros(0)
            sc = np.zeros(0)
        out = cls(int(h.nr_results), ids, ns, tot, sc, float(h.kernel_ms), float(h.walk_ms))
        lib().lm_hits_free(ctypes.byref(h))
        return out


class OrientationDB:
    """Orientation DB + HKL list resident on the device, with its spot table."""

    def __init__(self, ctx: Context, orients: np.ndarray, hkls: np.ndarray, geom: Geometry,
                 max_spots: int):
        o = np.ascontiguousarray(orients, dtype=np.float64).reshape(-1, 9)
        h = np.ascontiguousarray(hkls, dtype=np.int32).reshape(-1, 3)
        hd = ctypes.c_void_p()
        check(lib().lm_db_create(ctx.handle, _ptr(o), o.shape[0], _ptr(h), h.shape[0],
                                 ctypes.byref(geom.c), int(max_spots), ctypes.byref(hd)))
        self._h = hd
        self.ctx = ctx
        self.geom = geom
        self.n = o.shape[0]
        self.max_spots = int(max_spots)

    def close(self):
        if self._h:
            lib().lm_db_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def table_bytes(self) -> int:
        return int(lib().lm_db_table_bytes(self._h))

    @property
    def table_spots(self) -> int:
        return int(lib().lm_db_table_spots(self._h))

    def set_banded_tile_bits(self, bits: int) -> None:
        """Super-tile size (log2, 13..16) of the pixel-banded copy; 0 drops it.
        16 keeps 16-bit row sums (MaxNrLaueSpots <= 257) and recounts rows
        whose sum reaches MinIntensity."""
        check(lib().lm_db_set_option(self._h, 1, int(bits)))

    def set_batch_tile_bits(self, bits: int) -> None:
        """Super-tile size (log2, 9..13) of the batch banded copy; 0 drops it."""
        check(lib().lm_db_set_option(self._h, 2, int(bits)))

    def set_batch_width_tile_bits(self, width: int, bits: int) -> None:
        """Give batches of `width` images (4..64) their own batch copy of
        2^bits-row super-tiles (9..13; 0 = none)."""
        check(lib().lm_db_set_option(self._h, 9, (int(width).bit_length() - 1) << 8 | int(bits)))

    def set_batch_widths(self, mask: int) -> None:
        """Bit k set: batches of width 2**k read the batch banded copy."""
        check(lib().lm_db_set_option(self._h, 3, int(mask)))

    def set_banded_loads(self, loads: int) -> None:
        """Single-image banded walk: 0 = 4-byte loads x8, 1 = 16-byte x2, 2 = 16-byte x4,
        3 = image staged through LDS (2^14..2^16 super-tiles), 4 = 4-byte gathers,
        5, 6, 7 = LDS-staged with (entries per lane, chunk sets in flight)
        (8, 2), (4, 3), (8, 3); 3 is (4, 2); twice the entries for 2^16."""
        check(lib().lm_db_set_option(self._h, 5, int(loads)))

    def set_batch_grouped(self, mask: int) -> None:
        """Bit 6 / bit 7 set: table batches of 64 / 128 images use 16 images
        per lane (several lanes per orientation) instead of one lane holding
        every image."""
        check(lib().lm_db_set_option(self._h, 6, int(mask)))

    def set_batch_nibble(self, mask: int) -> None:
        """Bit 7 / bit 8 set: batches of 128 / 129..256 images read 4-bit
        upper bounds of the pixels and recount their candidates exactly;
        clear: 128 images take the uint8 walks, wider batches two calls."""
        check(lib().lm_db_set_option(self._h, 7, int(mask)))

    def set_batch_sum16(self, mask: int) -> None:
        """Bit k set (k = 2..6): batches of width 2**k read their batch copy
        with 16-bit sums (two images per LDS word, copies with
        width x 2^bits <= 65536) and recount the candidates exactly; set it
        before set_batch_width_tile_bits."""
        check(lib().lm_db_set_option(self._h, 10, int(mask)))

    def set_batch_split(self, mask: int) -> None:
        """Bit k set (k = 3..6): batches of width 2**k are scored as two calls
        of at most 2**(k-1) images; bit 2: 2..4 images one at a time."""
        check(lib().lm_db_set_option(self._h, 11, int(mask)))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64()
        check(lib().lm_db_get_option(self._h, int(option), ctypes.byref(v)))
        return v.value

    @property
    def table_on_host(self) -> bool:
        """True when the table did not fit the HBM budget and is kept in host
        memory, streamed to the device in chunks by every scoring call."""
        return self.get_option(8) == 1

    def set_banded_split(self, bits: int) -> None:
        """LDS-staged walk: each image chunk's entries in 2**bits balanced
        pieces (0..3), or -1 for the fewest giving pieces of at most 512
        entries on average."""
        check(lib().lm_db_set_option(self._h, 12, int(bits)))

    def banded_layout(self):
        """(super-tile bits, walk) of the single-image layout in use."""
        return self.get_option(1), self.get_option(5)

    def set_use_graphs(self, on: bool) -> None:
        """Replay captured launch sequences (hipGraph) in direct scoring calls."""
        check(lib().lm_db_set_option(self._h, 4, int(bool(on))))

    def autotune(self, sample_imgs_u8, max_batch: int = 64, reps: int = 3) -> dict:
        """Time every scoring layout on the sample images and keep the fastest
        (single-image banded copy, batch copy super-tile, widths using it)."""
        ims = [np.ascontiguousarray(i, dtype=np.uint8).ravel() for i in sample_imgs_u8]
        npx = self.geom.nrPxX * self.geom.nrPxY
        if not ims or any(i.size != npx for i in ims):
            raise ValueError("sample images must match the detector")
        ptrs = (ctypes.c_void_p * len(ims))(*[i.ctypes.data for i in ims])
        r = _lib.LmTuneReport()
        check(lib().lm_db_autotune(self._h, ptrs, len(ims), int(max_batch), int(reps),
                                   ctypes.byref(r)))
        widths = [r.widths[i] for i in range(r.n_widths)]
        single = {"table": r.ms_single_table}
        for i, tb in enumerate((13, 14, 15, 16)):
            for loads in range(8):
                if r.ms_single[i][loads] > 0:
                    single[f"banded{tb}_loads{loads}"] = r.ms_single[i][loads]
        return {
            "ms_single": single, "banded_loads": r.banded_loads,
            "ms_batch": {B: {**({"table": r.ms_lane[i]} if r.ms_lane[i] > 0 else {}),
                             **({"grouped": r.ms_grouped[i]} if r.ms_grouped[i] > 0 else {}),
                             **({"nibble": r.ms_nibble[i]} if r.ms_nibble[i] > 0 else {}),
                             **{f"banded{tb}": r.ms_banded[i][tb - 9]
                                for tb in range(9, 14) if r.ms_banded[i][tb - 9] > 0},
                             **{f"banded{tb}_sum16": r.ms_banded16[i][tb - 9]
                                for tb in range(9, 14) if r.ms_banded16[i][tb - 9] > 0}}
                         for i, B in enumerate(widths)},
            "banded_tile_bits": r.banded_tile_bits, "batch_tile_bits": r.batch_tile_bits,
            "batch_widths_mask": r.batch_widths_mask, "grouped_mask": r.grouped_mask,
            "nibble_mask": r.nibble_mask, "sum16_mask": r.sum16_mask,
            "split_mask": r.split_mask, "banded_split": r.banded_split,
            "ms_split": {1 << k: r.ms_split[k] for k in range(4) if r.ms_split[k] > 0},
            "batch_tile_bits_per_width": {B: r.batch_tile_bits_w[i] for i, B in enumerate(widths)},
        }

    def forward(self, mode: int = LM_FWD_FP64) -> LmFwdStats:
        st = LmFwdStats()
        check(lib().lm_db_forward(self._h, int(mode), ctypes.byref(st)))
        return st

    def checksum(self) -> int:
        """64-bit position-tagged checksum of the resident table."""
        s = ctypes.c_uint64()
        check(lib().lm_db_table_checksum(self._h, ctypes.byref(s)))
        return s.value

    def import_table(self, records: np.ndarray) -> None:
        r = np.ascontiguousarray(records, dtype=np.uint16).reshape(-1, 1 + 2 * self.max_spots)
        check(lib().lm_db_import_table(self._h, _ptr(r), r.shape[0]))

    def export_table(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.n - first if count is None else count
        out = np.zeros((count, 1 + 2 * self.max_spots), dtype=np.uint16)
        check(lib().lm_db_export_table(self._h, first, count, _ptr(out)))
        return out

    def score(self, img_u8: np.ndarray, im_scale: float, min_nr_spots: int,
              min_intensity: float) -> Hits:
        im = np.ascontiguousarray(img_u8, dtype=np.uint8).ravel()
        if im.size != self.geom.nrPxX * self.geom.nrPxY:
            raise ValueError("image size does not match the detector")
        h = LmHits()
        check(lib().lm_score(self._h, _ptr(im), float(im_scale), int(min_nr_spots),
                             float(min_intensity), ctypes.byref(h)))
        return Hits._take(h)

    def score_f64(self, img: np.ndarray, min_nr_spots: int, min_intensity: float) -> Hits:
        """Float64-image scoring of the reference DoFwd=1 branch."""
        im = np.ascontiguousarray(img, dtype=np.float64).ravel()
        if im.size != self.geom.nrPxX * self.geom.nrPxY:
            raise ValueError("image size does not match the detector")
        h = LmHits()
        check(lib().lm_score_f64(self._h, _ptr(im), int(min_nr_spots), float(min_intensity),
                                 ctypes.byref(h)))
        return Hits._take(h)

    def score_dev(self, dev_ptr: int, im_scale: float, min_nr_spots: int, min_intensity: float,
                  stream: int = 0) -> Hits:
        h = LmHits()
        check(lib().lm_score_dev(self._h, ctypes.c_void_p(dev_ptr), float(im_scale),
                                 int(min_nr_spots), float(min_intensity),
                                 ctypes.c_void_p(stream), ctypes.byref(h)))
        return Hits._take(h)

    def score_batch(self, imgs_u8, im_scales, min_nr_spots: int, min_intensity: float):
        ims = [np.ascontiguousarray(i, dtype=np.uint8).ravel() for i in imgs_u8]
        npx = self.geom.nrPxX * self.geom.nrPxY
        if any(i.size != npx for i in ims):
            raise ValueError("image size does not match the detector")
        n = len(ims)
        ptrs = (ctypes.c_void_p * n)(*[i.ctypes.data for i in ims])
        sc = (ctypes.c_double * n)(*[float(s) for s in im_scales])
        hs = (LmHits * n)()
        check(lib().lm_score_batch(self._h, n, ptrs, sc, int(min_nr_spots), float(min_intensity),
                                   hs))
        return [Hits._take(hs[i]) for i in range(n)]

    def score_batch_each(self, imgs_u8, im_scales, min_nr_spots, min_intensity):
        """Batch with thresholds per image (sequences as long as the batch)."""
        ims = [np.ascontiguousarray(i, dtype=np.uint8).ravel() for i in imgs_u8]
        npx = self.geom.nrPxX * self.geom.nrPxY
        n = len(ims)
        if any(i.size != npx for i in ims):
            raise ValueError("image size does not match the detector")
        if not (len(im_scales) == len(min_nr_spots) == len(min_intensity) == n):
            raise ValueError("one scale and one threshold pair per image")
        ptrs = (ctypes.c_void_p * n)(*[i.ctypes.data for i in ims])
        sc = (ctypes.c_double * n)(*[float(s) for s in im_scales])
        sp = (ctypes.c_int32 * n)(*[int(v) for v in min_nr_spots])
        mi = (ctypes.c_double * n)(*[float(v) for v in min_intensity])
        hs = (LmHits * n)()
        check(lib().lm_score_batch_each(self._h, n, ptrs, sc, sp, mi, hs))
        return [Hits._take(hs[i]) for i in range(n)]

    def score_batch_dev(self, dev_ptr: int, nimg: int, im_scales, min_nr_spots: int,
                        min_intensity: float, stream: int = 0):
        sc = (ctypes.c_double * nimg)(*[float(s) for s in im_scales])
        hs = (LmHits * nimg)()
        check(lib().lm_score_batch_dev(self._h, nimg, ctypes.c_void_p(dev_ptr), sc,
                                       int(min_nr_spots), float(min_intensity),
                                       ctypes.c_void_p(stream), hs))
        return [Hits._take(hs[i]) for i in range(nimg)]

    def score_fwd_batch_dev(self, dev_ptr: int, nimg: int, im_scales, min_nr_spots,
                            min_intensity, mode: int = LM_FWD_FP64, stream: int = 0):
        """On-the-fly scoring (lm_score_fwd_batch_dev): nimg device images
        back to back, thresholds per image (sequences, or one value for all);
        every orientation is projected in the same pass, no table is read.
        Returns (hits per image, LmFwdStats)."""
        if np.isscalar(min_nr_spots):
            min_nr_spots = [min_nr_spots] * nimg
        if np.isscalar(min_intensity):
            min_intensity = [min_intensity] * nimg
        sc = (ctypes.c_double * nimg)(*[float(s) for s in im_scales])
        sp = (ctypes.c_int32 * nimg)(*[int(v) for v in min_nr_spots])
        mi = (ctypes.c_double * nimg)(*[float(v) for v in min_intensity])
        hs = (LmHits * nimg)()
        st = LmFwdStats()
        check(lib().lm_score_fwd_batch_dev(self._h, nimg, ctypes.c_void_p(dev_ptr), sc, sp, mi,
                                           int(mode), ctypes.c_void_p(stream), hs,
                                           ctypes.byref(st)))
        return [Hits._take(hs[i]) for i in range(nimg)], st


def merge_candidates(orients, rows, scores, max_angle: float, space_group: int,
                     ctx: "Context | None" = None, threads: int = 8):
    """Merge of duplicate candidates (mergeDuplicateOrientations,
    src/LaueMatchingHeaders.h:969-1065): greedy clustering in list order by
    misorientation <= max_angle.  With a context the pair tests run on its
    device (same result).  Returns (best rows, best scores, cluster sizes)."""
    o = np.ascontiguousarray(orients, dtype=np.float64).reshape(-1, 9)
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    s = np.ascontiguousarray(scores, dtype=np.float64)
    nr = int(r.size)
    if s.size != nr:
        raise ValueError("one score per candidate row")
    cap = max(1, min(nr, 50000))
    best = np.zeros(cap, np.uint64)
    bs = np.zeros(cap)
    cs = np.zeros(cap, np.int32)
    k = ctypes.c_int32()
    if ctx is None:
        check(lib().lm_merge_candidates(_ptr(o), o.shape[0], _ptr(r), _ptr(s), nr,
                                        float(max_angle), int(space_group), int(threads),
                                        _ptr(best), _ptr(bs), _ptr(cs), ctypes.byref(k)))
    else:
        check(lib().lm_merge_candidates_dev(ctx.handle, _ptr(o), o.shape[0], _ptr(r), _ptr(s),
                                            nr, float(max_angle), int(space_group), _ptr(best),
                                            _ptr(bs), _ptr(cs), ctypes.byref(k)))
    return best[:k.value], bs[:k.value], cs[:k.value]


PREP_METHODS = {"adaptive": 0, "percentile": 1, "fixed": 2, "otsu": 3}


def preprocess(ctx: "Context", raw: np.ndarray, background: np.ndarray | None = None,
               method: str = "adaptive", fixed_value: float = 0.0, percentile: float = 90.0,
               min_area: int = 10, pixel_size: float = 0.2, distance: float = 0.513,
               orient_spacing: float = 0.4, d_u8: int = 0, want_blurred: bool = False,
               filter_radius: int = 0, median_passes: int = 1):
    """Image preprocessing on the device (preprocess_image,
    laue_index/preprocess.py:221-337, enhancement off): background
    subtraction, threshold, uint16 cast, 8-connected components with the
    small ones removed, Gaussian blur, uint8 quantization.  Without a
    background and with filter_radius > 0 the background is the median
    filter of median_background (preprocess.py:255-266).  Returns (u8 image,
    imScale, info dict[, blurred float64 image]); with d_u8 (a device pointer
    of npx bytes) the uint8 image is also left there for scoring."""
    img = np.ascontiguousarray(raw, dtype=np.float64)
    if img.ndim != 2:
        raise ValueError("raw must be a 2-D (NrPxY, NrPxX) image")
    H, W = img.shape
    bg = None
    if background is not None:
        bg = np.ascontiguousarray(background, dtype=np.float64)
        if bg.shape != img.shape:
            raise ValueError("background must have the image's shape")
    cfg = _lib.LmPrepConfig(PREP_METHODS[method], float(fixed_value), float(percentile),
                            int(min_area), float(pixel_size), float(distance),
                            float(orient_spacing), int(filter_radius), int(median_passes))
    u8 = np.zeros((H, W), np.uint8)
    blurred = np.zeros((H, W)) if want_blurred else None
    res = _lib.LmPrepResult()
    check(lib().lm_preprocess(ctx.handle, W, H, _ptr(img), None if bg is None else _ptr(bg),
                              ctypes.byref(cfg), _ptr(u8), ctypes.c_void_p(d_u8 or None),
                              None if blurred is None else _ptr(blurred), ctypes.byref(res)))
    info = {"threshold": res.threshold, "sigma": res.sigma, "n_components": res.n_components,
            "n_kept": res.n_kept, "ms": res.ms, "n_above": res.n_above}
    if want_blurred:
        return u8, res.im_scale, info, blurred
    return u8, res.im_scale, info


def median_background(ctx: "Context", image: np.ndarray, filter_radius: int = 101,
                      median_passes: int = 1):
    """compute_background (laue_index/preprocess.py:43-64) on the device:
    scipy.ndimage.median_filter(size=filter_radius) applied median_passes
    times, zeros when either is <= 0.  Returns (background float64 image,
    device ms)."""
    img = np.ascontiguousarray(image, dtype=np.float64)
    if img.ndim != 2:
        raise ValueError("image must be a 2-D (NrPxY, NrPxX) array")
    H, W = img.shape
    out = np.empty((H, W), np.float64)
    ms = ctypes.c_double()
    check(lib().lm_median_background(ctx.handle, W, H, _ptr(img), int(filter_radius),
                                     int(median_passes), _ptr(out), ctypes.byref(ms)))
    return out, ms.value


def quantize_u8(img: np.ndarray):
    """uint8 quantization + imScale (reference src/LaueMatchingCPU.c:355-377)."""
    im = np.ascontiguousarray(img, dtype=np.float64).ravel()
    out = np.zeros(im.size, dtype=np.uint8)
    s = ctypes.c_double()
    check(lib().lm_quantize_u8(_ptr(im), im.size, _ptr(out), ctypes.byref(s)))
    return out, s.value


_PARAM_KEYS = [  # (key, kind, count) in the reference's matching order
    ("LatticeParameter", float, 6), ("P_Array", float, 3), ("R_Array", float, 3),
    ("tol_c_over_a", float, 1), ("PxX", float, 1), ("PxY", float, 1), ("Elo", float, 1),
    ("Ehi", float, 1), ("DoFwd", int, 1), ("NrPxX", int, 1), ("NrPxY", int, 1),
    ("MaxNrLaueSpots", int, 1), ("BatchSize", int, 1), ("MinNrSpots", int, 1),
    ("SpaceGroup", int, 1), ("MinIntensity", float, 1), ("MaxAngle", float, 1),
    ("ForwardFile", str, 1), ("tol_LatC", float, 6), ("Optimizer", str, 1),
    ("ResultDir", str, 1), ("MinGoodSpots", int, 1),
]


def read_params(path: str):
    """Parameter file -> synth.Params, with the reference's prefix matching
    (src/LaueMatchingCPU.c:75-204) and C defaults (:60-70)."""
    from .synth import Params

    p = Params(MaxNrLaueSpots=500, MinNrSpots=5, MinIntensity=1000.0, DoFwd=1, SpaceGroup=225,
               Elo=5.0, Ehi=30.0, MaxAngle=2.0, NrPxX=0, NrPxY=0, PxX=0.0, PxY=0.0,
               P_Array=(0.0, 0.0, 0.0), R_Array=(0.0, 0.0, 0.0),
               LatticeParameter=(0.0,) * 6, ForwardFile="")
    with open(path) as f:
        for line in f:
            for key, kind, cnt in _PARAM_KEYS:
                if not line.startswith(key):
                    continue
                vals = line.split()[1:1 + cnt]
                try:
                    conv = [kind(v) if kind is not int else int(float(v)) for v in vals]
                except ValueError:
                    conv = []
                if conv:
                    if cnt == 1:
                        v = conv[0]
                    else:
                        v = tuple(conv + [0.0] * (cnt - len(conv)))
                    if hasattr(p, key):
                        setattr(p, key, v)
                    else:
                        p.extra[key] = v
                break
    return p
