"""ctypes binding of the CPU oracle (oracle/liblmoracle.so) for the tests.

The oracle is test infrastructure: it is loaded only by tests/, by
``__graft_entry__.smoke()`` and by bench.py's CPU-baseline leg.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIB = None


class OGeom(ctypes.Structure):
    _fields_ = [("recip", ctypes.c_double * 9), ("rotT", ctypes.c_double * 9),
                ("P", ctypes.c_double * 3), ("pxX", ctypes.c_double), ("pxY", ctypes.c_double),
                ("nrPxX", ctypes.c_int), ("nrPxY", ctypes.c_int),
                ("Elo", ctypes.c_double), ("Ehi", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ORACLE_DIR, "liblmoracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", ORACLE_DIR, "liblmoracle.so"], check=True,
                           capture_output=True)
        L = ctypes.CDLL(path)
        vp, sz, dbl, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int
        L.lmo_forward.restype = ctypes.c_long
        L.lmo_forward.argtypes = [vp, sz, vp, i32, ctypes.POINTER(OGeom), i32, vp, vp, i32, dbl,
                                  vp, i32]
        L.lmo_quantize_u8.restype = dbl
        L.lmo_quantize_u8.argtypes = [vp, sz, vp]
        L.lmo_compare_u8.restype = ctypes.c_long
        L.lmo_compare_u8.argtypes = [vp, sz, i32, vp, i32, dbl, i32, dbl, vp, vp, vp, i32]
        L.lmo_collect.restype = ctypes.c_long
        L.lmo_collect.argtypes = [vp, sz, vp, vp, ctypes.c_long]
        L.lmo_project.restype = i32
        L.lmo_project.argtypes = [vp, vp, ctypes.POINTER(OGeom), ctypes.POINTER(i32),
                                  ctypes.POINTER(i32), vp, ctypes.POINTER(dbl)]
        L.lmo_calc_recip.argtypes = [vp, i32, vp]
        L.lmo_detector_rotT.argtypes = [vp, vp]
        L.lmo_median_ranks.restype = i32
        L.lmo_median_ranks.argtypes = [vp, i32, i32, i32, ctypes.c_uint32, vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def calc_recip(lat, sg):
    out = np.zeros(9)
    lib().lmo_calc_recip(_p(np.asarray(lat, dtype=np.float64)), sg, _p(out))
    return out.reshape(3, 3)


def detector_rotT(r):
    out = np.zeros(9)
    lib().lmo_detector_rotT(_p(np.asarray(r, dtype=np.float64)), _p(out))
    return out.reshape(3, 3)


def geom(params) -> OGeom:
    """Geometry from a synth.Params-like object (attribute names of the
    reference parameter file)."""
    g = OGeom()
    g.recip[:] = calc_recip(params.LatticeParameter, params.SpaceGroup).ravel()
    g.rotT[:] = detector_rotT(params.R_Array).ravel()
    g.P[:] = list(params.P_Array)
    g.pxX, g.pxY = params.PxX, params.PxY
    g.nrPxX, g.nrPxY = params.NrPxX, params.NrPxY
    g.Elo, g.Ehi = params.Elo, params.Ehi
    return g


def project(this_orient, hkl, g):
    ipx, ipy, E = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(np.nan)
    qh = np.zeros(3)
    m = np.ascontiguousarray(this_orient, dtype=np.float64)
    h = np.ascontiguousarray(hkl, dtype=np.int32)
    ok = lib().lmo_project(_p(m), _p(h), ctypes.byref(g), ctypes.byref(ipx), ctypes.byref(ipy),
                           _p(qh), ctypes.byref(E))
    return bool(ok), ipx.value, ipy.value, qh, E.value


def forward(orients, hkls, g, max_spots, img=None, min_spots=5, min_int=1000.0, threads=8):
    o = np.ascontiguousarray(orients, dtype=np.float64).reshape(-1, 9)
    h = np.ascontiguousarray(hkls, dtype=np.int32).reshape(-1, 3)
    n = o.shape[0]
    table = np.zeros((n, 1 + 2 * max_spots), dtype=np.uint16)
    matched = np.zeros(n)
    im = None if img is None else np.ascontiguousarray(img, dtype=np.float64)
    hits = lib().lmo_forward(_p(o), n, _p(h), h.shape[0], ctypes.byref(g), max_spots,
                             _p(table), None if im is None else _p(im), min_spots, min_int,
                             _p(matched), threads)
    return table, matched, hits


def quantize(img):
    im = np.ascontiguousarray(img, dtype=np.float64).ravel()
    out = np.zeros(im.size, dtype=np.uint8)
    s = lib().lmo_quantize_u8(_p(im), im.size, _p(out))
    return out, s


def compare(table, max_spots, img_u8, nrPxX, im_scale, min_spots, min_int, threads=8):
    t = np.ascontiguousarray(table, dtype=np.uint16)
    n = t.shape[0]
    u8 = np.ascontiguousarray(img_u8, dtype=np.uint8).ravel()
    matched = np.zeros(n)
    ns = np.zeros(n, dtype=np.int32)
    tot = np.zeros(n)
    hits = lib().lmo_compare_u8(_p(t), n, max_spots, _p(u8), nrPxX, im_scale, min_spots,
                                min_int, _p(matched), _p(ns), _p(tot), threads)
    return matched, ns, tot, hits


def collect(matched):
    m = np.ascontiguousarray(matched, dtype=np.float64)
    k = lib().lmo_collect(_p(m), m.size, None, None, 0)
    rows = np.zeros(k, dtype=np.uint64)
    mA = np.zeros(k)
    lib().lmo_collect(_p(m), m.size, _p(rows), _p(mA), k)
    return rows, mA


def _merge_lib():
    L = lib()
    if not getattr(L, "_merge_ready", False):
        vp = ctypes.c_void_p
        L.lmo_quat_from_matrix.restype = None
        L.lmo_quat_from_matrix.argtypes = [vp, vp]
        L.lmo_misorientation_deg.restype = ctypes.c_double
        L.lmo_misorientation_deg.argtypes = [vp, vp, ctypes.c_int]
        L.lmo_merge.restype = ctypes.c_int
        L.lmo_merge.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, vp, vp, vp]
        L._merge_ready = True
    return L


def quat_from_matrix(m):
    q = np.zeros(4)
    _merge_lib().lmo_quat_from_matrix(_p(np.ascontiguousarray(m, dtype=np.float64).ravel()), _p(q))
    return q


def misorientation(q1, q2, sg=225):
    return _merge_lib().lmo_misorientation_deg(_p(np.ascontiguousarray(q1, dtype=np.float64)),
                                               _p(np.ascontiguousarray(q2, dtype=np.float64)), sg)


def merge(orients, rows, scores, max_angle, sg=225, threads=8):
    """Oracle merge: (best_row, best_score, cluster_size) per cluster."""
    o = np.ascontiguousarray(orients, dtype=np.float64).reshape(-1, 9)
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    s = np.ascontiguousarray(scores, dtype=np.float64)
    nr = int(r.size)
    if nr and int(r.max()) >= o.shape[0]:
        raise ValueError("candidate row outside the orientation DB")
    cap = max(1, min(nr, 50000))
    best = np.zeros(cap, np.uint64)
    bs = np.zeros(cap)
    cs = np.zeros(cap, np.int32)
    k = _merge_lib().lmo_merge(_p(o), _p(r), _p(s), nr, max_angle, sg, threads, _p(best),
                               _p(bs), _p(cs))
    return best[:k], bs[:k], cs[:k]


def median_background(image, size, passes=1):
    """compute_background's scipy branch (laue_index/preprocess.py:59-63)
    through the C oracle: the frame's values ranked densely (np.unique),
    lmo_median_ranks once per pass, ranks mapped back to values; zeros for
    size <= 0 or passes <= 0 (:53-54)."""
    a = np.asarray(image, dtype=np.float64)
    if size <= 0 or passes <= 0:
        return np.zeros_like(a)
    H, W = a.shape
    vals, inv = np.unique(a, return_inverse=True)
    rk = np.ascontiguousarray(inv.reshape(H, W), dtype=np.uint32)
    out = np.empty_like(rk)
    for _ in range(passes):
        rc = lib().lmo_median_ranks(_p(rk), W, H, int(size), int(vals.size), _p(out))
        if rc:
            raise RuntimeError(f"lmo_median_ranks failed ({rc})")
        rk, out = out, rk
    return vals[rk]
