"""Seeded synthetic inputs for the indexing hot path.

Everything here writes the reference's own file formats:

* orientation DB: N x 9 float64, row-major 3x3 matrices, no header
  (reference ``100MilOrients.bin.meta.json:10-12``);
* HKL file: text lines ``h k l |hkl|^2`` whose first three integers the C
  reader takes (``src/LaueMatchingCPU.c:289-312``);
* parameter file: ``Key value ...`` lines, prefix-matched by the C parser
  (``src/LaueMatchingCPU.c:75-209``);
* image: raw float64 W x H, row-major, X fastest (``scripts/RunImage.py:290-292``).

The image pipeline follows the benchmark description: Gaussian spots of k
seeded grains (sigma 1.5 px, amplitude U(200, 4000)) on a Poisson(50)
background, threshold at the 99th percentile, Gaussian blur sigma 2 px, and
quantization to integers 0..255 with maximum exactly 255, so that the
indexer's own quantization scale (max/255) is exactly 1.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

HC_KEVNM = 1.2398419739  # src/LaueMatchingHeaders.h:41

#: Rodrigues vector that puts 2theta = 90 deg (kf = +y) on the detector centre
#: for a detector at P = (0, 0, z): rotation by 120 deg about -(1,1,1)/sqrt(3).
R_ARRAY_90DEG = -(2.0 * math.pi / 3.0) / math.sqrt(3.0)

ORIENT_CHUNK = 1 << 20


# --------------------------------------------------------------------------
# Geometry parameters
# --------------------------------------------------------------------------

@dataclass
class Params:
    """Hot-path subset of the reference parameter file (same keys/units)."""

    LatticeParameter: tuple = (0.36, 0.36, 0.36, 90.0, 90.0, 90.0)
    SpaceGroup: int = 225
    P_Array: tuple = (0.0, 0.0, 0.5)
    R_Array: tuple = (R_ARRAY_90DEG, R_ARRAY_90DEG, R_ARRAY_90DEG)
    PxX: float = 0.0002
    PxY: float = 0.0002
    NrPxX: int = 2048
    NrPxY: int = 2048
    Elo: float = 5.0
    Ehi: float = 30.0
    MaxNrLaueSpots: int = 64
    MinNrSpots: int = 5
    MinIntensity: float = 1000.0
    MaxAngle: float = 2.0
    DoFwd: int = 1
    ForwardFile: str = "/dev/shm/lm_fwd.bin"
    extra: dict = field(default_factory=dict)

    def write(self, path: str) -> None:
        def fmt(v):
            if isinstance(v, (tuple, list)):
                return " ".join(repr(float(x)) if isinstance(x, float) else str(x) for x in v)
            return repr(float(v)) if isinstance(v, float) else str(v)

        keys = ["LatticeParameter", "SpaceGroup", "P_Array", "R_Array", "PxX", "PxY",
                "NrPxX", "NrPxY", "Elo", "Ehi", "MaxNrLaueSpots", "MinNrSpots",
                "MinIntensity", "MaxAngle", "DoFwd", "ForwardFile"]
        with open(path, "w") as f:
            for k in keys:
                f.write(f"{k} {fmt(getattr(self, k))}\n")
            for k, v in self.extra.items():
                f.write(f"{k} {fmt(v)}\n")


def baseline_params(**over) -> Params:
    """Cubic FCC Ni-like crystal (a = 0.36 nm), 2048^2 detector at 0.2 mm
    pitch, 500 mm from the sample, centred at 2theta = 90 deg, 5-30 keV."""
    p = Params()
    for k, v in over.items():
        setattr(p, k, v)
    return p


# --------------------------------------------------------------------------
# Orientation DB
# --------------------------------------------------------------------------

def quat_to_matrices(q: np.ndarray) -> np.ndarray:
    """Unit quaternions (w, x, y, z) -> (n, 9) row-major rotation matrices."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    out = np.empty((q.shape[0], 9), dtype=np.float64)
    out[:, 0] = 1 - 2 * (y * y + z * z)
    out[:, 1] = 2 * (x * y - w * z)
    out[:, 2] = 2 * (x * z + w * y)
    out[:, 3] = 2 * (x * y + w * z)
    out[:, 4] = 1 - 2 * (x * x + z * z)
    out[:, 5] = 2 * (y * z - w * x)
    out[:, 6] = 2 * (x * z - w * y)
    out[:, 7] = 2 * (y * z + w * x)
    out[:, 8] = 1 - 2 * (x * x + y * y)
    return out


def _orient_chunk(seed: int, k: int, m: int) -> np.ndarray:
    rng = np.random.default_rng([seed, k])
    q = rng.normal(size=(m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return quat_to_matrices(q)


def random_orientations(n: int, seed: int = 0, out: np.ndarray | None = None,
                        workers: int | None = None) -> np.ndarray:
    """Uniform random rotations: normalised seeded-normal quaternions.

    Chunk k of ``ORIENT_CHUNK`` rows draws from ``default_rng([seed, k])``, so
    any row can be regenerated alone and the DB is the same for any worker
    count.  Returns an (n, 9) float64 array (or fills ``out``)."""
    if out is None:
        out = np.empty((n, 9), dtype=np.float64)
    nchunks = (n + ORIENT_CHUNK - 1) // ORIENT_CHUNK

    def job(k):
        lo = k * ORIENT_CHUNK
        hi = min(n, lo + ORIENT_CHUNK)
        out[lo:hi] = _orient_chunk(seed, k, ORIENT_CHUNK)[: hi - lo]

    workers = workers or min(16, os.cpu_count() or 1)
    if nchunks <= 1 or workers <= 1:
        for k in range(nchunks):
            job(k)
    else:
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(job, range(nchunks)))
    return out


def _qmul(a, b):
    """Hamilton product of quaternion arrays (w, x, y, z)."""
    w1, x1, y1, z1 = np.moveaxis(a, -1, 0)
    w2, x2, y2, z2 = np.moveaxis(b, -1, 0)
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=-1)


def hexagonal_ops() -> np.ndarray:
    """The 12 proper rotations of point group 622 (c along z): six turns about
    z by k*60 deg and six 2-fold axes in the basal plane."""
    ops = []
    for k in range(6):
        a = math.radians(60 * k)
        ops.append([math.cos(a / 2), 0, 0, math.sin(a / 2)])
    for k in range(6):
        a = math.radians(30 * k)
        ops.append([0, math.cos(a), math.sin(a), 0])
    return np.array(ops)


def _orient_chunk_reduced(seed: int, k: int, m: int, ops: np.ndarray) -> np.ndarray:
    rng = np.random.default_rng([seed, k])
    q = rng.normal(size=(m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    cand = _qmul(q[:, None, :], ops[None, :, :])  # q * s for every symmetry s
    best = np.argmax(np.abs(cand[..., 0]), axis=1)
    q = cand[np.arange(m), best]
    q *= np.sign(q[:, :1] + (q[:, :1] == 0))
    return quat_to_matrices(q)


def random_orientations_reduced(n: int, seed: int, ops: np.ndarray,
                                out: np.ndarray | None = None, workers: int | None = None):
    """Uniform random rotations brought into the fundamental zone of the
    given point-group rotations (the equivalent q*s with the largest |w|):
    a symmetry-reduced DB.  Same chunked seeding as random_orientations."""
    if out is None:
        out = np.empty((n, 9), dtype=np.float64)
    nchunks = (n + ORIENT_CHUNK - 1) // ORIENT_CHUNK

    def job(k):
        lo = k * ORIENT_CHUNK
        hi = min(n, lo + ORIENT_CHUNK)
        out[lo:hi] = _orient_chunk_reduced(seed, k, ORIENT_CHUNK, ops)[: hi - lo]

    workers = workers or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max(1, workers)) as ex:
        list(ex.map(job, range(nchunks)))
    return out


def orientation_rows(seed: int, rows, ops: np.ndarray | None = None) -> np.ndarray:
    """Regenerate selected DB rows without building the whole DB (ops: the
    point-group rotations of a symmetry-reduced DB, or None)."""
    rows = np.asarray(rows, dtype=np.int64)
    res = np.empty((rows.size, 9))
    cache = {}
    for i, r in enumerate(rows):
        k = int(r) // ORIENT_CHUNK
        if k not in cache:
            cache[k] = (_orient_chunk(seed, k, ORIENT_CHUNK) if ops is None
                        else _orient_chunk_reduced(seed, k, ORIENT_CHUNK, ops))
        res[i] = cache[k][int(r) % ORIENT_CHUNK]
    return res


# --------------------------------------------------------------------------
# HKL lists
# --------------------------------------------------------------------------

def fcc_hkls(max_index: int = 8, rule: str = "box") -> np.ndarray:
    """FCC-allowed reflections (h, k, l all even or all odd), both signs and
    all harmonics, (0,0,0) excluded.

    ``rule="box"``: |h|,|k|,|l| <= max_index (1240 rows for 8).
    ``rule="sphere"``: h^2+k^2+l^2 <= max_index^2 (536 rows for 8, 1066 for 10).
    Sorted by h^2+k^2+l^2, ties by (h, k, l) descending, so lower-order
    reflections come first (the first hkl of a direction wins the pixel)."""
    r = np.arange(-max_index, max_index + 1)
    h, k, l = np.meshgrid(r, r, r, indexing="ij")
    h, k, l = h.ravel(), k.ravel(), l.ravel()
    par = (h % 2 == k % 2) & (k % 2 == l % 2)
    nz = (h != 0) | (k != 0) | (l != 0)
    m = par & nz
    if rule == "sphere":
        m &= (h * h + k * k + l * l) <= max_index * max_index
    elif rule != "box":
        raise ValueError(rule)
    hkl = np.stack([h[m], k[m], l[m]], axis=1)
    q2 = (hkl ** 2).sum(1)
    order = np.lexsort((-hkl[:, 2], -hkl[:, 1], -hkl[:, 0], q2))
    return hkl[order].astype(np.int32)


def hcp_hkls(max_index: int = 6) -> np.ndarray:
    """Hexagonal close-packed reflections as 3-index hkl from allowed hkil
    (i = -(h+k)), |h|,|k|,|i|,|l| <= max_index; extinction: l odd with
    (h + 2k) % 3 == 0 is absent.  Sorted like :func:`fcc_hkls`."""
    r = np.arange(-max_index, max_index + 1)
    h, k, l = np.meshgrid(r, r, r, indexing="ij")
    h, k, l = h.ravel(), k.ravel(), l.ravel()
    i = -(h + k)
    m = (np.abs(i) <= max_index) & ((h != 0) | (k != 0) | (l != 0))
    m &= ~(((h + 2 * k) % 3 == 0) & (l % 2 == 1))
    hkl = np.stack([h[m], k[m], l[m]], axis=1)
    q2 = (hkl ** 2).sum(1)
    order = np.lexsort((-hkl[:, 2], -hkl[:, 1], -hkl[:, 0], q2))
    return hkl[order].astype(np.int32)


def write_hkl_file(path: str, hkls: np.ndarray) -> None:
    with open(path, "w") as f:
        for h, k, l in np.asarray(hkls):
            f.write(f"{h} {k} {l} {h * h + k * k + l * l}\n")


def read_hkl_file(path: str) -> np.ndarray:
    """First three integers of each line (src/LaueMatchingCPU.c:301-311);
    a line without three integers reads as the zeros the C buffer holds."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            vals = []
            for p in parts[:3]:
                try:
                    vals.append(int(p))
                except ValueError:
                    break
            vals += [0] * (3 - len(vals))
            rows.append(vals)
    return np.asarray(rows, dtype=np.int32).reshape(-1, 3)


# --------------------------------------------------------------------------
# Images
# --------------------------------------------------------------------------

def _rodrigues(r):
    r = np.asarray(r, dtype=np.float64)
    ang = float(np.linalg.norm(r))
    if ang == 0.0:
        return np.eye(3)
    v = r / ang
    c, s = math.cos(ang), math.sin(ang)
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return c * np.eye(3) + (1 - c) * np.outer(v, v) + s * K


def _recip(lat):
    a, b, c = lat[:3]
    al, be, ga = np.radians(lat[3:6])
    ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
    phi = math.sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg)
    A = np.array([a, 0, 0])
    B = np.array([b * cg, b * sg, 0])
    C = np.array([c * cb, c * (ca - cb * cg) / sg, c * phi / sg])
    V = a * b * c * phi
    return np.stack([np.cross(B, C), np.cross(C, A), np.cross(A, B)], axis=1) * (2 * math.pi / V)


def spot_positions(orient: np.ndarray, hkls: np.ndarray, p: Params):
    """Continuous detector positions (px, py) of the reflections of one
    orientation that reach the detector inside the energy window, one per
    direction (first hkl wins).  Used only to draw synthetic images."""
    M = np.asarray(orient, dtype=np.float64).reshape(3, 3) @ _recip(p.LatticeParameter)
    q = hkls.astype(np.float64) @ M.T
    ql = np.linalg.norm(q, axis=1)
    qh = q / ql[:, None]
    kf = -2 * qh[:, 2:3] * qh
    kf[:, 2] += 1.0
    xyz = kf @ _rodrigues(p.R_Array)  # rows times R == (R^T kf)^T
    ok = xyz[:, 2] > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        X = xyz[:, 0] * p.P_Array[2] / xyz[:, 2] - p.P_Array[0]
        Y = xyz[:, 1] * p.P_Array[2] / xyz[:, 2] - p.P_Array[1]
        px = X / p.PxX + 0.5 * (p.NrPxX - 1)
        py = Y / p.PxY + 0.5 * (p.NrPxY - 1)
        E = HC_KEVNM * ql / (4 * math.pi * -qh[:, 2])
    ok &= (px > -1) & (px < p.NrPxX) & (py > -1) & (py < p.NrPxY)
    ok &= (E >= p.Elo) & (E <= p.Ehi)
    px, py = px[ok], py[ok]
    key = px.astype(np.int64) * p.NrPxY + py.astype(np.int64)
    _, first = np.unique(key, return_index=True)
    first.sort()
    return px[first], py[first]


def quantize_to_255(img: np.ndarray) -> np.ndarray:
    """Integers 0..255 with max exactly 255, same rounding rule as the
    indexer's quantizer (positive values below one step become 1)."""
    mx = float(img.max())
    if mx <= 0:
        return np.zeros_like(img, dtype=np.float64)
    v = img * (255.0 / mx)
    out = np.where(v >= 255.0, 255.0, np.where(v < 1.0, 1.0, np.floor(v)))
    out = np.where(img > 0, out, 0.0)
    out[img == mx] = 255.0  # the maximum maps to 255 exactly (imScale = 255/255)
    return out


def synth_image(p: Params, hkls: np.ndarray, grains: np.ndarray, seed: int,
                sigma_spot: float = 1.5, amp=(200.0, 4000.0), background: float = 50.0,
                pct: float = 99.0, sigma_blur: float = 2.0) -> np.ndarray:
    """Float64 H x W image of integers 0..255 (max 255) with the grains'
    Laue spots; ``grains`` is (k, 9)."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    W, H = p.NrPxX, p.NrPxY
    img = rng.poisson(background, size=(H, W)).astype(np.float64)
    rad = int(math.ceil(5 * sigma_spot))
    for g in np.asarray(grains).reshape(-1, 9):
        px, py = spot_positions(g, hkls, p)
        amps = rng.uniform(amp[0], amp[1], size=px.size)
        for cx, cy, a in zip(px, py, amps):
            x0, x1 = max(0, int(cx) - rad), min(W, int(cx) + rad + 1)
            y0, y1 = max(0, int(cy) - rad), min(H, int(cy) + rad + 1)
            if x0 >= x1 or y0 >= y1:
                continue
            xs = np.arange(x0, x1) + 0.5 - cx
            ys = np.arange(y0, y1) + 0.5 - cy
            img[y0:y1, x0:x1] += a * np.exp(-(ys[:, None] ** 2 + xs[None, :] ** 2)
                                            / (2 * sigma_spot ** 2))
    thr = np.percentile(img, pct)
    img = np.where(img > thr, img, 0.0)
    img = gaussian_filter(img, sigma_blur)
    return quantize_to_255(img)


def write_image(path: str, img: np.ndarray) -> None:
    np.ascontiguousarray(img, dtype=np.float64).tofile(path)
