"""GPU parity at the sizes of the benchmark configurations, through the C ABI,
against the CPU oracle on the same seeded inputs.

* The plain CPU configuration at full size: 1M uniform random orientations,
  FCC a = 0.36 nm with every FCC reflection |h|,|k|,|l| <= 8, 5-30 keV, a
  2048 x 2048 detector at 0.2 mm 500 mm from the sample, M = 64, one image of
  3 seeded grains, MinNrSpots 5 and MinIntensity 1000: the whole ForwardFile
  table (FP64 and FP32 with the FP64 guard band) byte for byte, the DoFwd=1
  float64-image scores, and the DoFwd=0 uint8 scores through every kept
  single-image layout.
* Batches of 8, 16, 32 and 64 images with seeded grain counts U{1..20} (spot
  density varying about 20x) over the same 1M-orientation table: every image
  of every batch against the oracle, through the orientation-major table
  walk, the batch banded copy and the autotuned choice.
"""
import sys

import numpy as np
import pytest

import oracle_bind as ob
from lauematching_amd import api, synth

pytestmark = pytest.mark.gpu

N = 1_000_000


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return api.Context(0)


@pytest.fixture(scope="module")
def cfg0(ctx):
    p = synth.baseline_params(MaxNrLaueSpots=64, MinNrSpots=5, MinIntensity=1000.0)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(N, seed=0)
    want, _, _ = ob.forward(O, hk, ob.geom(p), 64, threads=16)
    rng = np.random.default_rng(1)
    grains = np.sort(rng.choice(N, 3, replace=False))
    img = synth.synth_image(p, hk, O[grains], seed=2)
    _say("config0: 1M orientations, oracle table ready")
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), 64)
    yield p, hk, O, want, grains, img, db
    db.close()


def test_config0_forward_table_fp64_and_fp32_guard(cfg0):
    p, hk, O, want, grains, img, db = cfg0
    _say("config0 forward FP64")
    st = db.forward(api.LM_FWD_FP64)
    np.testing.assert_array_equal(db.export_table(), want)
    assert st.n_spots == int(want[:, 0].astype(np.int64).sum())
    c64 = db.checksum()
    _say("config0 forward FP32 + guard")
    st32 = db.forward(api.LM_FWD_FP32_GUARD)
    np.testing.assert_array_equal(db.export_table(), want)
    assert db.checksum() == c64
    _say(f"  recomputed {st32.n_recomputed / st32.n_pairs:.4%} of pairs")
    assert 0 < st32.n_recomputed < st32.n_pairs // 20


def test_config0_dofwd1_float64_image(cfg0):
    p, hk, O, want, grains, img, db = cfg0
    if db.table_spots == 0:
        db.forward()
    _say("config0 DoFwd=1 float64 image")
    _, matched, nres = ob.forward(O, hk, ob.geom(p), 64, img=img, min_spots=5, min_int=1000.0,
                                  threads=16)
    rows, mA = ob.collect(matched)
    h = db.score_f64(img, 5, 1000.0)
    assert h.nr_results == nres
    np.testing.assert_array_equal(h.ids, rows)
    np.testing.assert_array_equal(h.score, mA)
    assert set(grains.tolist()) <= set(h.ids.tolist())


@pytest.mark.parametrize("layout", [(0, 0), (14, 0), (15, 3), (16, 3), (16, 6)])
@pytest.mark.parametrize("thr", [(5, 1000.0), (5, 50.0), (8, 3000.0), (2, 0.5)])
def test_config0_uint8_scores(cfg0, layout, thr):
    p, hk, O, want, grains, img, db = cfg0
    if db.table_spots == 0:
        db.forward()
    tb, loads = layout
    db.set_banded_tile_bits(tb)
    db.set_banded_loads(loads)
    u8, s = api.quantize_u8(img)
    ms, mi = thr
    _say(f"config0 uint8 layout={layout} thresholds={thr}")
    matched, ns, tot, nres = ob.compare(want, 64, u8, p.NrPxX, s, ms, mi, threads=16)
    rows, mA = ob.collect(matched)
    h = db.score(u8, s, ms, mi)
    assert h.nr_results == nres
    np.testing.assert_array_equal(h.ids, rows)
    np.testing.assert_array_equal(h.score, mA)
    np.testing.assert_array_equal(h.nspots, ns[rows.astype(np.int64)])
    np.testing.assert_array_equal(h.totint, tot[rows.astype(np.int64)])
    if thr == (5, 1000.0):
        assert set(grains.tolist()) <= set(h.ids.tolist())


def _batch_images(p, hk, O, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = int(rng.integers(1, 21))
        rows = np.sort(rng.choice(N, k, replace=False))
        img = synth.synth_image(p, hk, O[rows], seed=int(rng.integers(1 << 30)))
        u8, s = api.quantize_u8(img)
        out.append((u8, s, rows))
    return out


@pytest.fixture(scope="module")
def batch_imgs(cfg0):
    p, hk, O, want, grains, img, db = cfg0
    return _batch_images(p, hk, O, 64, seed=77)


@pytest.mark.parametrize("B", [8, 16, 32, 64])
@pytest.mark.parametrize("walk", ["table", "banded", "per_width", "autotuned"])
def test_config2_batches_vs_oracle(cfg0, batch_imgs, B, walk):
    p, hk, O, want, grains, img, db = cfg0
    if db.table_spots == 0:
        db.forward()
    imgs = batch_imgs[:B]
    nz = [np.count_nonzero(u) for u, _, _ in imgs]
    if walk == "table":
        db.set_batch_tile_bits(0)
    elif walk == "banded":
        bits = {8: 12, 16: 11, 32: 10, 64: 9}[B]
        db.set_batch_tile_bits(bits)
        db.set_batch_widths(0x7f)
    elif walk == "per_width":
        # every width its own copy at once (2^12, 2^11, 2^10, 2^9 rows)
        db.set_batch_tile_bits(0)
        for w, bits in ((8, 12), (16, 11), (32, 10), (64, 9)):
            db.set_batch_width_tile_bits(w, bits)
        db.set_batch_widths(0x7f)
    else:
        db.autotune([u for u, _, _ in imgs[:3]], max_batch=64, reps=1)
    _say(f"config2 B={B} walk={walk} (nonzero pixels {min(nz)}..{max(nz)})")
    hs = db.score_batch([u for u, _, _ in imgs], [s for _, s, _ in imgs], 5, 1000.0)
    for i, (u8, s, rows) in enumerate(imgs):
        matched, ns, tot, nres = ob.compare(want, 64, u8, p.NrPxX, s, 5, 1000.0, threads=16)
        r, mA = ob.collect(matched)
        assert hs[i].nr_results == nres, i
        np.testing.assert_array_equal(hs[i].ids, r)
        np.testing.assert_array_equal(hs[i].score, mA)
        np.testing.assert_array_equal(hs[i].nspots, ns[r.astype(np.int64)])
    db.set_batch_tile_bits(0)
