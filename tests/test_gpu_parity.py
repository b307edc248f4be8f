"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Integer tables, counts, ids and sums must be bit-exact;
float64 scores and intensity sums must be bit-identical too (the oracle is the
reference's float64 arithmetic)."""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle_bind as ob
from lauematching_amd import _lib, api, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return api.Context(0)


def _db(ctx, p, O, hk):
    return api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), p.MaxNrLaueSpots)


def test_selftest_fp64_division_and_sqrt(ctx):
    """Device float64 division and sqrt are correctly rounded (bit-equal to
    the host's IEEE results) over random and hard inputs."""
    rng = np.random.default_rng(1)
    n = 1 << 22
    a = np.exp(rng.uniform(-40, 40, n)) * rng.choice([-1.0, 1.0], n)
    b = np.exp(rng.uniform(-40, 40, n)) * rng.choice([-1.0, 1.0], n)
    a[:1000] = np.nextafter(1.0, 2.0) ** np.arange(1000)  # near 1
    b[:1000] = 3.0
    a = np.abs(a) * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    q = np.zeros(n)
    s = np.zeros(n)
    _say("selftest_fp64")
    _lib.check(_lib.lib().lm_selftest_fp64(ctx.handle, a.ctypes.data, b.ctypes.data, n,
                                           q.ctypes.data, s.ctypes.data))
    np.testing.assert_array_equal(q.view(np.uint64), (a / b).view(np.uint64))
    pos = a > 0
    np.testing.assert_array_equal(s[pos].view(np.uint64), np.sqrt(a[pos]).view(np.uint64))
    assert np.isnan(s[~pos]).all()


GEOMS = {
    "baseline": dict(),
    "small_rect": dict(NrPxX=300, NrPxY=200, PxX=0.0008, PxY=0.0011, P_Array=(0.003, -0.002, 0.3)),
    "ni_fixture": dict(LatticeParameter=(0.35238,) * 3 + (90.0,) * 3,
                       P_Array=(0.028828, 0.002715, 0.512993),
                       R_Array=(-1.20161887, -1.21404493, -1.21852276)),
}


@pytest.mark.parametrize("geo", list(GEOMS))
@pytest.mark.parametrize("M", [64, 7, 100])
def test_forward_table_bit_identical(ctx, geo, M):
    p = synth.baseline_params(MaxNrLaueSpots=M, **GEOMS[geo])
    hk = synth.fcc_hkls(8, "box")
    n = 20000 if M != 100 else 3000
    O = synth.random_orientations(n, seed=11)
    _say(f"forward geo={geo} M={M} n={n}")
    db = _db(ctx, p, O, hk)
    st = db.forward()
    got = db.export_table()
    want, _, _ = ob.forward(O, hk, ob.geom(p), M)
    np.testing.assert_array_equal(got, want)
    assert st.n_spots == int(want[:, 0].astype(np.int64).sum()) == db.table_spots
    # partial export of a range that starts mid-tile
    np.testing.assert_array_equal(db.export_table(1234, 999), want[1234:1234 + 999])


@pytest.mark.parametrize("geo", list(GEOMS))
def test_forward_fp32_guard_bit_identical(ctx, geo):
    """FP32 projection with the FP64 guard band: the same table as FP64 and as
    the oracle; a small fraction of pairs is recomputed."""
    p = synth.baseline_params(MaxNrLaueSpots=64, **GEOMS[geo])
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(30000, seed=13)
    _say(f"forward fp32-guard geo={geo}")
    db = _db(ctx, p, O, hk)
    st = db.forward(api.LM_FWD_FP32_GUARD)
    want, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    np.testing.assert_array_equal(db.export_table(), want)
    frac = st.n_recomputed / st.n_pairs
    _say(f"  recomputed {st.n_recomputed} of {st.n_pairs} pairs ({frac:.4%})")
    assert 0 < st.n_recomputed and frac < 0.05
    c32 = db.checksum()
    db.forward(api.LM_FWD_FP64)
    assert db.checksum() == c32


# Rows of the 100M-orientation DB (seed 0, 2048^2 detector, FCC |hkl| <= 8)
# with a spot whose energy lies within a float32 rounding of Ehi + eE: an
# earlier classifier compared E with the rounded sum Ehi + eE and accepted
# them (E = 30.00045..30.00051 keV > Ehi = 30).
FP32_EDGE_ROWS = [13621346, 18944713, 21642751, 29386159, 30040475, 38000683, 40986199,
                  51687398, 56880516, 59475135, 61372750, 62127485, 66307574, 72056249,
                  75014891, 76294540, 78089371, 78238134, 79530517, 83598450, 84554566,
                  84604834, 86705729, 87291174, 89099544, 95102346, 97160938, 98347393,
                  98472412, 99041057, 99827004]


def test_forward_fp32_guard_energy_edge_rows(ctx):
    """Regression: pairs just above the energy window stay rejected."""
    p = synth.baseline_params(MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(8, "box")
    O = synth.orientation_rows(0, FP32_EDGE_ROWS)
    _say("forward fp32-guard energy-edge rows")
    db = _db(ctx, p, O, hk)
    db.forward(api.LM_FWD_FP32_GUARD)
    want, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    np.testing.assert_array_equal(db.export_table(), want)


def test_forward_fp32_guard_equals_fp64_4m(ctx):
    """4M random orientations: FP32+guard and FP64 tables have equal
    position-tagged checksums (the FP64 path is bit-exact to the oracle)."""
    p = synth.baseline_params(MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(4_000_000, seed=77)
    _say("forward fp32-guard vs fp64, 4M orientations")
    db = _db(ctx, p, O, hk)
    db.set_banded_tile_bits(0)
    db.forward(api.LM_FWD_FP64)
    c64 = db.checksum()
    db.forward(api.LM_FWD_FP32_GUARD)
    assert db.checksum() == c64
    db.close()


def test_forward_fp32_guard_config4_hkls(ctx):
    """400-reflection-class list (FCC |hkl| <= 10, 1066 rows), 5-40 keV."""
    p = synth.baseline_params(Ehi=40.0, MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(10, "sphere")
    O = synth.random_orientations(30000, seed=14)
    _say("forward fp32-guard |hkl|<=10, 5-40 keV")
    db = _db(ctx, p, O, hk)
    db.forward(api.LM_FWD_FP32_GUARD)
    want, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    np.testing.assert_array_equal(db.export_table(), want)


def test_checksum_detects_a_changed_spot(ctx):
    p = synth.baseline_params(NrPxX=256, NrPxY=256, PxX=0.0016, PxY=0.0016, MaxNrLaueSpots=16)
    O = synth.random_orientations(5000, seed=15)
    hk = synth.fcc_hkls(6)
    db = _db(ctx, p, O, hk)
    db.forward()
    rec = db.export_table()
    c0 = db.checksum()
    i = int(np.nonzero(rec[:, 0] > 1)[0][0])
    rec2 = rec.copy()
    rec2[i, 1] = (int(rec2[i, 1]) + 1) % 256
    db.import_table(rec2)
    assert db.checksum() != c0
    rec3 = rec.copy()  # swap two spots of one orientation: same set, other order
    rec3[i, 1:3], rec3[i, 3:5] = rec[i, 3:5].copy(), rec[i, 1:3].copy()
    db.import_table(rec3)
    assert db.checksum() != c0
    db.import_table(rec)
    assert db.checksum() == c0


def test_forward_hexagonal(ctx):
    p = synth.baseline_params(LatticeParameter=(0.295, 0.295, 0.468165, 90, 90, 120),
                              SpaceGroup=194, MaxNrLaueSpots=64)
    hk = synth.hcp_hkls(6)
    O = synth.random_orientations(8000, seed=12)
    _say("forward hexagonal")
    db = _db(ctx, p, O, hk)
    db.forward()
    want, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    np.testing.assert_array_equal(db.export_table(), want)


@pytest.fixture(scope="module")
def scored_case(ctx):
    p = synth.baseline_params(MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(8, "box")
    n = 60000
    O = synth.random_orientations(n, seed=21)
    grains = [123, 40000, 59999]
    img = synth.synth_image(p, hk, O[grains], seed=5)
    table, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    db = _db(ctx, p, O, hk)
    db.forward()
    return p, hk, O, grains, img, table, db


@pytest.mark.parametrize("min_sp,min_int", [(5, 1000.0), (5, 300.0), (8, 0.0), (0, 0.0),
                                             (3, 2500.0)])
def test_score_matches_oracle_unit_scale(scored_case, min_sp, min_int):
    p, hk, O, grains, img, table, db = scored_case
    u8, s = api.quantize_u8(img)
    assert s == 1.0
    _say(f"score unit scale min_sp={min_sp} min_int={min_int}")
    hits = db.score(u8, s, min_sp, min_int)
    matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, min_sp, min_int)
    rows, mA = ob.collect(matched)
    assert hits.nr_results == nres
    np.testing.assert_array_equal(hits.ids, rows)
    np.testing.assert_array_equal(hits.score, mA)
    np.testing.assert_array_equal(hits.nspots, ns[rows.astype(np.int64)])
    np.testing.assert_array_equal(hits.totint, tot[rows.astype(np.int64)])
    if min_int <= 1000:
        for g in grains:
            assert g in set(hits.ids.tolist())


@pytest.mark.parametrize("bits,loads", [(0, 0), (13, 0), (14, 0), (15, 0), (13, 1), (14, 1),
                                        (15, 1), (13, 2), (14, 2), (15, 2), (14, 3), (15, 3),
                                        (13, 4), (14, 4), (15, 4), (14, 5), (15, 5), (14, 6),
                                        (15, 6), (14, 7), (15, 7)])
def test_score_layouts_agree(scored_case, bits, loads):
    """Banded copy with 8K, 16K or 32K super-tiles walked with 4-byte or 16-byte
    entry loads, and no banded copy (the orientation-major kernel): identical
    results to the oracle."""
    p, hk, O, grains, img, table, db = scored_case
    u8, s = api.quantize_u8(img)
    _say(f"score banded bits={bits} loads={loads}")
    db.set_banded_tile_bits(bits)
    db.set_banded_loads(loads)
    try:
        for min_sp, min_int in [(5, 250.0), (0, 0.0)]:
            hits = db.score(u8, s, min_sp, min_int)
            matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, min_sp, min_int)
            rows, mA = ob.collect(matched)
            assert hits.nr_results == nres
            np.testing.assert_array_equal(hits.ids, rows)
            np.testing.assert_array_equal(hits.score, mA)
            np.testing.assert_array_equal(hits.nspots, ns[rows.astype(np.int64)])
    finally:
        db.set_banded_tile_bits(14)
        db.set_banded_loads(0)


@pytest.mark.parametrize("loads", [3, 5, 6, 7, 0])
def test_score_sum16_walk(scored_case, loads):
    """2^16-row super-tiles with 16-bit row sums: rows whose sum reaches
    MinIntensity are recounted from the table; thresholds with more such rows
    than the candidate list holds (4096 here), MinIntensity <= 0 and an
    unaligned device image take the table walk.  All bit-exact."""
    import torch

    p, hk, O, grains, img, table, db = scored_case
    u8, s = api.quantize_u8(img)
    _say(f"score sum16 loads={loads}")
    db.set_banded_tile_bits(16)
    db.set_banded_loads(loads)
    assert db.banded_layout()[0] == 16
    recounted = overflowed = 0
    try:
        for min_sp, min_int in [(5, 1000.0), (3, 2500.0), (0, 600.0), (5, 250.0), (8, 0.0),
                                (0, 0.0), (1, 1.0), (300, 900.0), (2, 1e9)]:
            hits = db.score(u8, s, min_sp, min_int)
            matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, min_sp, min_int)
            rows, mA = ob.collect(matched)
            assert hits.nr_results == nres, (min_sp, min_int)
            np.testing.assert_array_equal(hits.ids, rows)
            np.testing.assert_array_equal(hits.score, mA)
            np.testing.assert_array_equal(hits.nspots, ns[rows.astype(np.int64)])
            np.testing.assert_array_equal(hits.totint, tot[rows.astype(np.int64)])
            if min_int > 0:
                ncand = int(np.count_nonzero((tot >= min_int) & (tot >= max(min_sp, 1))))
                recounted += 0 < ncand <= 4096
                overflowed += ncand > 4096
        assert recounted >= 2 and overflowed >= 1
        # a device image at an odd address takes the table walk
        buf = torch.zeros(u8.size + 16, dtype=torch.uint8, device="cuda:0")
        buf[1:1 + u8.size] = torch.from_numpy(u8).to("cuda:0")
        torch.cuda.synchronize()
        hits = db.score_dev(buf.data_ptr() + 1, 1.0, 5, 1000.0)
        matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, 1.0, 5, 1000.0)
        rows, mA = ob.collect(matched)
        np.testing.assert_array_equal(hits.ids, rows)
        np.testing.assert_array_equal(hits.score, mA)
    finally:
        db.set_banded_tile_bits(14)
        db.set_banded_loads(0)


@pytest.mark.parametrize("loads", [3, 6])
def test_score_count16_walk_sparse_images(scored_case, loads):
    """Sparse images (a few percent of pixels nonzero, as preprocessed frames
    are) with MinNrSpots >= 2: the 2^16-row walk counts nonzero pixels per
    row instead of summing them, candidates are the rows with enough spots,
    and the recount gives the exact sums.  Thresholds as the reference's
    parameter files set them (MinIntensity 50) and others; imScale 1 and not."""
    p, hk, O, grains, img, table, db = scored_case
    sparse = np.where(img > np.percentile(img, 97.0), img, 0.0)
    cases = [api.quantize_u8(sparse), api.quantize_u8(sparse * 3.7)]
    db.set_banded_tile_bits(16)
    db.set_banded_loads(loads)
    try:
        for u8, s in cases:
            assert np.count_nonzero(u8) < u8.size // 4
            for min_sp, min_int in [(5, 50.0 * s), (2, 0.5), (3, 2000.0 * s), (7, 1e-3)]:
                _say(f"count16 loads={loads} s={s:.3f} min_sp={min_sp} min_int={min_int}")
                hits = db.score(u8, s, min_sp, min_int)
                matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, min_sp, min_int)
                rows, mA = ob.collect(matched)
                assert hits.nr_results == nres, (min_sp, min_int)
                np.testing.assert_array_equal(hits.ids, rows)
                np.testing.assert_array_equal(hits.score, mA)
                np.testing.assert_array_equal(hits.nspots, ns[rows.astype(np.int64)])
                np.testing.assert_array_equal(hits.totint, tot[rows.astype(np.int64)])
    finally:
        db.set_banded_tile_bits(14)
        db.set_banded_loads(0)


@pytest.mark.parametrize("bits", [14, 15, 16])
def test_lds_walk_skips_zero_chunks(scored_case, bits):
    """The LDS-staged walks skip the entries of image chunks (2^11 or 2^12
    pixels) with no nonzero pixel.  Images whose nonzero pixels sit in a few
    bands of detector rows, lone nonzero pixels at the first and last pixel
    of a chunk, and a blank image: every row still equals the oracle, with
    sums and with spot counts as the 16-bit counters."""
    p, hk, O, grains, img, table, db = scored_case
    base, _ = api.quantize_u8(img)
    H, W = p.NrPxY, p.NrPxX
    banded = base.reshape(H, W).copy()
    for y0 in range(0, H, 300):
        banded[y0 + 40:y0 + 300] = 0  # keep 40 rows of every 300
    edge = np.zeros(H * W, np.uint8)
    cb = 12 if bits == 14 else 11
    nch = (H * W) >> cb
    for c in (3, nch // 2, nch - 1):
        edge[(c << cb)] = 200
        edge[((c + 1) << cb) - 1] = 255
    hot = base.ravel()[np.flatnonzero(base.ravel() > 200)[:5000]]
    edge[np.flatnonzero(base.ravel() > 200)[:5000]] = hot  # the brightest spots of the grains
    images = [banded.ravel(), edge, np.zeros(H * W, np.uint8)]
    db.set_banded_tile_bits(bits)
    db.set_banded_loads(3)
    try:
        for k, u8 in enumerate(images):
            for min_sp, min_int in [(5, 300.0), (1, 1.0), (3, 50.0)]:
                _say(f"zero chunks bits={bits} image={k} min_sp={min_sp} min_int={min_int}")
                hits = db.score(u8, 1.0, min_sp, min_int)
                matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, 1.0, min_sp, min_int)
                rows, mA = ob.collect(matched)
                assert hits.nr_results == nres, (k, min_sp, min_int)
                np.testing.assert_array_equal(hits.ids, rows)
                np.testing.assert_array_equal(hits.score, mA)
                np.testing.assert_array_equal(hits.nspots, ns[rows.astype(np.int64)])
    finally:
        db.set_banded_tile_bits(14)
        db.set_banded_loads(0)


def test_sum16_needs_small_rows(ctx):
    """16-bit sums hold 255 * 257 at most: a DB with MaxNrLaueSpots > 257
    refuses 2^16 super-tiles."""
    p = synth.baseline_params(MaxNrLaueSpots=300)
    hk = synth.fcc_hkls(4, "box")
    O = synth.random_orientations(2000, seed=3)
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), 300)
    db.forward()
    with pytest.raises(_lib.LaueMatchError):
        db.set_banded_tile_bits(16)
    db.set_banded_tile_bits(15)
    db.close()


@pytest.mark.parametrize("geo", ["small_rect", "det1024"])
def test_score_other_detectors(ctx, geo):
    over = dict(GEOMS.get(geo, {}))
    if geo == "det1024":
        over = dict(NrPxX=1024, NrPxY=1024, PxX=0.0004, PxY=0.0004)
    p = synth.baseline_params(MaxNrLaueSpots=64, **over)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(40000, seed=31)
    img = synth.synth_image(p, hk, O[[3, 33333]], seed=8)
    _say(f"score detector {geo}")
    db = _db(ctx, p, O, hk)
    db.forward()
    table = db.export_table()
    u8, s = api.quantize_u8(img)
    assert s == 1.0
    for bits in (14, 13, 0):
        db.set_banded_tile_bits(bits)
        hits = db.score(u8, s, 4, 150.0)
        matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, 4, 150.0)
        rows, mA = ob.collect(matched)
        assert hits.nr_results == nres
        np.testing.assert_array_equal(hits.ids, rows)
        np.testing.assert_array_equal(hits.score, mA)


def test_chunk_pieces_match_oracle(ctx):
    """The LDS walk's work items are pieces of one image chunk's entries
    (LM_OPT_BANDED_SPLIT): a 1024 x 1024 detector (four times the entries per
    chunk of a 2048 x 2048 one) over 2^14..2^16-row super-tiles (the last one
    partial), both LDS walks, every split and the default, on a dense frame
    and on a frame whose rows are mostly zero (whole chunks skipped): every
    result bit-exact against the oracle."""
    p = synth.baseline_params(MaxNrLaueSpots=64, NrPxX=1024, NrPxY=1024, PxX=0.0004, PxY=0.0004)
    hk = synth.fcc_hkls(8, "box")
    n = 150000
    O = synth.random_orientations(n, seed=41)
    img = synth.synth_image(p, hk, O[[7, 77777, 149999]], seed=9)
    u8, s = api.quantize_u8(img)
    sparse = u8.copy().reshape(1024, 1024)
    sparse[np.arange(1024) % 5 != 0] = 0
    sparse = sparse.ravel()
    db = _db(ctx, p, O, hk)
    db.forward()
    table = db.export_table()
    want = {}
    for name, im, ms, mi in (("dense", u8, 5, 1000.0), ("sparse", sparse, 2, 60.0)):
        matched, ns, tot, nres = ob.compare(table, 64, im, p.NrPxX, s, ms, mi)
        want[name] = (nres,) + ob.collect(matched)
    try:
        for bits in (16, 15, 14):
            db.set_banded_tile_bits(bits)
            for loads in (3, 5):
                db.set_banded_loads(loads)
                for split in (-1, 0, 1, 2, 3):
                    db.set_banded_split(split)
                    if split >= 0:
                        assert db.get_option(12) == split
                    for name, im, ms, mi in (("dense", u8, 5, 1000.0), ("sparse", sparse, 2, 60.0)):
                        _say(f"pieces tb={bits} loads={loads} split={split} {name}")
                        hits = db.score(im, s, ms, mi)
                        nres, rows, mA = want[name]
                        assert hits.nr_results == nres, (bits, loads, split, name)
                        np.testing.assert_array_equal(hits.ids, rows)
                        np.testing.assert_array_equal(hits.score, mA)
    finally:
        db.close()


@pytest.mark.parametrize("keep", [0.002, 0.05, 0.2, 0.3])
def test_sum16_sums_or_counts_match_oracle(scored_case, keep):
    """The 16-bit walk counts nonzero pixels instead of summing them only
    where a row expects at most MinNrSpots / 4 of them: frames with 0.2% to
    30% of their pixels kept (the rest zeroed) at MinNrSpots 2..6, bit-exact
    against the oracle either way."""
    p, hk, O, grains, img, table, db = scored_case
    u8, s = api.quantize_u8(img)
    rng = np.random.default_rng(int(keep * 1000))
    frame = np.where(rng.random(u8.size) < keep, u8, 0).astype(np.uint8)
    was = db.banded_layout()
    db.set_banded_tile_bits(16)
    db.set_banded_loads(3)
    try:
        for ms, mi in ((2, 20.0), (4, 100.0), (6, 300.0)):
            _say(f"sums or counts keep={keep} min_sp={ms}")
            hits = db.score(frame, s, ms, mi)
            matched, ns, tot, nres = ob.compare(table, 64, frame, p.NrPxX, s, ms, mi)
            rows, mA = ob.collect(matched)
            assert hits.nr_results == nres, (keep, ms)
            np.testing.assert_array_equal(hits.ids, rows)
            np.testing.assert_array_equal(hits.score, mA)
    finally:
        db.set_banded_tile_bits(was[0])
        db.set_banded_loads(was[1])


def test_context_released_before_its_db():
    """A DB keeps its context: releasing the context first (as a garbage
    collector may) leaves the DB scoring exactly, and destroying the DB
    afterwards releases both."""
    p = synth.baseline_params(MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(6, "box")
    O = synth.random_orientations(5000, seed=51)
    img = synth.synth_image(p, hk, O[[10, 4000]], seed=3)
    u8, s = api.quantize_u8(img)
    table, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    c2 = api.Context(0)
    db = _db(c2, p, O, hk)
    db.forward()
    c2.close()
    _say("score after the context's release")
    hits = db.score(u8, s, 4, 200.0)
    matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, 4, 200.0)
    rows, mA = ob.collect(matched)
    assert hits.nr_results == nres
    np.testing.assert_array_equal(hits.ids, rows)
    np.testing.assert_array_equal(hits.score, mA)
    db.close()


@pytest.mark.parametrize("factor", [0.37, 3.3])
def test_score_matches_oracle_nonunit_scale(scored_case, factor):
    p, hk, O, grains, img, table, db = scored_case
    u8, s = api.quantize_u8(img * factor + (img > 0) * 0.01)
    assert s != 1.0
    for min_sp, min_int in [(5, 1000.0 * factor), (4, 200.0 * factor)]:
        _say(f"score scale={s} min_int={min_int}")
        hits = db.score(u8, s, min_sp, min_int)
        matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, min_sp, min_int)
        rows, mA = ob.collect(matched)
        assert hits.nr_results == nres
        np.testing.assert_array_equal(hits.ids, rows)
        np.testing.assert_array_equal(hits.totint, tot[rows.astype(np.int64)])
        np.testing.assert_array_equal(hits.score, mA)


@pytest.mark.parametrize("bits,loads", [(0, 0), (14, 0), (15, 3), (16, 3), (16, 5)])
def test_nonunit_scale_every_layout(scored_case, bits, loads):
    """Single images with imScale != 1 (every image whose maximum is not 255)
    through every layout: the integer walk keeps a candidate superset, the
    float64 in-order re-sum decides.  MinIntensity set exactly to one row's
    in-order totInt checks the boundary (that row passes, rows a few ulps
    below do not)."""
    p, hk, O, grains, img, table, db = scored_case
    u8, s = api.quantize_u8(img * 0.731 + (img > 0) * 0.013)
    assert s != 1.0
    _, _, tot0, _ = ob.compare(table, 64, u8, p.NrPxX, s, 0, -1.0)
    order = np.argsort(tot0)
    edge = float(tot0[order[int(len(order) * 0.9995)]])  # a row near the top
    db.set_banded_tile_bits(bits)
    db.set_banded_loads(loads)
    try:
        for min_sp, min_int in [(5, 700.0 * s), (3, edge), (0, 0.0), (4, np.nextafter(edge, 0))]:
            _say(f"nonunit tb={bits} loads={loads} min_sp={min_sp} min_int={min_int!r}")
            hits = db.score(u8, s, min_sp, min_int)
            matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, s, min_sp, min_int)
            rows, mA = ob.collect(matched)
            assert hits.nr_results == nres, (min_sp, min_int)
            np.testing.assert_array_equal(hits.ids, rows)
            np.testing.assert_array_equal(hits.totint, tot[rows.astype(np.int64)])
            np.testing.assert_array_equal(hits.score, mA)
    finally:
        db.set_banded_tile_bits(14)
        db.set_banded_loads(0)


@pytest.mark.parametrize("min_sp,min_int", [(5, 300.0), (0, 0.0), (6, 4000.0)])
def test_score_f64_matches_dofwd1_oracle(scored_case, min_sp, min_int):
    """Float64 image (non-integer values): the DoFwd=1 scoring of the oracle's
    forward pass (src/LaueMatchingCPU.c:647-653, 683-689)."""
    p, hk, O, grains, img, table, db = scored_case
    rng = np.random.default_rng(9)
    im = img * rng.uniform(0.3, 3.0, img.shape)
    im[rng.random(img.shape) < 0.05] = -0.5
    _say(f"score f64 min_sp={min_sp} min_int={min_int}")
    hits = db.score_f64(im, min_sp, min_int)
    _, matched, nres = ob.forward(O, hk, ob.geom(p), 64, img=im.ravel(), min_spots=min_sp,
                                  min_int=min_int)
    rows, mA = ob.collect(matched)
    assert hits.nr_results == nres
    np.testing.assert_array_equal(hits.ids, rows)
    np.testing.assert_array_equal(hits.score, mA)


def test_import_matches_forward_and_roundtrips(ctx, scored_case):
    p, hk, O, grains, img, table, db = scored_case
    _say("import/export roundtrip")
    db2 = _db(ctx, p, O, hk)
    db2.import_table(table)
    np.testing.assert_array_equal(db2.export_table(), table)
    u8, s = api.quantize_u8(img)
    a = db.score(u8, s, 5, 500.0)
    b = db2.score(u8, s, 5, 500.0)
    np.testing.assert_array_equal(a.ids, b.ids)
    np.testing.assert_array_equal(a.score, b.score)


def test_import_rejects_bad_records(ctx):
    p = synth.baseline_params(NrPxX=64, NrPxY=32, MaxNrLaueSpots=4)
    O = synth.random_orientations(10, seed=1)
    db = _db(ctx, p, O, synth.fcc_hkls(4))
    rec = np.zeros((10, 9), np.uint16)
    rec[3, 0] = 5  # count > M
    with pytest.raises(_lib.LaueMatchError):
        db.import_table(rec)
    rec[3, 0] = 1
    rec[3, 1], rec[3, 2] = 10, 40  # y outside the detector
    with pytest.raises(_lib.LaueMatchError):
        db.import_table(rec)
    rec[3, 2] = 31
    db.import_table(rec)
    np.testing.assert_array_equal(db.export_table(), rec)


@pytest.mark.parametrize("nimg", [2, 5, 8, 16, 64, 100, 128])
def test_batch_equals_single(scored_case, nimg):
    p, hk, O, grains, img, table, db = scored_case
    rng = np.random.default_rng(nimg)
    imgs, scales = [], []
    base, _ = api.quantize_u8(img)
    for i in range(nimg):
        perm = rng.permutation(base.size) if i % 3 == 1 else None
        u8 = base if perm is None else base[perm]
        if i % 4 == 3:
            u8, s = api.quantize_u8(img * 0.5 + 0.01 * (img > 0))
        else:
            s = 1.0
        imgs.append(u8)
        scales.append(s)
    _say(f"batch nimg={nimg}")
    got = db.score_batch(imgs, scales, 5, 600.0)
    for i in (0, nimg // 2, nimg - 1):
        single = db.score(imgs[i], scales[i], 5, 600.0)
        assert got[i].nr_results == single.nr_results
        np.testing.assert_array_equal(got[i].ids, single.ids)
        np.testing.assert_array_equal(got[i].score, single.score)
        np.testing.assert_array_equal(got[i].nspots, single.nspots)


@pytest.mark.parametrize("nimg,grouped", [(128, 0xC0), (100, 0xC0), (64, 0xC0), (64, 0),
                                          (128, 0), (16, 0xC0)])
def test_batch_thresholds_per_image(scored_case, nimg, grouped):
    """Each image of a batch with its own thresholds (as the server mixes
    requests): rows with no spot passing, thresholds no row reaches, negative
    intensities, images with imScale != 1 and blank images, through the
    grouped (16 images per lane) and the one-lane table walks; every image
    bit-exact against the oracle with its own thresholds."""
    p, hk, O, grains, img, table, db = scored_case
    rng = np.random.default_rng(500 + nimg)
    base, _ = api.quantize_u8(img)
    pairs = [(5, 600.0), (0, 0.0), (0, 300.0), (1, 0.0), (300, 10.0), (5, 1e12), (3, -5.0),
             (8, 1500.0), (2, 2500.5), (256, 0.0), (-3, 0.0)]
    imgs, scales, msp, mint = [], [], [], []
    for i in range(nimg):
        if i % 6 == 5:
            u8, s = api.quantize_u8(img * 0.43 + 0.02 * (img > 0))
        elif i % 11 == 7:
            u8, s = np.zeros_like(base), 1.0
        else:
            u8, s = (base if i == 0 else base[rng.permutation(base.size)]), 1.0
        ms, mi = pairs[i % len(pairs)]
        imgs.append(u8)
        scales.append(s)
        msp.append(ms)
        mint.append(mi * s if mi > 0 else mi)
    _say(f"batch per-image thresholds nimg={nimg} grouped={grouped:#x}")
    db.set_batch_grouped(grouped)
    try:
        got = db.score_batch_each(imgs, scales, msp, mint)
    finally:
        db.set_batch_grouped(0xC0)
    for i in range(nimg):
        matched, ns, tot, nres = ob.compare(table, 64, imgs[i], p.NrPxX, scales[i], msp[i],
                                            mint[i])
        rows, mA = ob.collect(matched)
        assert got[i].nr_results == nres, (i, msp[i], mint[i])
        np.testing.assert_array_equal(got[i].ids, rows)
        np.testing.assert_array_equal(got[i].score, mA)
        np.testing.assert_array_equal(got[i].nspots, ns[rows.astype(np.int64)])


@pytest.mark.parametrize("tb,nimg", [(9, 64), (9, 20), (10, 32), (10, 12), (11, 16), (11, 5),
                                     (12, 8), (13, 4), (13, 3), (0, 16)])
def test_batch_banded_copy_matches_oracle(scored_case, tb, nimg):
    """Batches scored from the batch banded copy (super-tiles of 2^tb rows,
    per-(image, row) LDS counters), and from the orientation-major table
    (tb=0): every image bit-exact against the oracle."""
    p, hk, O, grains, img, table, db = scored_case
    rng = np.random.default_rng(100 + nimg)
    base, _ = api.quantize_u8(img)
    imgs, scales = [], []
    for i in range(nimg):
        if i % 5 == 4:
            u8, s = api.quantize_u8(img * 0.37 + 0.02 * (img > 0))
        elif i % 5 == 2:
            u8, s = np.zeros_like(base), 1.0
        else:
            u8, s = (base if i == 0 else base[rng.permutation(base.size)]), 1.0
        imgs.append(u8)
        scales.append(s)
    _say(f"batch banded tb={tb} nimg={nimg}")
    db.set_batch_tile_bits(tb)
    try:
        got = db.score_batch(imgs, scales, 5, 500.0)
    finally:
        db.set_batch_tile_bits(0)
    for i in range(nimg):
        matched, ns, tot, nres = ob.compare(table, 64, imgs[i], p.NrPxX, scales[i], 5, 500.0)
        rows, mA = ob.collect(matched)
        assert got[i].nr_results == nres, i
        np.testing.assert_array_equal(got[i].ids, rows)
        np.testing.assert_array_equal(got[i].score, mA)
        np.testing.assert_array_equal(got[i].nspots, ns[rows.astype(np.int64)])


@pytest.mark.parametrize("tb,nimg,low", [(12, 16, False), (13, 8, False), (11, 32, False),
                                         (10, 64, False), (9, 64, False), (13, 5, False),
                                         (12, 9, False), (12, 16, True)])
def test_batch_sum16_copy_matches_oracle(scored_case, tb, nimg, low):
    """Batches read from the batch copy with 16-bit sums (two images per LDS
    word, twice the rows) and an exact recount of the candidates: every image
    bit-exact against the oracle, with per-image thresholds, imScale != 1,
    blank images, a MinIntensity <= 0 call (read from the table instead) and,
    with low thresholds, more candidates than the list holds (rerun over the
    table)."""
    p, hk, O, grains, img, table, db = scored_case
    rng = np.random.default_rng(300 + nimg + tb)
    base, _ = api.quantize_u8(img)
    imgs, scales, msp, mint = [], [], [], []
    for i in range(nimg):
        if i % 5 == 4:
            u8, s = api.quantize_u8(img * 0.37 + 0.02 * (img > 0))
        elif i % 5 == 2:
            u8, s = np.zeros_like(base), 1.0
        else:
            u8, s = (base if i == 0 else base[rng.permutation(base.size)]), 1.0
        imgs.append(u8)
        scales.append(s)
        msp.append(1 if low else int(rng.integers(3, 7)))
        mint.append(1.0 if low else float(rng.uniform(300, 900)) * s)
    k = (nimg - 1).bit_length() if nimg > 4 else 2
    _say(f"batch sum16 tb={tb} nimg={nimg} low={low}")
    db.set_batch_sum16(1 << k)
    db.set_batch_width_tile_bits(1 << k, tb)
    db.set_batch_widths(1 << k)
    try:
        got = db.score_batch_each(imgs, scales, msp, mint)
        mint0 = list(mint)
        mint0[1] = -1.0  # one image with MinIntensity <= 0: the table walk
        got0 = db.score_batch_each(imgs, scales, msp, mint0)
    finally:
        db.set_batch_sum16(0)
        db.set_batch_width_tile_bits(1 << k, 0)
        db.set_batch_widths(0x7F)
    for gg, mi in ((got, mint), (got0, mint0)):
        for i in range(nimg):
            matched, ns, tot, nres = ob.compare(table, 64, imgs[i], p.NrPxX, scales[i], msp[i],
                                                mi[i])
            rows, mA = ob.collect(matched)
            assert gg[i].nr_results == nres, (i, msp[i], mi[i])
            np.testing.assert_array_equal(gg[i].ids, rows)
            np.testing.assert_array_equal(gg[i].score, mA)
            np.testing.assert_array_equal(gg[i].nspots, ns[rows.astype(np.int64)])


@pytest.mark.parametrize("nimg", [2, 4, 13, 40])
def test_batch_split_matches_oracle(scored_case, nimg):
    """A width scored as two calls of half the width (LM_OPT_BATCH_SPLIT), and
    2..4 images scored one at a time: every image bit-exact against the
    oracle, per-image thresholds kept."""
    p, hk, O, grains, img, table, db = scored_case
    rng = np.random.default_rng(500 + nimg)
    base, _ = api.quantize_u8(img)
    imgs = [base] + [base[rng.permutation(base.size)] for _ in range(nimg - 1)]
    msp = [int(v) for v in rng.integers(3, 7, nimg)]
    mint = [float(v) for v in rng.uniform(300, 900, nimg)]
    _say(f"batch split nimg={nimg}")
    db.set_batch_split(0x7C)
    try:
        got = db.score_batch_each(imgs, [1.0] * nimg, msp, mint)
    finally:
        db.set_batch_split(0)
    for i in range(nimg):
        matched, ns, tot, nres = ob.compare(table, 64, imgs[i], p.NrPxX, 1.0, msp[i], mint[i])
        rows, mA = ob.collect(matched)
        assert got[i].nr_results == nres, i
        np.testing.assert_array_equal(got[i].ids, rows)
        np.testing.assert_array_equal(got[i].score, mA)


def test_autotuner_picks_timed_layouts_and_stays_exact(scored_case):
    """The autotuner times every layout, keeps the fastest and leaves the DB
    scoring bit-exactly (single image and batches) with what it chose."""
    p, hk, O, grains, img, table, db = scored_case
    base, _ = api.quantize_u8(img)
    rng = np.random.default_rng(9)
    samples = [base, base[rng.permutation(base.size)]]
    _say("autotune")
    rep = db.autotune(samples, max_batch=32, reps=2)
    try:
        assert set(rep["ms_batch"]) == {4, 8, 16, 32}  # max_batch=32
        singles = rep["ms_single"]
        assert len(singles) == 25 and all(v > 0 for v in singles.values())
        tb = rep["banded_tile_bits"]
        chosen = "table" if tb == 0 else f"banded{tb}_loads{rep['banded_loads']}"
        # the kept layout is within 1% of the fastest (a later one must win by more)
        assert singles[chosen] <= min(singles.values()) / 0.99
        # an LDS walk is kept with one of its timed splits of the chunks into pieces
        split = rep["banded_split"]
        if tb >= 14 and rep["banded_loads"] in (3, 5, 6, 7):
            assert sorted(rep["ms_split"]) == [1, 2, 4, 8] and 0 <= split <= 3
            assert rep["ms_split"][1 << split] <= min(rep["ms_split"].values()) / 0.99
            assert db.get_option(12) == split
        else:
            assert split == -1
        u8 = samples[0]
        one = db.score(u8, 1.0, 5, 500.0)
        matched, ns, tot, nres = ob.compare(table, 64, u8, p.NrPxX, 1.0, 5, 500.0)
        rows, mA = ob.collect(matched)
        np.testing.assert_array_equal(one.ids, rows)
        np.testing.assert_array_equal(one.score, mA)
        for B, row in rep["ms_batch"].items():
            assert row["table"] > 0
            use = (rep["batch_widths_mask"] >> (B.bit_length() - 1)) & 1
            wtb = rep["batch_tile_bits_per_width"][B]
            assert bool(use) == bool(wtb)
            if use:  # the width's own copy, the fastest of its copies, beats the table
                s16 = (rep["sum16_mask"] >> (B.bit_length() - 1)) & 1
                key = f"banded{wtb}" + ("_sum16" if s16 else "")
                assert row[key] < row["table"]
                assert row[key] == min(v for k, v in row.items() if k.startswith("banded"))
        got = db.score_batch(samples * 8, [1.0] * 16, 5, 500.0)
        for i in (0, 1):
            matched, ns, tot, nres = ob.compare(table, 64, samples[i], p.NrPxX, 1.0, 5, 500.0)
            rows, mA = ob.collect(matched)
            for g in got[i::2]:
                assert g.nr_results == nres
                np.testing.assert_array_equal(g.ids, rows)
                np.testing.assert_array_equal(g.score, mA)
    finally:
        db.set_batch_sum16(0)
        db.set_batch_split(0)
        db.set_batch_tile_bits(0)
        db.set_batch_widths(0x7F)
        db.set_banded_tile_bits(14)
        db.set_banded_loads(0)


def test_graph_replay_of_direct_calls(scored_case):
    """Direct calls with LM_OPT_USE_GRAPHS replay a captured launch sequence;
    several replays with eager work between them all equal the eager result."""
    p, hk, O, grains, img, table, db = scored_case
    base, _ = api.quantize_u8(img)
    rng = np.random.default_rng(3)
    imgs = [base] + [base[rng.permutation(base.size)] for _ in range(7)]
    eager = db.score_batch(imgs, [1.0] * 8, 5, 400.0)
    db.set_use_graphs(True)
    try:
        for rep in range(4):
            _say(f"graph replay {rep}")
            got = db.score_batch(imgs, [1.0] * 8, 5, 400.0)
            single = db.score(imgs[rep], 1.0, 5, 400.0)  # eager-sized work in between
            np.testing.assert_array_equal(single.ids, eager[rep].ids)
            for a, b in zip(got, eager):
                assert a.nr_results == b.nr_results
                np.testing.assert_array_equal(a.ids, b.ids)
                np.testing.assert_array_equal(a.score, b.score)
    finally:
        db.set_use_graphs(False)


@pytest.mark.parametrize("bits,loads", [(16, 3), (15, 3)])
def test_graph_replay_single_image(scored_case, bits, loads):
    """A captured single-image sequence (counter fill, LDS walk and, for 2^16
    rows, the candidate recount) replays bit-exactly, with eager calls of
    another layout and thresholds between replays; eager calls time the
    compare kernel alone (walk_ms <= kernel_ms)."""
    p, hk, O, grains, img, table, db = scored_case
    base, _ = api.quantize_u8(img)
    rng = np.random.default_rng(12)
    imgs = [base] + [base[rng.permutation(base.size)] for _ in range(3)]
    db.set_banded_tile_bits(bits)
    db.set_banded_loads(loads)
    try:
        eager = [db.score(u, 1.0, 5, 400.0) for u in imgs]
        for h in eager:
            assert 0 < h.walk_ms <= h.kernel_ms
        db.set_use_graphs(True)
        for rep in range(6):
            _say(f"single-image graph replay tb={bits} rep={rep}")
            i = rep % len(imgs)
            got = db.score(imgs[i], 1.0, 5, 400.0)
            assert got.nr_results == eager[i].nr_results
            np.testing.assert_array_equal(got.ids, eager[i].ids)
            np.testing.assert_array_equal(got.score, eager[i].score)
            np.testing.assert_array_equal(got.nspots, eager[i].nspots)
            db.set_use_graphs(False)
            other = db.score(imgs[i], 1.0, 3, 900.0)  # eager work between replays
            assert other.nr_results >= 0
            db.set_use_graphs(True)
        matched, ns, tot, nres = ob.compare(table, 64, imgs[0], p.NrPxX, 1.0, 5, 400.0)
        assert eager[0].nr_results == nres
        # imScale != 1: the float64 re-sum of the hits is part of the
        # captured sequence; alternate with imScale == 1 calls (no re-sum)
        u8n, sn = api.quantize_u8(img * 0.57 + 0.02 * (img > 0))
        assert sn != 1.0
        matched, ns, tot, nres = ob.compare(table, 64, u8n, p.NrPxX, sn, 5, 400.0 * sn)
        rows, mA = ob.collect(matched)
        for rep in range(4):
            _say(f"single-image graph replay, imScale {sn:.4f}, rep={rep}")
            got = db.score(u8n, sn, 5, 400.0 * sn)
            assert got.nr_results == nres
            np.testing.assert_array_equal(got.ids, rows)
            np.testing.assert_array_equal(got.score, mA)
            np.testing.assert_array_equal(got.totint, tot[rows.astype(np.int64)])
            one = db.score(imgs[rep], 1.0, 5, 400.0)
            np.testing.assert_array_equal(one.ids, eager[rep].ids)
    finally:
        db.set_use_graphs(False)
        db.set_banded_tile_bits(14)
        db.set_banded_loads(0)


def test_budget_admission(ctx):
    small = api.Context(0, hbm_budget=1 << 20)
    p = synth.baseline_params()
    O = synth.random_orientations(20000, seed=2)  # 1.44 MB of orientations
    with pytest.raises(_lib.LaueMatchError) as e:
        api.OrientationDB(small, O, synth.fcc_hkls(8), api.Geometry.from_params(p), 64)
    assert e.value.code == -4


@pytest.mark.parametrize("guard", [False, True])
def test_forward_without_room_for_kept_lists(ctx, guard):
    """With an HBM budget that holds the DB and its table but not every
    orientation's spot list (n x M words), the forward simulation projects
    twice instead of keeping the lists: the same table, bit for bit."""
    p = synth.baseline_params(MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(8, "box")
    n = 20000
    O = synth.random_orientations(n, seed=19)
    want, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    spots = int(want[:, 0].astype(np.int64).sum())
    # orientations + hkls + counts/tile offsets + entries + banded copy, with
    # slack, and well under the n * 64 * 4 bytes of kept lists on top
    need = n * 72 + hk.size * 4 + n * 2 + 4096 + 2 * (spots + 64) * 4 + (1 << 20)
    assert need + n * 64 * 4 // 2 < need + n * 64 * 4
    small = api.Context(0, hbm_budget=need + n * 64 * 4 // 2)
    _say(f"forward with a small budget guard={guard}")
    db = api.OrientationDB(small, O, hk, api.Geometry.from_params(p), 64)
    db.set_banded_tile_bits(0)
    db.forward(api.LM_FWD_FP32_GUARD if guard else api.LM_FWD_FP64)
    np.testing.assert_array_equal(db.export_table(), want)
    db.close()


def test_state_kept_between_calls_follows_the_stream(scored_case):
    """A call reuses the predicate the previous call uploaded only on the
    stream that uploaded it: single-image calls alternating between the
    default stream and a side stream, thresholds changing and repeating, with
    batch and on-the-fly calls and graph replays in between; every result
    equals the oracle's."""
    import torch

    p, hk, O, grains, img, table, db = scored_case
    u8, s = api.quantize_u8(img)
    assert s == 1.0
    perm = u8[np.random.default_rng(3).permutation(u8.size)]
    dev = torch.from_numpy(np.stack([u8, perm])).to("cuda:0")
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    want = {}

    def check_single(i, stream, ms, mi):
        im = (u8, perm)[i]
        key = (i, ms, mi)
        if key not in want:
            matched, ns, tot, nres = ob.compare(table, 64, im, p.NrPxX, 1.0, ms, mi)
            want[key] = ob.collect(matched)
        rows, mA = want[key]
        h = db.score_dev(dev[i].data_ptr(), 1.0, ms, mi, stream)
        np.testing.assert_array_equal(h.ids, rows)
        np.testing.assert_array_equal(h.score, mA)

    seq = [(0, 0, 5, 1000.0), (0, 0, 5, 1000.0), (1, side.cuda_stream, 5, 1000.0),
           (0, 0, 5, 1000.0), (1, 0, 3, 500.0), (0, side.cuda_stream, 3, 500.0),
           "batch", (0, 0, 3, 500.0), "fwd", (1, 0, 5, 1000.0), "graphs", (0, 0, 5, 1000.0),
           (0, side.cuda_stream, 5, 1000.0), "nographs", (1, side.cuda_stream, 3, 500.0),
           (1, 0, 3, 500.0)]
    try:
        for k, step in enumerate(seq):
            _say(f"state between calls step {k}: {step}")
            if step == "batch":
                hs = db.score_batch_dev(dev.data_ptr(), 2, [1.0, 1.0], 4, 700.0)
                for i, h in enumerate(hs):
                    matched, ns, tot, nres = ob.compare(table, 64, (u8, perm)[i], p.NrPxX, 1.0,
                                                        4, 700.0)
                    rows, mA = ob.collect(matched)
                    np.testing.assert_array_equal(h.ids, rows)
            elif step == "fwd":
                hs, _ = db.score_fwd_batch_dev(dev.data_ptr(), 2, [1.0, 1.0], 4, 700.0)
                for i, h in enumerate(hs):
                    matched, ns, tot, nres = ob.compare(table, 64, (u8, perm)[i], p.NrPxX, 1.0,
                                                        4, 700.0)
                    rows, mA = ob.collect(matched)
                    np.testing.assert_array_equal(h.ids, rows)
            elif step == "graphs":
                db.set_use_graphs(True)
            elif step == "nographs":
                db.set_use_graphs(False)
            else:
                check_single(*step)
            torch.cuda.synchronize()
    finally:
        db.set_use_graphs(False)
