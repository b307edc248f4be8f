"""Batches of 128 and 129..256 images through the 4-bit walk (LM_OPT_BATCH_NIBBLE):
the walk keeps every (row, image) pair whose nonzero count and 17 x nibble sum
can pass, and the recount decides them from the uint8 images, so every image
must equal the oracle bit for bit with its own thresholds: boundary
thresholds equal to a row's exact sum (and just above it), thresholds at
multiples of 17, sparse and blank images, images with imScale != 1, rows
that fail only on their spot count.  Candidate overflow, batches the 4-bit
walk cannot take and the option turned off must give the same results
through the uint8 walks (two calls above 128 images)."""
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
def case():
    ctx = api.Context(0)
    p = synth.baseline_params(MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(8, "box")
    n = 60000
    O = synth.random_orientations(n, seed=31)
    grains = [7, 31000, 59998]
    img = synth.synth_image(p, hk, O[grains], seed=11)
    table, _, _ = ob.forward(O, hk, ob.geom(p), 64)
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), 64)
    db.forward()
    return p, img, table, db


def _check(p, table, imgs, scales, msp, mint, got):
    for i in range(len(imgs)):
        matched, ns, tot, nres = ob.compare(table, 64, imgs[i], p.NrPxX, scales[i], msp[i],
                                            mint[i])
        rows, mA = ob.collect(matched)
        assert got[i].nr_results == nres, (i, msp[i], mint[i])
        np.testing.assert_array_equal(got[i].ids, rows)
        np.testing.assert_array_equal(got[i].score, mA)
        np.testing.assert_array_equal(got[i].nspots, ns[rows.astype(np.int64)])
        np.testing.assert_array_equal(got[i].totint, tot[rows.astype(np.int64)])


def _mixed(p, img, table, nimg, seed):
    """nimg images with per-image thresholds, all MinIntensity > 0."""
    rng = np.random.default_rng(seed)
    base, _ = api.quantize_u8(img)
    imgs, scales, msp, mint = [], [], [], []
    fixed = [(5, 600.0), (3, 2500.5), (8, 1500.0), (300, 10.0), (5, 1e12), (0, 300.0),
             (2, 17.0 * 40), (4, 1000.0), (1, 17.0 * 3)]
    for i in range(nimg):
        kind = i % 9
        if kind == 4:
            u8, s = api.quantize_u8(img * 0.61 + 0.015 * (img > 0))  # imScale != 1
        elif kind == 6:
            u8, s = np.where(base > np.percentile(base, 97.0), base, 0).astype(np.uint8), 1.0
        elif kind == 8 and i % 2:
            u8, s = np.zeros_like(base), 1.0
        else:
            u8, s = (base if i < 9 else base[rng.permutation(base.size)]), 1.0
        if kind in (1, 7) and i < 60:
            # boundary: a row's exact sum passes, the next representable value does not
            _, _, tot, _ = ob.compare(table, 64, u8, p.NrPxX, s, 1, 1e-300)
            top = float(tot.max())
            ms, mi = 1, (top if kind == 1 else float(np.nextafter(top, np.inf)))
        elif kind == 6:
            ms, mi = 5, 50.0
        else:
            ms, mi = fixed[i % len(fixed)]
            mi = mi * s
        imgs.append(u8)
        scales.append(s)
        msp.append(ms)
        mint.append(mi)
    return imgs, scales, msp, mint


@pytest.mark.parametrize("nimg,mask", [(256, 0x180), (200, 0x180), (129, 0x180), (128, 0x180),
                                       (97, 0x180), (160, 0x80), (128, 0x100)])
def test_nibble_batch_matches_oracle(case, nimg, mask):
    p, img, table, db = case
    imgs, scales, msp, mint = _mixed(p, img, table, nimg, seed=nimg)
    _say(f"nibble batch nimg={nimg} mask={mask:#x}")
    db.set_batch_nibble(mask)
    try:
        got = db.score_batch_each(imgs, scales, msp, mint)
    finally:
        db.set_batch_nibble(0x180)
    _check(p, table, imgs, scales, msp, mint, got)
    assert any(len(g.ids) for g in got)


def test_nibble_candidate_overflow_and_zero_threshold(case):
    """MinIntensity 1 on dense images makes nearly every (row, image) pair a
    candidate (more than the list holds): the batch is rescored through the
    uint8 walks.  A batch with one MinIntensity <= 0 never takes the 4-bit
    walk."""
    p, img, table, db = case
    base, _ = api.quantize_u8(img)
    rng = np.random.default_rng(2)
    imgs = [base] + [base[rng.permutation(base.size)] for _ in range(255)]
    _say("nibble candidate overflow, 256 images")
    got = db.score_batch(imgs, [1.0] * 256, 1, 1.0)
    _check(p, table, imgs[:6], [1.0] * 6, [1] * 6, [1.0] * 6, got[:6])
    _check(p, table, imgs[-3:], [1.0] * 3, [1] * 3, [1.0] * 3, got[-3:])
    msp = [5] * 256
    mint = [600.0] * 256
    mint[130] = 0.0
    _say("nibble batch with a zero threshold, 256 images")
    got = db.score_batch_each(imgs, [1.0] * 256, msp, mint)
    idx = [0, 1, 129, 130, 131, 255]
    _check(p, table, [imgs[i] for i in idx], [1.0] * 6, [msp[i] for i in idx],
           [mint[i] for i in idx], [got[i] for i in idx])


def test_nibble_graph_replay(case):
    """Captured nibble sequences (pack, counter fill, walk, recount) replayed
    with eager work between them equal the eager result every time."""
    p, img, table, db = case
    base, _ = api.quantize_u8(img)
    rng = np.random.default_rng(4)
    imgs = [base] + [base[rng.permutation(base.size)] for _ in range(199)]
    eager = db.score_batch(imgs, [1.0] * 200, 5, 500.0)
    assert len(eager[0].ids) > 0
    db.set_use_graphs(True)
    try:
        for rep in range(3):
            _say(f"nibble graph replay {rep}")
            got = db.score_batch(imgs, [1.0] * 200, 5, 500.0)
            one = db.score(imgs[rep], 1.0, 5, 500.0)
            np.testing.assert_array_equal(one.ids, eager[rep].ids)
            for a, b in zip(got, eager):
                assert a.nr_results == b.nr_results
                np.testing.assert_array_equal(a.ids, b.ids)
                np.testing.assert_array_equal(a.score, b.score)
    finally:
        db.set_use_graphs(False)


def test_nibble_option_is_checked(case):
    p, img, table, db = case
    for bad in (0x40, 0x200, -1, 0x181):
        with pytest.raises(_lib.LaueMatchError):
            db.set_batch_nibble(bad)
    assert db.get_option(7) == 0x180


def test_nibble_unaligned_device_images(case):
    """Device images at an address that is not 4-byte aligned cannot be packed
    as nibbles: 256 such images are scored as two uint8 calls, 128 through the
    uint8 walks, with the same results as aligned images."""
    import torch

    p, img, table, db = case
    base, _ = api.quantize_u8(img)
    rng = np.random.default_rng(6)
    imgs = [base] + [base[rng.permutation(base.size)] for _ in range(255)]
    npx = base.size
    flat = torch.zeros(256 * npx + 16, dtype=torch.uint8, device="cuda:0")
    host = torch.from_numpy(np.concatenate(imgs))
    flat[1:1 + 256 * npx].copy_(host)
    aligned = torch.from_numpy(np.stack(imgs)).to("cuda:0")
    torch.cuda.synchronize()
    for nimg in (256, 128):
        _say(f"nibble unaligned device images nimg={nimg}")
        got = db.score_batch_dev(flat.data_ptr() + 1, nimg, [1.0] * nimg, 5, 500.0)
        want = db.score_batch_dev(aligned.data_ptr(), nimg, [1.0] * nimg, 5, 500.0)
        torch.cuda.synchronize()
        for a, b in zip(got, want):
            assert a.nr_results == b.nr_results
            np.testing.assert_array_equal(a.ids, b.ids)
            np.testing.assert_array_equal(a.score, b.score)
        _check(p, table, imgs[:2], [1.0] * 2, [5] * 2, [500.0] * 2, got[:2])
