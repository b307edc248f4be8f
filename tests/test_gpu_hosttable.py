"""Spot tables larger than the HBM budget (row a12 of the reference's
drivers, src/LaueMatchingGPU.cu:660-749, which streams the ForwardFile to the
device in chunks): with a context budget that holds the orientations but not
the table, the forward simulation and the ForwardFile import keep the table
on the host and every scoring call streams it through a device buffer in
several chunks.  Tables, hits, scores and nrResults equal the oracle's and
the resident table's."""
import sys

import numpy as np
import pytest

import oracle_bind as ob
from lauematching_amd import _lib, api, synth

pytestmark = pytest.mark.gpu

M = 40
N = 500_000


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def case():
    p = synth.baseline_params(MaxNrLaueSpots=M, NrPxX=512, NrPxY=384, PxX=0.0008, PxY=0.0008)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(N, seed=23)
    want, _, _ = ob.forward(O, hk, ob.geom(p), M, threads=16)
    spots = int(want[:, 0].astype(np.int64).sum())
    rng = np.random.default_rng(3)
    imgs = [synth.synth_image(p, hk, O[np.sort(rng.choice(N, int(rng.integers(2, 12)),
                                                          replace=False))], seed=40 + i)
            for i in range(4)]
    # orientations + hkls + 16 MB: the spot table (4 bytes per spot) does not fit
    base = N * 72 + hk.size * 4 + 64
    budget = base + (N + 64) * 2 + 16 * (1 << 20)
    assert spots * 4 > 16 * (1 << 20) + (1 << 20)
    return p, hk, O, want, imgs, budget


def _check_scores(db, p, want, imgs):
    for img in imgs:
        u8, s = api.quantize_u8(img)
        for ms, mi in ((5, 400.0), (3, 60.0)):
            matched, ns, tot, nres = ob.compare(want, M, u8, p.NrPxX, s, ms, mi, threads=16)
            rows, mA = ob.collect(matched)
            h = db.score(u8, s, ms, mi)
            assert h.nr_results == nres
            np.testing.assert_array_equal(h.ids, rows)
            np.testing.assert_array_equal(h.score, mA)
            np.testing.assert_array_equal(h.nspots, ns[rows.astype(np.int64)])
    u8s, scales = zip(*[api.quantize_u8(im) for im in imgs])
    hs = db.score_batch(list(u8s), list(scales), 5, 400.0)
    for h, u8, s in zip(hs, u8s, scales):
        matched, _, _, nres = ob.compare(want, M, u8, p.NrPxX, s, 5, 400.0, threads=16)
        rows, mA = ob.collect(matched)
        assert h.nr_results == nres
        np.testing.assert_array_equal(h.ids, rows)
        np.testing.assert_array_equal(h.score, mA)


@pytest.mark.parametrize("mode", [api.LM_FWD_FP64, api.LM_FWD_FP32_GUARD])
def test_forward_over_budget_keeps_table_on_host(case, mode):
    p, hk, O, want, imgs, budget = case
    ctx = api.Context(0, hbm_budget=budget)
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), M)
    _say(f"forward over budget mode={mode}")
    st = db.forward(mode)
    assert db.table_on_host
    assert st.n_spots == int(want[:, 0].astype(np.int64).sum()) == db.table_spots
    np.testing.assert_array_equal(db.export_table(), want)
    np.testing.assert_array_equal(db.export_table(12345, 777), want[12345:12345 + 777])
    # the same checksum as the resident table's
    big = api.Context(0)
    ref = api.OrientationDB(big, O, hk, api.Geometry.from_params(p), M)
    ref.forward()
    assert not ref.table_on_host
    assert db.checksum() == ref.checksum()
    ref.close()
    big.close()
    _say("score over budget")
    _check_scores(db, p, want, imgs)
    # DoFwd=1 scoring of the float64 image, streamed the same way
    _, matched, nres = ob.forward(O, hk, ob.geom(p), M, img=imgs[0], min_spots=5, min_int=400.0,
                                  threads=16)
    rows, mA = ob.collect(matched)
    h = db.score_f64(imgs[0], 5, 400.0)
    assert h.nr_results == nres
    np.testing.assert_array_equal(h.ids, rows)
    np.testing.assert_array_equal(h.score, mA)
    # what a host-kept table cannot do: autotune, serving, the no-allocation call
    with pytest.raises(_lib.LaueMatchError) as e:
        db.autotune([api.quantize_u8(imgs[0])[0]], max_batch=8, reps=1)
    assert e.value.code == -6
    from lauematching_amd.serve import Server
    srv = Server(ctx, max_batch=4, workers=1)
    with pytest.raises(_lib.LaueMatchError):
        srv.add_db(db)
    srv.close()
    db.close()
    ctx.close()


def test_import_over_budget_keeps_table_on_host(case):
    p, hk, O, want, imgs, budget = case
    ctx = api.Context(0, hbm_budget=budget)
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), M)
    _say("import over budget")
    db.import_table(want)
    assert db.table_on_host
    np.testing.assert_array_equal(db.export_table(), want)
    _check_scores(db, p, want, imgs[:2])
    bad = want.copy()
    bad[7, 0] = M + 1
    with pytest.raises(_lib.LaueMatchError):
        db.import_table(bad)
    db.close()
    ctx.close()
