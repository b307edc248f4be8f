"""lm_score_dev_noalloc (scoring with every device buffer carved from the
library's fixed scratch array) against lm_score_dev on the same DB and device
image, through the C ABI, on two calls in a row -- the second with a smaller
DB and detector -- and against the CPU oracle.  Each call is also repeated
after calls on the other DB, since both share the one scratch array."""
import ctypes
import sys

import numpy as np
import pytest

import oracle_bind as ob
from lauematching_amd import _lib, api, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _noalloc(db, dev_ptr, scale, min_sp, min_int, cap=1 << 16):
    ids = np.zeros(cap, np.uint64)
    ns = np.zeros(cap, np.int32)
    tot = np.zeros(cap)
    sc = np.zeros(cap)
    count, nres = ctypes.c_uint64(), ctypes.c_uint64()
    _lib.check(_lib.lib().lm_score_dev_noalloc(db._h, dev_ptr, scale, min_sp, min_int, None,
                                               cap, ids.ctypes.data, ns.ctypes.data,
                                               tot.ctypes.data, sc.ctypes.data,
                                               ctypes.byref(count), ctypes.byref(nres)))
    k = count.value
    return ids[:k], ns[:k], tot[:k], sc[:k], nres.value


def _case(ctx, n, det, seed, factor):
    import torch
    over = {} if det == 2048 else dict(NrPxX=det, NrPxY=det, PxX=0.0002 * 2048 / det,
                                       PxY=0.0002 * 2048 / det)
    p = synth.baseline_params(MaxNrLaueSpots=64, **over)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(n, seed=seed)
    img = synth.synth_image(p, hk, O[[1, n // 2, n - 3]], seed=seed + 1) * factor
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), 64)
    db.forward()
    table = db.export_table()
    u8, s = api.quantize_u8(img)
    d = torch.from_numpy(u8).to("cuda:0")
    torch.cuda.synchronize()
    return p, db, table, u8, s, d


def _check(db, table, p, u8, s, d, min_sp, min_int):
    ref = db.score_dev(d.data_ptr(), s, min_sp, min_int)
    ids, ns, tot, sc, nres = _noalloc(db, d.data_ptr(), s, min_sp, min_int)
    assert nres == ref.nr_results
    np.testing.assert_array_equal(ids, ref.ids)
    np.testing.assert_array_equal(ns, ref.nspots)
    np.testing.assert_array_equal(tot, ref.totint)
    np.testing.assert_array_equal(sc, ref.score)
    matched, ons, otot, onres = ob.compare(table, 64, u8, p.NrPxX, s, min_sp, min_int)
    rows, mA = ob.collect(matched)
    assert nres == onres
    np.testing.assert_array_equal(ids, rows)
    np.testing.assert_array_equal(sc, mA)
    return len(ids)


def test_noalloc_equals_score_dev_two_calls():
    ctx = api.Context(0)
    big = _case(ctx, 60000, 2048, 41, 1.0)
    small = _case(ctx, 7000, 1024, 43, 0.61)  # imScale != 1
    for rnd in range(2):
        for name, (p, db, table, u8, s, d) in (("60000 x 2048^2", big), ("7000 x 1024^2", small)):
            for layout in ((16, 3), (15, 3), (0, 0)):
                db.set_banded_tile_bits(layout[0])
                db.set_banded_loads(layout[1])
                for min_sp, min_int in ((5, 400.0 * (1.0 if s == 1.0 else 0.61)), (3, 50.0)):
                    _say(f"noalloc round {rnd} {name} layout {layout} thr ({min_sp}, {min_int})")
                    _check(db, table, p, u8, s, d, min_sp, min_int)
    # output arrays too small for the candidate records: a clear error, the
    # number needed in count
    p, db, table, u8, s, d = big
    ids, ns, tot, sc, nres = _noalloc(db, d.data_ptr(), s, 3, 50.0)
    assert len(ids) > 1
    with pytest.raises(_lib.LaueMatchError) as e:
        _noalloc(db, d.data_ptr(), s, 3, 50.0, cap=1)
    assert e.value.code == -5
