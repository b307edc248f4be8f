"""Continuous-batching server over resident orientation DBs (lm_server_* of
include/lauematch.h): the serving counterpart of the reference's streaming
daemon (src/LaueMatchingGPUStream.cu:1167-1257)."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import LmHits, LmServerConfig, LmServerStats, check, lib
from .api import Context, Hits, OrientationDB


class Server:
    """Requests are images for one of the added DBs; results are the same
    ascending hit lists lm_score gives for each image alone."""

    def __init__(self, ctx: Context, max_batch: int = 32, workers: int = 2,
                 max_inflight: int = 0, use_graphs: bool = True, batch_after_ms: int = 0):
        cfg = LmServerConfig(max_batch, workers, max_inflight, 1 if use_graphs else 0,
                             batch_after_ms)
        h = ctypes.c_void_p()
        check(lib().lm_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._dbs = []

    def add_db(self, db: OrientationDB) -> int:
        i = ctypes.c_int32()
        check(lib().lm_server_add_db(self._h, db._h, ctypes.byref(i)))
        self._dbs.append(db)  # keep the DB alive while the server uses it
        return i.value

    def submit(self, db_id: int, img_u8: np.ndarray, im_scale: float, min_nr_spots: int,
               min_intensity: float) -> int:
        im = np.ascontiguousarray(img_u8, dtype=np.uint8).ravel()
        db = self._dbs[db_id]
        if im.size != db.geom.nrPxX * db.geom.nrPxY:
            raise ValueError("image size does not match the DB's detector")
        t = ctypes.c_uint64()
        check(lib().lm_server_submit(self._h, db_id, ctypes.c_void_p(im.ctypes.data),
                                     float(im_scale), int(min_nr_spots), float(min_intensity),
                                     ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int):
        h = LmHits()
        lat = ctypes.c_double()
        check(lib().lm_server_wait(self._h, ticket, ctypes.byref(h), ctypes.byref(lat)))
        return Hits._take(h), lat.value

    def poll(self, ticket: int):
        """(hits, latency_ms) if the ticket's result is ready, else None."""
        h = LmHits()
        lat = ctypes.c_double()
        ready = ctypes.c_int32()
        check(lib().lm_server_poll(self._h, ticket, ctypes.byref(h), ctypes.byref(lat),
                                   ctypes.byref(ready)))
        return (Hits._take(h), lat.value) if ready.value else None

    def stats(self) -> dict:
        s = LmServerStats()
        check(lib().lm_server_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in LmServerStats._fields_}

    def reset_stats(self) -> None:
        lib().lm_server_reset_stats(self._h)

    def close(self):
        if self._h:
            lib().lm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
