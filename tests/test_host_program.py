"""bin/LaueMatchingGPU: the reference's command-line contract
(src/LaueMatchingGPU.cu:107-111; laue_index/indexer.py:33-45, 56-97) and an
end-to-end run against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from lauematching_amd import api, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bin", "LaueMatchingGPU")


def test_program_is_where_the_python_wrapper_looks():
    # the reference wrapper resolves <repo_root>/bin/LaueMatchingGPU for GPU runs
    assert os.path.isfile(EXE) and os.access(EXE, os.X_OK)


def test_usage_on_wrong_argument_count():
    r = subprocess.run([EXE, "a", "b"], capture_output=True, text=True)
    assert r.returncode == 0 and "Arguments" in r.stdout


def test_missing_parameter_file(tmp_path):
    r = subprocess.run([EXE, str(tmp_path / "none.txt"), "o", "h", "i", "1"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Could not open parameter file" in r.stdout


def _case(tmp_path, n=20000, grains=(17, 4321, 19999)):
    p = synth.baseline_params(NrPxX=1024, NrPxY=1024, PxX=0.0004, PxY=0.0004,
                              MaxNrLaueSpots=40, MinNrSpots=5, MinIntensity=400.0,
                              ForwardFile=str(tmp_path / "fwd.bin"), DoFwd=1)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(n, seed=4)
    img = synth.synth_image(p, hk, O[list(grains)], seed=6)
    O.tofile(tmp_path / "orients.bin")
    synth.write_hkl_file(str(tmp_path / "hkls.csv"), hk)
    synth.write_image(str(tmp_path / "image.bin"), img)
    return p, hk, O, img


def _run(tmp_path, p):
    p.write(str(tmp_path / "params.txt"))
    r = subprocess.run([EXE, str(tmp_path / "params.txt"), str(tmp_path / "orients.bin"),
                        str(tmp_path / "hkls.csv"), str(tmp_path / "image.bin"), "4"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _hits_file(path):
    rows = [l.split() for l in open(path) if not l.startswith("%")]
    return (np.array([int(r[0]) for r in rows], dtype=np.uint64),
            np.array([float(r[3]) for r in rows]))


def _free_port():
    import socket
    s = socket.socket()
    s.bind(("127.0.0.1", 0))
    port = s.getsockname()[1]
    s.close()
    return port


@pytest.mark.gpu
def test_stream_daemon_matches_single_image_program(tmp_path):
    """bin/LaueMatchingGPUStream: wire v1 frames (uint16 LE image number +
    float32 pixels), SIGTERM drain, ImageNr-prefixed rows equal to the
    single-image program's rows for the same images."""
    import signal
    import socket
    import struct
    import time

    p, hk, O, img = _case(tmp_path)
    p.extra = {"MinGoodSpots": p.MinNrSpots, "ResultDir": str(tmp_path / "res")}
    p.DoFwd = 1
    p.write(str(tmp_path / "params.txt"))
    rng = np.random.default_rng(2)
    frames = {1: img}
    for num in (2, 7):
        g = O[rng.choice(len(O), 3, replace=False)]
        frames[num] = synth.synth_image(p, hk, g, seed=num)
    port = _free_port()
    env = dict(os.environ, LM_STREAM_PORT=str(port))
    exe = os.path.join(ROOT, "bin", "LaueMatchingGPUStream")
    proc = subprocess.Popen([exe, str(tmp_path / "params.txt"), str(tmp_path / "orients.bin"),
                             str(tmp_path / "hkls.csv"), "4"], stdout=subprocess.PIPE,
                            stderr=subprocess.STDOUT, text=True, env=env)
    try:
        deadline = time.time() + 120
        sock = None
        while time.time() < deadline:
            try:
                sock = socket.create_connection(("127.0.0.1", port), timeout=2)
                break
            except OSError:
                assert proc.poll() is None, proc.stdout.read()
                time.sleep(0.3)
        assert sock is not None
        for num, im in frames.items():
            sock.sendall(struct.pack("<H", num) + im.astype(np.float32).tobytes())
        sock.close()
        time.sleep(2.0)
        proc.send_signal(signal.SIGTERM)
        out, _ = proc.communicate(timeout=120)
    finally:
        if proc.poll() is None:
            proc.kill()
    assert proc.returncode == 0, out
    assert "exited cleanly" in out
    rows = [l.rstrip("\n").split("\t") for l in open(tmp_path / "res" / "solutions.txt")]
    assert rows[0][0] == "%ImageNr"
    got = {}
    for r in rows[1:]:
        got.setdefault(int(r[0]), []).append(r[1:])
    p.DoFwd = 0
    for num, im in frames.items():
        synth.write_image(str(tmp_path / "image.bin"), im)
        _run(tmp_path, p)
        want = [l.rstrip("\n").split("\t") for l in open(tmp_path / "image.bin.solutions.txt")][1:]
        assert got.get(num, []) == want, num


def _v2_frame(num, db, img, dtype):
    import struct
    codes = {np.float32: 0, np.float64: 1, np.uint16: 2, np.uint8: 3}
    H, W = img.shape
    return struct.pack("<IIHHIII", 0x32464D4C, num, db, codes[dtype], W, H, 0) + \
        np.ascontiguousarray(img, dtype=dtype).tobytes()


@pytest.mark.gpu
def test_stream_daemon_two_dbs_and_wire_v2(tmp_path):
    """Two DBs (different detectors and orientation sets) in one daemon: v1
    frames to DB 0, v2 frames (float64 -- with values float32 cannot hold --,
    uint16, uint8, float32) to both; a malformed v2 header closes only its
    own connection.  Every result row
    equals the single-image program's for the same image and DB, in
    solutions.txt (DB 0) and solutions_db1.txt (DB 1)."""
    import signal
    import socket
    import time

    p0, hk, O0, img0 = _case(tmp_path)
    p0.extra = {"MinGoodSpots": p0.MinNrSpots, "ResultDir": str(tmp_path / "res")}
    p0.DoFwd = 1
    p0.write(str(tmp_path / "params.txt"))
    d1 = tmp_path / "db1"
    d1.mkdir()
    p1 = synth.baseline_params(NrPxX=512, NrPxY=384, PxX=0.0008, PxY=0.0008, MaxNrLaueSpots=40,
                               MinNrSpots=4, MinIntensity=300.0,
                               ForwardFile=str(d1 / "fwd.bin"), DoFwd=1)
    p1.extra = {"MinGoodSpots": 4, "ResultDir": str(tmp_path / "res")}
    O1 = synth.random_orientations(12000, seed=8)
    O1.tofile(d1 / "orients.bin")
    synth.write_hkl_file(str(d1 / "hkls.csv"), hk)
    p1.write(str(d1 / "params.txt"))
    rng = np.random.default_rng(5)

    def frame_for(p, O, seed, top):
        # top 0: float32 pixels; else integers 0..top (uint8 frames: 250 goes
        # through the daemon's level table, 255 is its identity case)
        im = synth.synth_image(p, hk, O[rng.choice(len(O), 3, replace=False)], seed=seed)
        return im.astype(np.float32) if top == 0 else np.floor(im * (top / im.max()))

    # a float64 frame float32 cannot hold: integers 1..254 lowered by 1e-9
    # (maximum exactly 255, so imScale is 1), which truncate one level lower
    # in float64 than after a float32 rounding
    f64 = synth.synth_image(p0, hk, O0[rng.choice(len(O0), 3, replace=False)], seed=2)
    f64 = np.floor(f64 * (255.0 / f64.max()))
    f64 = f64 - 1e-9 * ((f64 > 0) & (f64 < 255))
    assert not np.array_equal(f64.astype(np.float32).astype(np.float64), f64)
    sends = [  # (wire, number, db, image, dtype)
        ("v1", 1, 0, img0.astype(np.float32), np.float32),
        ("v2", 2, 0, f64, np.float64),
        ("v2", 3, 1, frame_for(p1, O1, 3, 60000), np.uint16),
        ("v2", 4, 1, frame_for(p1, O1, 4, 0), np.float32),
        ("v2", 5, 0, frame_for(p0, O0, 5, 250), np.uint8),
        ("v2", 6, 0, frame_for(p0, O0, 6, 255), np.uint8),
    ]
    port = _free_port()
    env = dict(os.environ, LM_STREAM_PORT=str(port))
    exe = os.path.join(ROOT, "bin", "LaueMatchingGPUStream")
    proc = subprocess.Popen([exe, str(tmp_path / "params.txt"), str(tmp_path / "orients.bin"),
                             str(tmp_path / "hkls.csv"), "4", str(d1 / "params.txt"),
                             str(d1 / "orients.bin"), str(d1 / "hkls.csv")],
                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
    try:
        deadline = time.time() + 120
        socks = None
        while time.time() < deadline:
            try:
                socks = (socket.create_connection(("127.0.0.1", port), timeout=2),
                         socket.create_connection(("127.0.0.1", port + 1), timeout=2))
                break
            except OSError:
                assert proc.poll() is None, proc.stdout.read()
                time.sleep(0.3)
        assert socks is not None
        bad = socket.create_connection(("127.0.0.1", port + 1), timeout=2)
        bad.sendall(b"\x00" * 24)  # wrong magic: this connection is closed
        bad.close()
        for wire, num, db, im, dt in sends:
            if wire == "v1":
                socks[0].sendall(np.array([num], "<u2").tobytes() + im.astype(np.float32).tobytes())
            else:
                socks[1].sendall(_v2_frame(num, db, im, dt))
        for sk in socks:
            sk.close()
        time.sleep(2.0)
        proc.send_signal(signal.SIGTERM)
        out, _ = proc.communicate(timeout=120)
    finally:
        if proc.poll() is None:
            proc.kill()
    assert proc.returncode == 0, out
    assert "bad frame header" in out and "exited cleanly" in out

    def rows_of(path):
        got = {}
        for r in [l.rstrip("\n").split("\t") for l in open(path)][1:]:
            got.setdefault(int(r[0]), []).append(r[1:])
        return got

    got0 = rows_of(tmp_path / "res" / "solutions.txt")
    got1 = rows_of(tmp_path / "res" / "solutions_db1.txt")
    p0.DoFwd = 0
    p1.DoFwd = 0
    found = 0
    for wire, num, db, im, dt in sends:
        base = tmp_path if db == 0 else d1
        p = p0 if db == 0 else p1
        synth.write_image(str(base / "image.bin"),
                          im if dt == np.float64 else im.astype(np.float32).astype(np.float64))
        p.write(str(base / "params.txt"))
        r = subprocess.run([EXE, str(base / "params.txt"), str(base / "orients.bin"),
                            str(base / "hkls.csv"), str(base / "image.bin"), "4"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        want = [l.rstrip("\n").split("\t") for l in open(base / "image.bin.solutions.txt")][1:]
        got = (got0 if db == 0 else got1).get(num, [])
        assert got == want, (wire, num, db)
        found += bool(want)
    assert found >= 4  # the frames hold grains of their DB, so solutions appear


@pytest.mark.gpu
def test_end_to_end_dofwd1_then_dofwd0(tmp_path):
    p, hk, O, img = _case(tmp_path)
    g = ob.geom(p)
    # DoFwd=1: forward simulation + float64 image
    out = _run(tmp_path, p)
    table, matched, nres = ob.forward(O, hk, g, p.MaxNrLaueSpots, img=img.ravel(),
                                      min_spots=p.MinNrSpots, min_int=p.MinIntensity)
    got = np.fromfile(tmp_path / "fwd.bin", dtype=np.uint16).reshape(table.shape)
    np.testing.assert_array_equal(got, table)
    rows, mA = ob.collect(matched)
    ids, sc = _hits_file(tmp_path / "image.bin.hits.txt")
    np.testing.assert_array_equal(ids, rows)
    np.testing.assert_array_equal(sc, mA)
    assert f"Initial solutions: {nres} " in out
    sol = [l for l in open(tmp_path / "image.bin.solutions.txt")]
    assert sol[0].startswith("%GrainNr\tNumberOfSolutions") and len(sol) >= 2
    assert all(len(l.rstrip("\n").split("\t")) >= 34 for l in sol[1:])
    best_rows = {int(l.split()[-1]) for l in sol[1:]}
    assert {17, 4321, 19999} <= best_rows
    # DoFwd=0: the ForwardFile is read back, the uint8 image is scored
    p.DoFwd = 0
    out = _run(tmp_path, p)
    u8, s = ob.quantize(img.ravel())
    matched, _, _, nres = ob.compare(table, p.MaxNrLaueSpots, u8, p.NrPxX, s, p.MinNrSpots,
                                     p.MinIntensity)
    rows, mA = ob.collect(matched)
    ids, sc = _hits_file(tmp_path / "image.bin.hits.txt")
    np.testing.assert_array_equal(ids, rows)
    np.testing.assert_array_equal(sc, mA)
    assert "file was found. Will not do forward simulation." in out
    assert f"Initial solutions: {nres} " in out


@pytest.mark.gpu
def test_missing_forward_file_runs_the_simulation(tmp_path):
    """DoFwd=0 with no ForwardFile: the program switches to simulation mode
    (src/LaueMatchingCPU.c:422-433), writes the file and gives the DoFwd=1
    results."""
    p, hk, O, img = _case(tmp_path)
    p.DoFwd = 0
    assert not (tmp_path / "fwd.bin").exists()
    out = _run(tmp_path, p)
    assert "Could not read the forward simulation file. Running in simulation mode!" in out
    table, matched, nres = ob.forward(O, hk, ob.geom(p), p.MaxNrLaueSpots, img=img.ravel(),
                                      min_spots=p.MinNrSpots, min_int=p.MinIntensity)
    np.testing.assert_array_equal(
        np.fromfile(tmp_path / "fwd.bin", dtype=np.uint16).reshape(table.shape), table)
    rows, mA = ob.collect(matched)
    ids, sc = _hits_file(tmp_path / "image.bin.hits.txt")
    np.testing.assert_array_equal(ids, rows)
    np.testing.assert_array_equal(sc, mA)


@pytest.mark.gpu
def test_truncated_forward_file_is_an_error(tmp_path):
    """A ForwardFile shorter than N x (1 + 2M) uint16 records ends the run
    with an error instead of scoring past the end of the data (the
    reference would read beyond it)."""
    p, hk, O, img = _case(tmp_path)
    _run(tmp_path, p)  # DoFwd=1 writes the file
    full = (tmp_path / "fwd.bin").read_bytes()
    (tmp_path / "fwd.bin").write_bytes(full[: len(full) // 2 + 7])
    p.DoFwd = 0
    p.write(str(tmp_path / "params.txt"))
    r = subprocess.run([EXE, str(tmp_path / "params.txt"), str(tmp_path / "orients.bin"),
                        str(tmp_path / "hkls.csv"), str(tmp_path / "image.bin"), "4"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 1
    assert "Failed to read forward simulation file" in r.stdout


@pytest.mark.gpu
def test_stream_daemon_survives_broken_clients(tmp_path):
    """A client that closes in the middle of a frame's pixels, one that sends
    only a partial header, and one that connects and leaves: none of their
    partial frames is scored, the daemon keeps serving (a complete frame on
    a new connection gets its rows) and exits cleanly on SIGTERM
    (src/LaueMatchingGPUStream.cu:295-383)."""
    import signal
    import socket
    import struct
    import time

    p, hk, O, img = _case(tmp_path)
    p.extra = {"MinGoodSpots": p.MinNrSpots, "ResultDir": str(tmp_path / "res")}
    p.DoFwd = 1
    p.write(str(tmp_path / "params.txt"))
    port = _free_port()
    env = dict(os.environ, LM_STREAM_PORT=str(port), LM_STREAM_AUTOTUNE="0")
    exe = os.path.join(ROOT, "bin", "LaueMatchingGPUStream")
    proc = subprocess.Popen([exe, str(tmp_path / "params.txt"), str(tmp_path / "orients.bin"),
                             str(tmp_path / "hkls.csv"), "4"], stdout=subprocess.PIPE,
                            stderr=subprocess.STDOUT, text=True, env=env)
    payload = img.astype(np.float32).tobytes()
    try:
        deadline = time.time() + 120
        while time.time() < deadline:
            try:
                socket.create_connection(("127.0.0.1", port), timeout=2).close()  # leaves
                break
            except OSError:
                assert proc.poll() is None, proc.stdout.read()
                time.sleep(0.3)
        s1 = socket.create_connection(("127.0.0.1", port), timeout=2)
        s1.sendall(struct.pack("<H", 11) + payload[: len(payload) // 3])  # mid-payload
        s1.close()
        s2 = socket.create_connection(("127.0.0.1", port), timeout=2)
        s2.sendall(b"\x0c")  # one byte of the 2-byte header
        s2.close()
        s3 = socket.create_connection(("127.0.0.1", port), timeout=2)
        s3.sendall(struct.pack("<H", 13) + payload)  # a complete frame
        s3.close()
        time.sleep(2.0)
        proc.send_signal(signal.SIGTERM)
        out, _ = proc.communicate(timeout=120)
    finally:
        if proc.poll() is None:
            proc.kill()
    assert proc.returncode == 0, out
    assert "exited cleanly" in out
    assert "Received image 11" not in out and "Received image 12" not in out
    assert "Received image 13" in out
    rows = [l.rstrip("\n").split("\t") for l in open(tmp_path / "res" / "solutions.txt")][1:]
    nums = {int(r[0]) for r in rows}
    assert nums == {13}
