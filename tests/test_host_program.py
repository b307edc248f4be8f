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
