"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/*.h declares, and its host-only functions match the oracle bit for
bit (no GPU needed)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle_bind as ob
from lauematching_amd import _lib, api, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(lm_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(_lib.SIGNATURES) >= names


def test_version_and_errors():
    assert b"gfx950" in _lib.lib().lm_version()
    h = ctypes.c_void_p()
    rc = _lib.lib().lm_ctx_create(0, 0, ctypes.byref(h))
    try:
        import torch
        gpu = torch.cuda.is_available()
    except Exception:
        gpu = False
    if not gpu:
        assert rc != 0 and _lib.lib().lm_last_error()
    elif rc == 0:
        _lib.lib().lm_ctx_destroy(h)


@pytest.mark.parametrize("lat,sg,R", [
    ((0.36, 0.36, 0.36, 90, 90, 90), 225, (synth.R_ARRAY_90DEG,) * 3),
    ((0.295, 0.295, 0.468165, 90, 90, 120), 194, (-1.20161887, -1.21404493, -1.21852276)),
    ((0.51, 0.51, 0.51, 60, 60, 60), 166, (0.2, -0.4, 1.1)),
    ((0.3, 0.41, 0.52, 88, 95, 101), 2, (0.0, 1e-3, 0.0)),
])
def test_geometry_bit_identical_to_oracle(lat, sg, R):
    p = synth.baseline_params(LatticeParameter=lat, SpaceGroup=sg, R_Array=R)
    g = api.Geometry.from_params(p)
    og = ob.geom(p)
    assert g.recip.tobytes() == np.array(og.recip).reshape(3, 3).tobytes()
    assert g.rotT.tobytes() == np.array(og.rotT).reshape(3, 3).tobytes()


def test_quantize_bit_identical_to_oracle():
    rng = np.random.default_rng(4)
    for scale in (1.0, 0.37, 1234.5):
        img = rng.gamma(0.3, 20.0, 100000) * scale
        img[rng.random(img.size) < 0.2] = 0
        img[rng.random(img.size) < 0.01] = -1
        a, sa = api.quantize_u8(img)
        b, sb = ob.quantize(img)
        assert sa == sb
        np.testing.assert_array_equal(a, b)


def test_quantize_f32_equals_quantize_of_widened_pixels():
    """lm_quantize_u8_f32 (the daemon's float32 frames) gives the oracle's
    quantization of the same pixels widened to float64."""
    import ctypes
    L = _lib.lib()
    rng = np.random.default_rng(9)
    for scale in (1.0, 0.37, 1234.5, 3e-30):
        img = (rng.gamma(0.3, 20.0, 100000) * scale).astype(np.float32)
        img[rng.random(img.size) < 0.2] = 0
        img[rng.random(img.size) < 0.01] = -1
        img[:3] = [np.nan, np.inf, -np.inf] if scale == 1.0 else img[:3]
        out = np.zeros(img.size, np.uint8)
        s = ctypes.c_double()
        _lib.check(L.lm_quantize_u8_f32(img.ctypes.data, img.size, out.ctypes.data,
                                        ctypes.byref(s)))
        b, sb = ob.quantize(img.astype(np.float64))
        assert s.value == sb
        np.testing.assert_array_equal(out, b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_quantize_level_boundaries_and_ragged_lengths(dtype):
    """Pixels on and one ulp either side of every level boundary x = k * mx /
    255 (the 1 and 255 clamps included), subnormals, signed zeros, NaN and
    -inf, at every length 0..47 (the vector loops' tails) and in full: both
    quantizers are byte-equal to the oracle."""
    import ctypes
    L = _lib.lib()
    rng = np.random.default_rng(12)
    tiny = np.finfo(dtype).smallest_subnormal
    for mx in (1000.0, 3.0, 255.0, 7.3e-3):
        mx = float(dtype(mx))
        base = np.arange(0, 257, dtype=np.float64) * (mx / 255.0)
        vals = np.concatenate([base, np.nextafter(base, np.inf), np.nextafter(base, -np.inf)])
        vals = vals.astype(dtype)
        vals = vals[vals <= mx]
        special = np.array([0.0, -0.0, np.nan, -np.inf, -1.0, tiny, -tiny, mx], dtype)
        img = np.concatenate([vals, special])
        img = img[rng.permutation(img.size)]
        for n in list(range(48)) + [img.size]:
            x = np.ascontiguousarray(img[:n])
            want, sw = ob.quantize(x.astype(np.float64))
            if dtype == np.float64:
                got, sg = api.quantize_u8(x)
            else:
                got = np.zeros(n, np.uint8)
                s = ctypes.c_double()
                _lib.check(L.lm_quantize_u8_f32(x.ctypes.data, n, got.ctypes.data,
                                                ctypes.byref(s)))
                sg = s.value
            assert sg == sw, (mx, n)
            np.testing.assert_array_equal(got, want, err_msg=f"mx={mx} n={n}")


def test_invalid_geometry_rejected():
    with pytest.raises(_lib.LaueMatchError):
        api.Geometry((0, 0, 0, 90, 90, 90), 225, (0, 0, .5), (0.1, 0, 0), 1e-4, 1e-4, 64, 64, 5, 30)
    with pytest.raises(_lib.LaueMatchError):
        api.Geometry((.3, .3, .3, 90, 90, 90), 225, (0, 0, .5), (0.1, 0, 0), 1e-4, 1e-4, 0, 64, 5, 30)


def test_read_params_prefix_matching(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text("LatticeParameter 0.35 0.35 0.35 90 90 90\nNrPxX 1024 # comment\nNrPxY 512\n"
                 "MaxNrLaueSpotsXYZ 30\nElo 6\nEhi 25.5\nMinIntensity 12.5\nDoFwd 0\n"
                 "ForwardFile /dev/shm/x.bin\nP_Array 0 0 0.4\n")
    p = api.read_params(str(f))
    assert p.LatticeParameter[0] == 0.35 and p.NrPxX == 1024 and p.NrPxY == 512
    assert p.MaxNrLaueSpots == 30  # prefix match, as the C parser does
    assert p.Elo == 6 and p.Ehi == 25.5 and p.MinIntensity == 12.5 and p.DoFwd == 0
    assert p.ForwardFile == "/dev/shm/x.bin" and p.P_Array == (0.0, 0.0, 0.4)
    assert p.MinNrSpots == 5  # C default


def test_one_hip_runtime_per_process():
    """With PyTorch installed, the library binds to the HIP runtime torch has
    loaded: two runtimes side by side cannot both open the device."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "from lauematching_amd import _lib; _lib.lib()\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(libs))\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "1"
