"""Pins of the CPU oracle (oracle/lm_oracle.c) against reference output and
against what the physics and geometry fix independently of its formulas.

* Reference C output: ``tests/golden/ni_spots_golden.json`` (hkl -> pixel and
  unit scattering vector, written by the reference indexer).
* Bragg's law: E = hc / (2 d sin(theta)) with d from the lattice and 2theta
  from the detector geometry (inverse projection with scipy's rotation).
* A closed-form case: a (111) reflection sent to 2theta = 90 deg lands on the
  detector centre; its first harmonic inside 5-30 keV is (222).
* Composition: the table builder equals "first accepted, unseen pixel wins,
  stop at MaxNrLaueSpots" over the single-step projector.
* Quantization known answers, brute-force gathers for the compare step.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_bind as ob
from lauematching_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
HC = 1.2398419739


def _fixture():
    with open(os.path.join(HERE, "golden", "ni_spots_golden.json")) as f:
        return json.load(f)


def _params_from_fixture(fx):
    g = fx["geometry"]
    return synth.Params(LatticeParameter=tuple(g["LatticeParameter"]), P_Array=tuple(g["P_Array"]),
                        R_Array=tuple(g["R_Array"]), PxX=g["PxX"], PxY=g["PxY"],
                        NrPxX=g["NrPxX"], NrPxY=g["NrPxY"], Elo=g["Elo"], Ehi=g["Ehi"])


def test_reference_spots_golden():
    """Every spot the reference printed is reproduced: same truncated pixel,
    q-hat equal to the printed 6 decimals."""
    fx = _fixture()
    g = ob.geom(_params_from_fixture(fx))
    n = 0
    for grain in fx["grains"]:
        M = np.array(grain["recip_rowmajor"]).reshape(3, 3)
        for s in grain["spots"]:
            ok, ipx, ipy, qh, E = ob.project(M, s["hkl"], g)
            assert ok, s
            assert (ipx, ipy) == (s["X"], s["Y"]), s
            np.testing.assert_allclose(qh, s["qhat"], atol=1.5e-6)
            assert g.Elo <= E <= g.Ehi
            n += 1
    assert n == 26


def test_reference_rotation_matches_scipy():
    """Detector rotation: the transposed Rodrigues matrix equals scipy's
    active rotation by the R_Array rotation vector, transposed."""
    for r in ([-1.20161887, -1.21404493, -1.21852276], [0.3, -0.2, 0.9], [1e-3, 0, 0]):
        rt = ob.detector_rotT(r)
        np.testing.assert_allclose(rt, Rotation.from_rotvec(r).as_matrix().T, atol=1e-14)


def test_recip_known_answers():
    """Cubic: B = (2 pi / a) I.  Hexagonal a=0.295, c=0.468165: |a*| =
    4 pi / (sqrt(3) a), |c*| = 2 pi / c, a*.c* = 0.  Rhombohedral (SG 166)
    alpha=90 reduces to cubic in a rotated frame: B^T B = (2 pi / a)^2 I."""
    B = ob.calc_recip([0.36, 0.36, 0.36, 90, 90, 90], 225)
    np.testing.assert_allclose(B, np.eye(3) * 2 * math.pi / 0.36, rtol=1e-14, atol=1e-9)
    H = ob.calc_recip([0.295, 0.295, 0.468165, 90, 90, 120], 194)
    a_star, c_star = H[:, 0], H[:, 2]
    assert abs(np.linalg.norm(a_star) - 4 * math.pi / (math.sqrt(3) * 0.295)) < 1e-9
    assert abs(np.linalg.norm(c_star) - 2 * math.pi / 0.468165) < 1e-9
    assert abs(a_star @ c_star) < 1e-9
    R = ob.calc_recip([0.4, 0.4, 0.4, 90, 90, 90], 166)
    np.testing.assert_allclose(R.T @ R, np.eye(3) * (2 * math.pi / 0.4) ** 2, rtol=1e-12)


def _align(u, v):
    """Rotation matrix taking unit vector u to unit vector v."""
    u = np.asarray(u, float) / np.linalg.norm(u)
    v = np.asarray(v, float) / np.linalg.norm(v)
    ax = np.cross(u, v)
    s, c = np.linalg.norm(ax), u @ v
    return Rotation.from_rotvec(ax / s * math.atan2(s, c)).as_matrix()


def test_detector_centre_closed_form():
    p = synth.baseline_params()
    g = ob.geom(p)
    a = 0.36
    U = _align([1, 1, 1], [0, 1, -1])
    M = U @ (np.eye(3) * 2 * math.pi / a)
    ok, *_ , E111 = ob.project(M, [1, 1, 1], g)
    # Bragg: d111 = a/sqrt(3), theta = 45 deg
    e111 = HC / (2 * (a / math.sqrt(3)) * math.sin(math.pi / 4))
    assert abs(E111 - e111) < 1e-9 * e111
    assert not ok and e111 < p.Elo  # 4.22 keV is below the window
    ok, ipx, ipy, qh, E222 = ob.project(M, [2, 2, 2], g)
    assert ok and abs(E222 - 2 * e111) < 1e-9 * e111
    assert (ipx, ipy) == (1023, 1023)  # trunc(0 + 0.5*2047)
    # inclusive window edges (src/LaueMatchingCPU.c:635)
    for lo, hi, want in [(E222, 30.0, True), (5.0, E222, True),
                         (np.nextafter(E222, 99), 30.0, False), (5.0, np.nextafter(E222, 0), False)]:
        g2 = ob.geom(synth.baseline_params(Elo=lo, Ehi=hi))
        assert ob.project(M, [2, 2, 2], g2)[0] == want


def _inverse_qhat(ipx, ipy, p):
    """Scattering vector of the centre of pixel (ipx, ipy): detector point in
    the detector frame, rotated to the lab with scipy, kf - ki normalised."""
    R = Rotation.from_rotvec(p.R_Array).as_matrix()
    xp = (ipx + 0.5 - 0.5 * (p.NrPxX - 1)) * p.PxX + p.P_Array[0]
    yp = (ipy + 0.5 - 0.5 * (p.NrPxY - 1)) * p.PxY + p.P_Array[1]
    kf = R @ np.array([xp, yp, p.P_Array[2]])
    kf /= np.linalg.norm(kf)
    q = kf - np.array([0.0, 0.0, 1.0])
    return q / np.linalg.norm(q), kf


@pytest.mark.parametrize("geo", ["baseline", "ni"])
def test_pixel_inverse_and_bragg_energy(geo):
    p = synth.baseline_params() if geo == "baseline" else _params_from_fixture(_fixture())
    g = ob.geom(p)
    hk = synth.fcc_hkls(8, "box")
    lat = p.LatticeParameter
    B = np.eye(3) * 2 * math.pi / lat[0]
    O = synth.random_orientations(60, seed=7)
    # pixel angular size bounds the q-hat error of a truncated pixel
    tol = 2.5 * p.PxX / p.P_Array[2]
    nacc = 0
    for o in O:
        U = o.reshape(3, 3)
        M = U @ ob.calc_recip(lat, 225)
        for h in hk:
            ok, ipx, ipy, qh, E = ob.project(M, h, g)
            q = U @ B @ h
            qref = q / np.linalg.norm(q)
            if not ok:
                continue
            nacc += 1
            np.testing.assert_allclose(qh, qref, atol=1e-12)
            qpix, kf = _inverse_qhat(ipx, ipy, p)
            assert np.linalg.norm(qpix - qref) < tol
            theta = 0.5 * math.acos(np.clip(kf[2], -1, 1))
            d = lat[0] / math.sqrt(float(h @ h))
            Eb = HC / (2 * d * math.sin(theta))
            assert abs(Eb - E) < 5e-3 * E
            assert p.Elo <= E <= p.Ehi
    assert nacc > 60 * 10


def test_table_is_first_unseen_pixel_capped():
    p = synth.baseline_params()
    g = ob.geom(p)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(40, seed=3)
    for M_cap in (64, 7):
        table, _, _ = ob.forward(O, hk, g, M_cap)
        for i, o in enumerate(O):
            Mx = o.reshape(3, 3) @ np.array(g.recip).reshape(3, 3)
            want, seen = [], set()
            for h in hk:
                ok, ipx, ipy, _, _ = ob.project(Mx, h, g)
                if ok and (ipx, ipy) not in seen:
                    seen.add((ipx, ipy))
                    want.append((ipx, ipy))
                    if len(want) == M_cap:
                        break
            rec = np.zeros(1 + 2 * M_cap, dtype=np.uint16)
            rec[0] = len(want)
            for s, (x, y) in enumerate(want):
                rec[1 + 2 * s], rec[2 + 2 * s] = x, y
            np.testing.assert_array_equal(table[i], rec)


def test_quantize_known_answers():
    img = np.array([0.0, -3.0, 1e-9, 510.0, 1000.0, 1020.0, 254.9999 * 4, 3.9])
    u8, s = ob.quantize(img)
    assert s == 1020.0 / 255.0
    np.testing.assert_array_equal(u8, [0, 0, 1, 127, 250, 255, 254, 1])
    u8, s = ob.quantize(np.array([0.0, 7.0, 255.0, 128.0]))
    assert s == 1.0
    np.testing.assert_array_equal(u8, [0, 7, 255, 128])
    u8, s = ob.quantize(np.zeros(5))
    assert s == 1.0 and not u8.any()


def test_compare_bruteforce_and_collect():
    p = synth.baseline_params(NrPxX=64, NrPxY=48)
    rng = np.random.default_rng(11)
    n, M = 3000, 9
    table = np.zeros((n, 1 + 2 * M), dtype=np.uint16)
    cnt = rng.integers(0, M + 1, n)
    table[:, 0] = cnt
    for i in range(n):
        table[i, 1:1 + 2 * cnt[i]:2] = rng.integers(0, 64, cnt[i])
        table[i, 2:2 + 2 * cnt[i]:2] = rng.integers(0, 48, cnt[i])
    img = rng.integers(0, 256, 64 * 48).astype(np.uint8)
    img[rng.random(img.size) < 0.3] = 0
    # brute force, imScale = 1 (integer sums are exact)
    vals = [img[table[i, 2:2 + 2 * cnt[i]:2].astype(int) * 64 + table[i, 1:1 + 2 * cnt[i]:2]]
            for i in range(n)]
    ns = np.array([(v > 0).sum() for v in vals])
    tot = np.array([int(v.astype(np.int64).sum()) for v in vals])
    for min_sp, min_int in [(3, 500.0), (0, 0.0), (5, 1e9)]:
        matched, ns_o, tot_o, hits = ob.compare(table, M, img, 64, 1.0, min_sp, min_int)
        np.testing.assert_array_equal(ns_o, ns)
        np.testing.assert_array_equal(tot_o, tot)
        hit = (ns >= min_sp) & (tot >= min_int)
        assert hits == hit.sum()
        exp = np.where(hit, tot * np.sqrt(ns.astype(float)), 0.0)
        np.testing.assert_array_equal(matched, exp)
        rows, mA = ob.collect(matched)
        nz = np.nonzero(exp)[0]  # a zero score (0 spots) is not collected
        np.testing.assert_array_equal(rows, nz)
        np.testing.assert_array_equal(mA, exp[nz])
    # a non-unit scale: sequential float64 sum in record order
    s = 0.7312
    matched, ns_o, tot_o, _ = ob.compare(table, M, img, 64, s, 1, 0.0)
    for i in range(0, n, 37):
        acc = 0.0
        for v in vals[i]:
            if v > 0:
                acc += float(v) * s
        assert tot_o[i] == acc
    # thread-count invariance
    m1 = ob.compare(table, M, img, 64, s, 2, 100.0, threads=1)[0]
    m8 = ob.compare(table, M, img, 64, s, 2, 100.0, threads=8)[0]
    np.testing.assert_array_equal(m1, m8)


def test_forward_f64_scoring_matches_compare_on_integer_image():
    """DoFwd=1 scores the float64 image (src/LaueMatchingCPU.c:647-653); on an
    integer image with max 255 the uint8 path (:668-682) gives the same hits."""
    p = synth.baseline_params(NrPxX=512, NrPxY=512, PxX=0.0008, PxY=0.0008)
    g = ob.geom(p)
    hk = synth.fcc_hkls(8, "box")
    O = synth.random_orientations(2000, seed=5)
    img = synth.synth_image(p, hk, O[[10, 500]], seed=2)
    table, matched_f, hits_f = ob.forward(O, hk, g, 64, img=img.ravel(), min_spots=5,
                                          min_int=300.0)
    u8, s = ob.quantize(img.ravel())
    assert s == 1.0
    matched_u, _, _, hits_u = ob.compare(table, 64, u8, 512, s, 5, 300.0)
    np.testing.assert_array_equal(matched_f, matched_u)
    assert hits_f == hits_u and matched_f[10] > 0 and matched_f[500] > 0
