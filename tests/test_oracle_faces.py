"""Pins the dycore's face primitives in the oracle (CPU only).

weno_face is the reference's WENO5 (c_kernels.c:64-104) rewritten with one
division; upwind_weno picks the reconstruction of the upwind side; up5 is
the 5th-order upwind flux of Wicker & Skamarock (2002).  Checks: agreement
with the restated reference formula and with the golden vectors written from
the reference's own c_kernels.c, exactness on polynomials of the stencils'
order, the linear-weight limit, mirror symmetry, and the upwind choice.
"""
import ctypes
import json
import os

import numpy as np

from oracle_bind import oracle_lib

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def ref_weno(q):
    l, r = ctypes.c_double(), ctypes.c_double()
    oracle_lib().orc_weno5(*q, ctypes.byref(l), ctypes.byref(r))
    return l.value, r.value


def face(q):
    return oracle_lib().orc_weno_face(*q)


def test_weno_face_equals_reference_formula():
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(20000):
        q = rng.normal(size=5) * rng.choice([1e-3, 1.0, 300.0])
        if rng.random() < 0.3:  # smooth data, where all three weights matter
            q = np.polyval(rng.normal(size=3), np.arange(5) * 0.1) + 0.0
        l, _ = ref_weno(q)
        f = face(q)
        # rounding differs from the three-division form only: the error is
        # measured against the scale of the data
        worst = max(worst, abs(f - l) / max(np.max(np.abs(q)), 1e-300))
    assert worst < 1e-14, worst


def test_weno_face_golden_vectors():
    data = json.load(open(os.path.join(GOLD, "weno5.json")))
    for c in data["cases"]:
        q = c["q"]
        scale = max(np.max(np.abs(q)), 1e-300)
        assert abs(face(q) - c["q_left"]) <= 4e-15 * scale, q
        # the right-biased value is the mirrored stencil
        assert abs(face(q[::-1]) - c["q_right"]) <= 4e-15 * scale, q


def test_weno_face_exact_on_constants_and_linear_limit():
    for v in (0.0, 1.0, -3.25, 300.0):
        assert abs(face([v] * 5) - v) <= 4.5e-16 * abs(v)  # to rounding
    # on a linear profile all smoothness indicators are equal, so the weights
    # reduce to (0.1, 0.6, 0.3) and the value is the 5th-order linear stencil
    q = np.array([1.0, 2.0, 3.0, 4.0, 5.0]) * 0.5 + 7.0
    lin = (2 * q[0] - 13 * q[1] + 47 * q[2] + 27 * q[3] - 3 * q[4]) / 60.0
    assert abs(face(q) - lin) < 1e-13
    assert abs(face(q) - 8.75) < 1e-13  # the face value of the line itself


def test_upwind_weno_picks_upwind_side():
    L = oracle_lib()
    q = [0.2, -0.4, 1.1, 0.9, 2.3, -1.0]  # cells f-3 .. f+2
    Up = L.orc_upwind_weno(2.0, *q)
    Um = L.orc_upwind_weno(-2.0, *q)
    assert abs(Up - 2.0 * face(q[0:5])) < 1e-15
    assert abs(Um - (-2.0) * face(q[5:0:-1])) < 1e-15
    assert L.orc_upwind_weno(0.0, *q) == 0.0


def test_up5_flux_properties():
    L = oracle_lib()
    # constant field: flux U q exactly
    assert abs(L.orc_up5(3.0, *([2.5] * 6)) - 7.5) < 1e-14
    # 5th-order: exact face value for polynomials up to degree 4 in the
    # centred part; the dissipation d5 vanishes on degree <= 4
    x = np.arange(-2, 4) - 0.5  # cell centres relative to the face at 0
    for deg in range(5):
        cc = np.random.default_rng(deg).normal(size=deg + 1)
        # cell averages of the polynomial over [x-1/2, x+1/2]
        P = np.polyint(cc)
        qa = np.polyval(P, x + 0.5) - np.polyval(P, x - 0.5)
        for U in (1.0, -1.0):
            got = L.orc_up5(U, *qa)
            want = U * np.polyval(cc, 0.0)
            if deg <= 4:
                assert abs(got - want) < 1e-12 * max(1.0, abs(want)), (deg, U, got, want)
    # upwinding: a step (0,0,0,1,1,1) gives a flux closer to the upwind value
    step = [0, 0, 0, 1, 1, 1]
    assert L.orc_up5(1.0, *step) < 0.5 < L.orc_up5(-1.0, *step) * -1.0
