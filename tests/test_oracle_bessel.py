"""Pins the oracle's J0/Y0/J1/Y1 (cbessel.cpp:1707-2017 branch structure)
against mpmath fixtures, the reference's own known answers and the Wronskian."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fixture():
    with open(os.path.join(GOLD, "bessel_mpmath.json")) as f:
        return json.load(f)["rows"]


def test_bessel_vs_mpmath(O):
    rows = _fixture()
    x = np.array([r["x"] for r in rows])
    j0, y0, j1, y1 = O.jy01(x)
    for k, (name, got) in enumerate((("J0", j0), ("Y0", y0), ("J1", j1), ("Y1", y1))):
        ref = np.array([float(r[name]) for r in rows])
        amp = np.sqrt(2.0 / (np.pi * x))
        # absolute error relative to the envelope (J, Y oscillate with amplitude ~ sqrt(2/(pi x)))
        scale = np.maximum(np.abs(ref), np.minimum(amp, 1.0))
        err = np.abs(got - ref) / scale
        big = x > 1e6   # leading-order asymptotics: relative error O(1/x) by design
        assert np.all(err[~big] < 2e-13), (name, x[~big][np.argmax(err[~big])], err[~big].max())
        assert np.all(err[big] < 1e-6), (name, err[big].max())


def test_reference_known_answers(O):
    with open(os.path.join(GOLD, "reference_known_answers.json")) as f:
        ka = json.load(f)
    j0, y0, j1, y1 = O.jy01([ka["J0_at_pi"]["x"], ka["H0prime_at_0p1"]["x"]])
    assert abs(j0[0] - ka["J0_at_pi"]["value"]) <= ka["J0_at_pi"]["tol"]
    assert abs(-j1[0] - ka["J0prime_at_pi"]["value"]) <= ka["J0prime_at_pi"]["tol"]
    h0p = -(j1[1] + 1j * y1[1])
    assert abs(h0p.real - ka["H0prime_at_0p1"]["re"]) <= ka["H0prime_at_0p1"]["tol"]
    assert abs(h0p.imag - ka["H0prime_at_0p1"]["im"]) <= ka["H0prime_at_0p1"]["tol"] * 10


def test_nonpositive_argument(O):
    j0, y0, j1, y1 = O.jy01([0.0, -1.0])
    assert np.all(j0 == 1.0) and np.all(j1 == 0.0) and np.all(y0 == 0.0) and np.all(y1 == 0.0)


def test_wronskian(O):
    x = np.concatenate([np.geomspace(1e-7, 8, 200), np.linspace(8, 600, 300)])
    j0, y0, j1, y1 = O.jy01(x)
    w = j1 * y0 - j0 * y1          # = 2/(pi x)
    assert np.max(np.abs(w * np.pi * x / 2 - 1.0)) < 5e-13
