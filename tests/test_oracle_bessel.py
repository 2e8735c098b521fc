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


def test_complex_hankel_vs_mpmath(O):
    """H0, H1 of complex argument (the builder's path for k not real positive,
    galerkin_builder.cpp:224-235) against 50-digit mpmath values."""
    with open(os.path.join(GOLD, "hankel_cplx_mpmath.json")) as f:
        rows = json.load(f)["rows"]
    z = np.array([r["re"] + 1j * r["im"] for r in rows])
    h0, h1 = O.hankel01_cplx(z)
    for k, r in enumerate(rows):
        for got, ref in ((h0[k], r["h0"]), (h1[k], r["h1"])):
            ref = float(ref[0]) + 1j * float(ref[1])
            # the phase of e^{iz} is conditioned by |z| eps; values below 1e-300
            # are subnormal in double precision
            assert abs(got - ref) <= (4e-14 + 2e-16 * abs(z[k])) * abs(ref) + 1e-300, (z[k], got, ref)


def test_complex_hankel_zero_and_continuity(O):
    """z = 0 gives 0 (non-finite values are replaced by zeros,
    galerkin_builder.cpp:216-222); approaching the positive real axis the
    complex path meets the real one."""
    h0, h1 = O.hankel01_cplx([0.0])
    assert h0[0] == 0 and h1[0] == 0
    x = np.array([0.3, 1.9, 2.1, 5.0, 11.0, 40.0])
    j0, y0, j1, y1 = O.jy01(x)
    c0, c1 = O.hankel01_cplx(x + 1e-300j)
    assert np.allclose(c0, j0 + 1j * y0, rtol=1e-14, atol=0)
    assert np.allclose(c1, j1 + 1j * y1, rtol=1e-14, atol=0)
