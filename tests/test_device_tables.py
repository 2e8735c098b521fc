"""Pins the coefficient table of the device J0/Y0/J1/Y1 (hankel_tables.inc,
read by hbem_device.hpp) on the CPU: the table's polynomials, the three-part
pi/2 of the sin/cos reduction and the branch formulas, evaluated in double
precision with exactly rounded FMAs (as the device does), against the mpmath
fixture (50 digits) and against numpy's sin/cos."""
import json
import os
import re

import mpmath as mp
import numpy as np
import pytest

HERE = os.path.dirname(__file__)
INC = os.path.join(HERE, "..", "helmholtzbem_amd", "csrc", "hankel_tables.inc")


def _table():
    txt = open(INC).read()
    defs = {k: int(v) for k, v in re.findall(r"#define (HBEM_\w+) (\d+)\n", txt)}
    body = txt.split("#define HBEM_BESSEL_TABLE_INIT {", 1)[1].split("}", 1)[0]
    vals = [float(v) for v in re.findall(r"[-+0-9.eE]+", body.replace("\\", " "))]
    assert len(vals) == defs["HBEM_TAB_SIZE"]
    return defs, vals


def fma(a, b, c):
    """a * b + c rounded once (mpmath exact arithmetic, then to double)."""
    return float(mp.fadd(mp.fmul(a, b, exact=True), c, exact=True))


def horner4(c, off, deg, z):
    v = [fma(z, c[off + 4 * deg + f], c[off + 4 * deg - 4 + f]) for f in range(4)]
    for i in range(deg - 2, -1, -1):
        v = [fma(v[f], z, c[off + 4 * i + f]) for f in range(4)]
    return v


def sincos_med(x, defs, c):
    T = defs["HBEM_TAB_SC"]
    deg = defs["HBEM_SC_DEG"]
    n = float(np.rint(x * c[T]))
    r = fma(-n, c[T + 1], x)
    r = fma(-n, c[T + 2], r)
    r = fma(-n, c[T + 3], r)
    u = r * r
    k = T + 4
    ps = fma(u, c[k + 2 * deg], c[k + 2 * deg - 2])
    pc = fma(u, c[k + 2 * deg + 1], c[k + 2 * deg - 1])
    for i in range(deg - 2, -1, -1):
        ps = fma(ps, u, c[k + 2 * i])
        pc = fma(pc, u, c[k + 2 * i + 1])
    sr = fma(r * u, ps, r)
    cr = fma(u * u, pc, fma(-0.5, u, 1.0))
    q = int(n) & 3
    s0, c0 = (cr, sr) if q & 1 else (sr, cr)
    return (-s0 if q & 2 else s0), (-c0 if (q + 1) & 2 else c0)


def device_jy01(x, defs, c):
    """J0, Y0, J1, Y1 as bessel_jy01 in hbem_device.hpp computes them."""
    two_pi = 2.0 / np.pi
    if x < 8.0:
        y = (0.125 * x) ** 2
        A, B, C, D = horner4(c, defs["HBEM_TAB_SMALL"], defs["HBEM_SMALL_DEG"], y)
        L2 = two_pi * np.log(0.5 * x)
        J1 = x * B
        return A, fma(L2, A, C), J1, fma(L2, J1, fma(x, D, -two_pi * (1.0 / x)))
    s, co = sincos_med(x, defs, c)
    b = 1.0 / np.sqrt(np.pi * x)
    cs, sc = co + s, s - co
    xinv = 1.0 / x
    z = 64.0 * xinv * xinv
    P0, Q0, P1, Q1 = horner4(c, defs["HBEM_TAB_PQ"], defs["HBEM_PQ_DEG"], z)
    Q0 *= xinv
    Q1 *= xinv
    return (b * fma(P0, cs, -Q0 * sc), b * fma(P0, sc, Q0 * cs), b * fma(P1, sc, Q1 * cs), b * fma(Q1, sc, -P1 * cs))


def device_jy01_touching(x, defs, c):
    """J0, Y0, J1, Y1 as the assembly kernels of touching panel pairs compute
    them (bessel_jy01_f<true> in hbem_assembly.hip): below HBEM_TINY_X the entire
    parts come from the low-degree table fitted on (0, HBEM_TINY_X)."""
    if x >= defs["HBEM_TINY_X"]:
        return device_jy01(x, defs, c)
    two_pi = 2.0 / np.pi
    y = (0.125 * x) ** 2
    A, B, C, D = horner4(c, defs["HBEM_TAB_TINY"], defs["HBEM_TINY_DEG"], y)
    L2 = two_pi * np.log(0.5 * x)
    J1 = x * B
    return A, fma(L2, A, C), J1, fma(L2, J1, fma(x, D, -two_pi * (1.0 / x)))


def device_jy01_coinciding(x, defs, c):
    """As the assembly kernel of coinciding panel pairs computes them
    (bessel_jy01_f<2>): below HBEM_TINY1_X from the degree-HBEM_TINY1_DEG table."""
    if x >= defs["HBEM_TINY1_X"]:
        return device_jy01_touching(x, defs, c)
    two_pi = 2.0 / np.pi
    y = (0.125 * x) ** 2
    A, B, C, D = horner4(c, defs["HBEM_TAB_TINY1"], defs["HBEM_TINY1_DEG"], y)
    L2 = two_pi * np.log(0.5 * x)
    J1 = x * B
    return A, fma(L2, A, C), J1, fma(L2, J1, fma(x, D, -two_pi * (1.0 / x)))


def test_table_layout():
    defs, c = _table()
    assert defs["HBEM_TAB_SMALL"] == 0
    assert defs["HBEM_TAB_PQ"] == 4 * (defs["HBEM_SMALL_DEG"] + 1)
    assert defs["HBEM_TAB_SC"] == defs["HBEM_TAB_PQ"] + 4 * (defs["HBEM_PQ_DEG"] + 1)
    assert defs["HBEM_TAB_TINY"] == defs["HBEM_TAB_SC"] + 4 + 2 * (defs["HBEM_SC_DEG"] + 1)
    assert defs["HBEM_TAB_TINY1"] == defs["HBEM_TAB_TINY"] + 4 * (defs["HBEM_TINY_DEG"] + 1)
    assert defs["HBEM_TAB_SIZE"] == defs["HBEM_TAB_TINY1"] + 4 * (defs["HBEM_TINY1_DEG"] + 1)
    T = defs["HBEM_TAB_SC"]
    assert c[T] == 2.0 / np.pi and c[T + 1] == np.pi / 2
    # the three parts sum to pi/2 far beyond double precision
    mp.mp.dps = 60
    assert abs(mp.mpf(c[T + 1]) + c[T + 2] + c[T + 3] - mp.pi / 2) < mp.mpf(10) ** -45


def test_device_sincos_matches_numpy():
    defs, c = _table()
    rng = np.random.default_rng(5)
    xs = np.concatenate([np.linspace(8.0, 200.0, 3001), rng.uniform(200.0, 1e6, 2000),
                         (np.arange(5, 200000, 997) * np.pi / 2)])
    err = 0.0
    for x in xs:
        s, co = sincos_med(float(x), defs, c)
        err = max(err, abs(s - np.sin(x)), abs(co - np.cos(x)))
    assert err < 4e-16, err


def test_device_bessel_vs_mpmath():
    defs, c = _table()
    with open(os.path.join(HERE, "golden", "bessel_mpmath.json")) as f:
        rows = [r for r in json.load(f)["rows"] if 0.0 < r["x"] <= 1e6]
    assert len(rows) > 20
    for r in rows:
        x = r["x"]
        got = device_jy01(x, defs, c)
        amp = min(np.sqrt(2.0 / (np.pi * x)), 1.0)
        for name, g in zip(("J0", "Y0", "J1", "Y1"), got):
            ref = float(r[name])
            # the generator's stated fit errors: 2.2e-14 absolute for the x < 8
            # entire parts, 1.5e-16 relative for the asymptotic amplitudes
            tol = 5e-14 if x < 8.0 else 2e-15
            assert abs(g - ref) <= tol * max(abs(ref), amp), (name, x, g, ref)


def test_device_bessel_touching_pairs_vs_mpmath():
    """The low-degree table of touching panel pairs against the 50-digit
    fixture below HBEM_TINY_X, and against the full table: it must not be less
    accurate than the (0, 8) fit it replaces there."""
    defs, c = _table()
    xt = defs["HBEM_TINY_X"]
    with open(os.path.join(HERE, "golden", "bessel_mpmath.json")) as f:
        rows = [r for r in json.load(f)["rows"] if 0.0 < r["x"] < xt]
    assert len(rows) > 5
    for r in rows:
        x = r["x"]
        got = device_jy01_touching(x, defs, c)
        for name, g in zip(("J0", "Y0", "J1", "Y1"), got):
            ref = float(r[name])
            assert abs(g - ref) <= 2e-15 * max(abs(ref), 1.0), (name, x, g, ref)
    # dense sweep of (0, HBEM_TINY_X) against mpmath directly, for both the
    # adjacent (x < 2 table) and the coinciding (x < 1 table first) forms
    mp.mp.dps = 30
    for x in np.concatenate([np.geomspace(1e-5, 1.0, 25), np.linspace(0.5, xt * 0.9999, 25)]):
        x = float(x)
        ref = [mp.besselj(0, x), mp.bessely(0, x), mp.besselj(1, x), mp.bessely(1, x)]
        for fn in (device_jy01_touching, device_jy01_coinciding):
            got = fn(x, defs, c)
            for g, e in zip(got, ref):
                assert abs(g - float(e)) <= 2e-15 * max(abs(float(e)), 1.0), (fn.__name__, x, g, e)
