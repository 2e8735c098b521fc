"""Incoming waves for the transmission solve (host side).

Restates the reference's ``incoming`` namespace (incoming.hpp:29-143,
incoming.cpp:20-232): solutions of Delta u + k^2 u = 0 summed from a list of
wave specifications, with their gradients, for the right-hand side of
``tp::direct_second_kind``.

File format (``load``, incoming.cpp:47-92), one wave per line, all whitespace
removed first, empty lines and lines starting with ``#`` skipped::

    Plane, <amplitude>, <x0_1>, <x0_2>, <angle>
    CircularJ | CircularY | FourierHankel1 | FourierHankel2, <amplitude>, <x0_1>, <x0_2>, <order>
    Herglotz, <amplitude>, <x0_1>, <x0_2>, <angle>, <eps>

Numbers are read like std::stod / std::stoi: the longest numeric prefix of a
field counts, a field without one or out of range fails the load, and so does
an unknown wave name, a missing field or an extra one.  As in the reference, a
file that cannot be opened loads as an empty wave.

Waves at x (p = x - x0, r = |p|, theta = atan2(p_2, p_1), d = (cos a, sin a)):

    Plane            exp(i k p.d)
    CircularJ / Y    J_l(k r) e^{i l theta} / Y_l(k r) e^{i l theta}
    FourierHankel1/2 CircularJ +- i CircularY
    Herglotz         sum_j w_j exp(i k (p_1 cos t_j + p_2 sin t_j)),
                     t_j = a - eps + 2 eps s_j, (s_j, w_j) the 21-point
                     Gauss-Legendre rule on [0, 1] (weights summing to 1)

Bessel functions of integer order and complex argument come from
scipy.special (Amos' algorithms, as the reference's complex_bessel).  Points
may be one 2-vector or an array (..., 2); values broadcast over the leading
axes and gradients carry a trailing axis of 2.
"""
from __future__ import annotations

import enum
import math
import re
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np
from scipy import special

_GL_X, _GL_W = np.polynomial.legendre.leggauss(21)
_GL_X = 0.5 * (_GL_X + 1.0)      # getGaussQR(21, 0, 1) (incoming.cpp:22)
_GL_W = 0.5 * _GL_W


class WaveType(enum.Enum):
    """incoming.hpp:35."""
    Plane = 0
    CircularJ = 1
    CircularY = 2
    FourierHankel1 = 3
    FourierHankel2 = 4
    Herglotz = 5


@dataclass
class wavespec_t:
    """incoming.hpp:37-44."""
    type: WaveType
    amplitude: float = 0.0
    eps: float = 0.0
    angle: float = 0.0
    order: int = 0
    x0: Tuple[float, float] = (0.0, 0.0)


_C_SPACE = " \t\n\v\f\r"
_DEC = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")
_HEX = re.compile(r"[+-]?0[xX](?:[0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)(?:[pP][+-]?\d+)?")
_SPECIAL = re.compile(r"[+-]?(?:inf(?:inity)?|nan(?:\([0-9A-Za-z_]*\))?)", re.IGNORECASE)
_INT = re.compile(r"[+-]?\d+")
_DBL_MIN = 2.2250738585072014e-308


def stod(tok: str) -> Optional[float]:
    """std::stod: the value of the longest numeric prefix, None where stod
    throws (no conversion, or a result out of range)."""
    m = _HEX.match(tok)
    if m:
        txt = m.group(0)
        sign = -1.0 if txt.startswith("-") else 1.0
        body = txt.lstrip("+-")
        if "p" not in body.lower():
            body += "p0"
        try:
            v = sign * float.fromhex(body)
        except OverflowError:
            return None
        return v
    m = _SPECIAL.match(tok)
    if m:
        return float(m.group(0).split("(")[0])
    m = _DEC.match(tok)
    if not m:
        return None
    txt = m.group(0)
    v = float(txt)
    if math.isinf(v):
        return None                                   # ERANGE overflow
    mant = txt.lstrip("+-").split("e")[0].split("E")[0]
    if (v == 0.0 and any(c in "123456789" for c in mant)) or (v != 0.0 and abs(v) < _DBL_MIN):
        return None                                   # ERANGE underflow
    return v


def stoi(tok: str) -> Optional[int]:
    """std::stoi: the longest integer prefix, None on no conversion or when
    the value does not fit an int."""
    m = _INT.match(tok)
    if not m:
        return None
    v = int(m.group(0))
    return v if -2 ** 31 <= v < 2 ** 31 else None


_NAMES = {t.name: t for t in WaveType}


def load(fname: str) -> Optional[List[wavespec_t]]:
    """The wave of a file (incoming.cpp:47-92), or None where the reference's
    load returns false."""
    try:
        with open(fname) as f:
            lines = f.read().split("\n")
    except OSError:
        return []
    w: List[wavespec_t] = []
    for line in lines:
        s = "".join(ch for ch in line if ch not in _C_SPACE)
        if not s or s[0] == "#":
            continue
        toks = s.split(",")
        if toks[-1] == "":
            toks.pop()            # getline after a final ',' extracts nothing
        wt = _NAMES.get(toks[0])
        if wt is None:
            return None
        fields = toks[1:]
        nreal = 3 + (1 if wt == WaveType.Plane else 2 if wt == WaveType.Herglotz else 0)
        nint = 0 if wt in (WaveType.Plane, WaveType.Herglotz) else 1
        if len(fields) != nreal + nint:
            return None
        vals = [stod(t) for t in fields[:nreal]]
        if any(v is None for v in vals):
            return None
        p = wavespec_t(wt, amplitude=vals[0], x0=(vals[1], vals[2]))
        if wt == WaveType.Plane:
            p.angle = vals[3]
        elif wt == WaveType.Herglotz:
            p.angle, p.eps = vals[3], vals[4]
        else:
            o = stoi(fields[3])
            if o is None:
                return None
            p.order = o
        w.append(p)
    return w


def _rel(x, x0):
    x = np.asarray(x, dtype=np.float64)
    return x[..., 0] - x0[0], x[..., 1] - x0[1]


def _polar(x, x0):
    p1, p2 = _rel(x, x0)
    return np.sqrt(p1 * p1 + p2 * p2), np.arctan2(p2, p1)


def _vec(a, b):
    return np.stack(np.broadcast_arrays(a, b), axis=-1)


def plane(x, x0, angle: float, k):
    """exp(i k (x - x0).d), d = (cos angle, sin angle) (incoming.cpp:139-142)."""
    p1, p2 = _rel(x, x0)
    return np.exp(1j * complex(k) * (p1 * math.cos(angle) + p2 * math.sin(angle)))


def plane_del(x, x0, angle: float, k):
    """incoming.cpp:144-147."""
    u = 1j * complex(k) * plane(x, x0, angle, k)
    return _vec(u * math.cos(angle), u * math.sin(angle))


def circular_J(x, x0, l: int, k):
    """J_l(k r) e^{i l theta} (incoming.cpp:149-154)."""
    r, th = _polar(x, x0)
    return special.jv(l, complex(k) * r) * np.exp(1j * l * th)


def _circ_del(x, x0, l, k, f, origin_J):
    r, th = _polar(x, x0)
    k = complex(k)
    a = k * r
    dfdr = k * (f(l - 1, a) - f(l + 1, a)) / 2.0
    dfdt = f(l, a) * (1j * l)
    s, c = np.sin(th), np.cos(th)
    with np.errstate(divide="ignore", invalid="ignore"):
        g1 = dfdr * c - dfdt * s / r
        g2 = dfdr * s + dfdt * c / r
    at0 = r == 0
    if np.any(at0):
        if origin_J:
            # the limit of the gradient at the center (incoming.cpp:163-165)
            h1 = dfdr * c - (1j * s * k * 0.5 if abs(l) == 1 else 0.0)
            h2 = dfdr * s + (1j * c * k * 0.5 if abs(l) == 1 else 0.0)
        else:
            h1 = h2 = 0.0     # circular_Y_del returns 0 there (incoming.cpp:178)
        g1 = np.where(at0, h1, g1)
        g2 = np.where(at0, h2, g2)
    e = np.exp(1j * l * th)
    if not origin_J:
        e = np.where(at0, 1.0, e)
    return _vec(g1 * e, g2 * e)


def circular_J_del(x, x0, l: int, k):
    """incoming.cpp:156-168."""
    return _circ_del(x, x0, l, k, special.jv, True)


def circular_Y(x, x0, l: int, k):
    """Y_l(k r) e^{i l theta} (incoming.cpp:170-175)."""
    r, th = _polar(x, x0)
    return special.yv(l, complex(k) * r) * np.exp(1j * l * th)


def circular_Y_del(x, x0, l: int, k):
    """incoming.cpp:177-189."""
    return _circ_del(x, x0, l, k, special.yv, False)


def fourier_hankel(x, x0, kind: int, l: int, k):
    """CircularJ + i CircularY (kind 1) or - i (kind 2) (incoming.cpp:191-193)."""
    return circular_J(x, x0, l, k) + (1j if kind == 1 else -1j) * circular_Y(x, x0, l, k)


def fourier_hankel_del(x, x0, kind: int, l: int, k):
    """incoming.cpp:195-197."""
    return circular_J_del(x, x0, l, k) + (1j if kind == 1 else -1j) * circular_Y_del(x, x0, l, k)


def _herglotz_terms(x, x0, angle, eps, k):
    p1, p2 = _rel(x, x0)
    th = (angle - eps) + (2.0 * eps) * _GL_X
    e = np.exp(1j * complex(k) * (np.asarray(p1)[..., None] * np.cos(th) + np.asarray(p2)[..., None] * np.sin(th)))
    return th, e


def herglotz(x, x0, angle: float, eps: float, k):
    """Mean of the plane waves with directions in [angle - eps, angle + eps]
    (incoming.cpp:199-205)."""
    _, e = _herglotz_terms(x, x0, angle, eps, k)
    return (e * _GL_W).sum(axis=-1)


def herglotz_del(x, x0, angle: float, eps: float, k):
    """incoming.cpp:207-218."""
    th, e = _herglotz_terms(x, x0, angle, eps, k)
    ik = 1j * complex(k)
    return _vec(ik * (np.cos(th) * e * _GL_W).sum(axis=-1), ik * (np.sin(th) * e * _GL_W).sum(axis=-1))


def _one(wp: wavespec_t, x, k):
    t = wp.type
    if t == WaveType.Plane:
        return plane(x, wp.x0, wp.angle, k)
    if t == WaveType.CircularJ:
        return circular_J(x, wp.x0, wp.order, k)
    if t == WaveType.CircularY:
        return circular_Y(x, wp.x0, wp.order, k)
    if t == WaveType.FourierHankel1:
        return fourier_hankel(x, wp.x0, 1, wp.order, k)
    if t == WaveType.FourierHankel2:
        return fourier_hankel(x, wp.x0, 2, wp.order, k)
    return herglotz(x, wp.x0, wp.angle, wp.eps, k)


def _one_del(wp: wavespec_t, x, k):
    t = wp.type
    if t == WaveType.Plane:
        return plane_del(x, wp.x0, wp.angle, k)
    if t == WaveType.CircularJ:
        return circular_J_del(x, wp.x0, wp.order, k)
    if t == WaveType.CircularY:
        return circular_Y_del(x, wp.x0, wp.order, k)
    if t == WaveType.FourierHankel1:
        return fourier_hankel_del(x, wp.x0, 1, wp.order, k)
    if t == WaveType.FourierHankel2:
        return fourier_hankel_del(x, wp.x0, 2, wp.order, k)
    return herglotz_del(x, wp.x0, wp.angle, wp.eps, k)


def compute(w: List[wavespec_t], x, k):
    """Sum of amplitude * wave over the specification (incoming.cpp:95-120)."""
    res = np.zeros(np.asarray(x).shape[:-1], dtype=np.complex128)
    for wp in w:
        res = res + wp.amplitude * _one(wp, x, k)
    return res[()] if res.ndim == 0 else res


def compute_del(w: List[wavespec_t], x, k):
    """Gradient of compute (incoming.cpp:122-137)."""
    res = np.zeros(np.asarray(x).shape[:-1] + (2,), dtype=np.complex128)
    for wp in w:
        res = res + wp.amplitude * _one_del(wp, x, k)
    return res


def as_functions(w: List[wavespec_t], k):
    """(u_inc, u_inc_del) as callables of (x1, x2), the form the transmission
    solver takes (the reference drivers wrap compute / compute_del the same
    way)."""
    return (lambda x1, x2: complex(compute(w, (x1, x2), k)),
            lambda x1, x2: compute_del(w, (x1, x2), k))
