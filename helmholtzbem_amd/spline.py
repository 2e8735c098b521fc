"""Complex Akima splines for smooth scatterers (host side).

Restates the reference's ComplexSpline (cspline.hpp, cspline.cpp:12-76): the
real and imaginary parts of y(x) interpolated by Akima splines, periodic or
not, with values and first and second derivatives.  The reference takes them
from GSL (gsl_interp_akima / gsl_interp_akima_periodic, GSL 2.x, not vendored);
here the spline is built from Akima's published method (H. Akima, A new method
of interpolation and smooth curve fitting based on local procedures, J. ACM
17, 1970):

    m_j = (y_{j+1} - y_j) / (x_{j+1} - x_j)                      secant slopes
    t_i = (|m_{i+1} - m_i| m_{i-1} + |m_{i-1} - m_{i-2}| m_i)
          / (|m_{i+1} - m_i| + |m_{i-1} - m_{i-2}|)             node slopes
    on [x_j, x_{j+1}], d = x - x_j, h = x_{j+1} - x_j:
    y = y_j + t_j d + (3 m_j - 2 t_j - t_{j+1}) / h d^2 + (t_j + t_{j+1} - 2 m_j) / h^2 d^3

When the denominator of t_i is 0 (two exactly straight runs meet at node i)
the spline follows GSL's akima: the piece starting at node i is linear with
slope m_i, and the piece ending there takes m_{i-1} as its right-end slope.

The secants beyond the ends are continued periodically (m_{-1} = m_{n-2},
m_{-2} = m_{n-3}, m_{n-1} = m_0, m_n = m_1, for data with y_{n-1} = y_0) or
by Akima's linear extrapolation (m_{-1} = 2 m_0 - m_1, m_{-2} = 3 m_0 - 2 m_1,
and alike at the right end).

``rows()`` packs the pieces for the device panel table (3 rows of 3 doubles
per piece, one closing row): {x_j, Re y_j, Im y_j}, {b, c, d} of the real part,
{b, c, d} of the imaginary part, ..., {x_K, Re y_K, Im y_K}.
"""
from __future__ import annotations

import numpy as np

_GL_X, _GL_W = np.polynomial.legendre.leggauss(16)


def _akima_slopes(x: np.ndarray, y: np.ndarray, periodic: bool) -> np.ndarray:
    n = x.size
    m = np.diff(y) / np.diff(x)                      # m_0 .. m_{n-2}
    if periodic:
        ext = np.concatenate([m[-2:], m, m[:2]])      # m_{-2}, m_{-1}, ..., m_{n-1}, m_n
    else:
        left = [3.0 * m[0] - 2.0 * m[1], 2.0 * m[0] - m[1]]
        right = [2.0 * m[-1] - m[-2], 3.0 * m[-1] - 2.0 * m[-2]]
        ext = np.concatenate([left, m, right])
    t = np.empty(n)
    flat = np.zeros(n, dtype=bool)
    for i in range(n):
        mm2, mm1, m0, mp1 = ext[i], ext[i + 1], ext[i + 2], ext[i + 3]
        w1, w2 = abs(mp1 - m0), abs(mm1 - mm2)
        flat[i] = w1 + w2 == 0.0
        t[i] = 0.0 if flat[i] else (w1 * mm1 + w2 * m0) / (w1 + w2)
    return t, flat


def _akima_pieces(x: np.ndarray, v: np.ndarray, periodic: bool):
    """(b, c, d) of every piece.  Where both weights at node j vanish (two
    exactly straight runs meet there) the piece from node j is linear with
    slope m_j, and a piece ending at such a node ends with its own secant
    slope -- GSL's convention, which keeps both runs straight."""
    h = np.diff(x)
    m = np.diff(v) / h
    t, flat = _akima_slopes(x, v, periodic)
    tl = t[:-1]
    tr = np.where(flat[1:], m, t[1:])
    b = np.where(flat[:-1], m, tl)
    c = np.where(flat[:-1], 0.0, (3.0 * m - 2.0 * tl - tr) / h)
    d = np.where(flat[:-1], 0.0, (tl + tr - 2.0 * m) / (h * h))
    return b, c, d


class ComplexSpline:
    """Akima spline of complex data y over increasing x (cspline.cpp:12-28)."""

    def __init__(self, x, y, periodic: bool = False):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.complex128)
        if x.size != y.size or x.size < (4 if periodic else 5) or np.any(np.diff(x) <= 0):
            raise ValueError("need increasing abscissae and enough points for an Akima spline")
        self.x, self.y, self.periodic = x, y, periodic
        self.coef = [_akima_pieces(x, part, periodic) for part in (y.real, y.imag)]

    def _piece(self, t):
        t = np.asarray(t, dtype=np.float64)
        j = np.clip(np.searchsorted(self.x, t, side="right") - 1, 0, self.x.size - 2)
        return t, j, t - self.x[j]

    def eval(self, t):
        """Value at t (cspline.cpp:64-66)."""
        t, j, d = self._piece(t)
        out = []
        for part, (b, c, e) in zip((self.y.real, self.y.imag), self.coef):
            out.append(part[j] + d * (b[j] + d * (c[j] + d * e[j])))
        return out[0] + 1j * out[1]

    def eval_der(self, t, order: int = 1):
        """Derivative of order 0, 1 or 2 at t (cspline.cpp:68-76)."""
        if order < 0 or order > 2:
            raise RuntimeError("Order of the spline derivative must be 0, 1 or 2")
        if order == 0:
            return self.eval(t)
        t, j, d = self._piece(t)
        out = []
        for b, c, e in self.coef:
            out.append(b[j] + d * (2.0 * c[j] + 3.0 * d * e[j]) if order == 1 else 2.0 * c[j] + 6.0 * d * e[j])
        return out[0] + 1j * out[1]

    def arc_length(self, a: float, b: float) -> float:
        """Length of the curve t -> y(t) over [a, b]: 16-point Gauss-Legendre on
        every piece of the spline that meets [a, b] (|y'| is smooth on a piece)."""
        if b <= a:
            return 0.0
        ks = self.x
        lo = max(int(np.searchsorted(ks, a, side="right")) - 1, 0)
        hi = min(int(np.searchsorted(ks, b, side="left")), ks.size - 1)
        total = 0.0
        for j in range(lo, hi):
            u, v = max(a, ks[j]), min(b, ks[j + 1])
            if v <= u:
                continue
            tt = 0.5 * (v - u) * _GL_X + 0.5 * (u + v)
            total += 0.5 * (v - u) * float(np.dot(_GL_W, np.abs(self.eval_der(tt))))
        return total

    def rows(self):
        """Device table rows of the pieces (see the module docstring)."""
        out = []
        (bx, cx, dx), (by, cy, dy) = self.coef
        for j in range(self.x.size - 1):
            out.append([self.x[j], self.y[j].real, self.y[j].imag])
            out.append([bx[j], cx[j], dx[j]])
            out.append([by[j], cy[j], dy[j]])
        out.append([self.x[-1], self.y[-1].real, self.y[-1].imag])
        return out

    @property
    def pieces(self) -> int:
        return self.x.size - 1
