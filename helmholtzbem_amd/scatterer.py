"""Scatterer input files and polygonal meshing (host side).

Restates the reference's ``scatterer`` namespace (scatterer.hpp:39-92,
scatterer.cpp:29-262): the on-disk format of ``data/scatterer_*.dat`` and the
two ways of splitting a polygon into line panels that the reference drivers
use before the sigma_min(k) path.

File format (``read_array``, scatterer.cpp:29-72): one row per line, numbers
separated by commas; all whitespace is removed first, empty lines are skipped.
A token is read as std::stod reads it: the longest prefix strtod accepts (so
"1.5abc" is 1.5); a token with no such prefix, or out of the double range,
makes the whole read fail.  Rows shorter than the widest row are padded with
zeros.  ``read`` requires exactly two columns
(x, y) and multiplies by ``scale`` (scatterer.cpp:74-81).
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Tuple

import numpy as np

from .geometry import ParametrizedLine, ParametrizedMesh, ParametrizedPolynomial
from .spline import ComplexSpline

_MAX_ITER_ALPHA = 5000


_libc = ctypes.CDLL(None, use_errno=True)
_libc.strtod.restype = ctypes.c_double
_libc.strtod.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]


def _stod(tok: str) -> Optional[float]:
    """std::stod: strtod on the token, None where stod throws (nothing
    converted: invalid_argument; ERANGE: out_of_range)."""
    buf = ctypes.create_string_buffer(tok.encode())
    end = ctypes.c_char_p()
    ctypes.set_errno(0)
    v = _libc.strtod(buf, ctypes.byref(end))
    used = ctypes.cast(end, ctypes.c_void_p).value - ctypes.addressof(buf)
    if used == 0 or ctypes.get_errno() != 0:
        return None
    return float(v)


def _cround(v: float) -> float:
    """std::round for v >= 0: halves away from zero."""
    r = math.floor(v)
    return r + 1.0 if v - r >= 0.5 else r


def read_array(fname: str) -> Optional[np.ndarray]:
    """The numeric table of a scatterer / incoming-wave file, or None when a
    token is not a number or the file has no rows (scatterer.cpp:29-72)."""
    rows = []
    width = 0
    try:
        with open(fname) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    for line in lines:
        s = "".join(ch for ch in line if not ch.isspace())
        if not s:
            continue
        toks = s.split(",")
        if toks[-1] == "":
            toks.pop()          # std::getline yields no field after a trailing comma
        row = []
        for tok in toks:
            v = _stod(tok)
            if v is None:
                return None
            row.append(v)
        rows.append(row)
        width = max(width, len(row))
    if not rows:
        return None
    out = np.zeros((len(rows), width))
    for i, r in enumerate(rows):
        out[i, :len(r)] = r
    return out


def read(fname: str, scale: float = 1.0) -> Optional[Tuple[np.ndarray, np.ndarray]]:
    """Polygon vertices (x, y) of a scatterer file, or None (scatterer.cpp:74-81)."""
    ar = read_array(fname)
    if ar is None or ar.shape[1] != 2:
        return None
    return scale * ar[:, 0], scale * ar[:, 1]


def _edges(x, y):
    """Vertices and edge lengths |p_{i+1} - p_i| = sqrt(dx^2 + dy^2) (the
    reference's Vector2d::norm); plain sequential arithmetic throughout, so the
    C++ header (include/hbem/scatterer.hpp) reproduces every panel bit for bit."""
    x = [float(v) for v in x]
    y = [float(v) for v in y]
    n = len(x)
    d = []
    for i in range(n):
        j = i + 1 if i + 1 < n else 0
        dx, dy = x[j] - x[i], y[j] - y[i]
        d.append(math.sqrt(dx * dx + dy * dy))
    return x, y, d


def length(x, y) -> float:
    """Perimeter of the closed polygon (scatterer.cpp:83-92)."""
    L = 0.0
    for v in _edges(x, y)[2]:
        L += v
    return L


def auto_num_panels(x, y, f: float) -> int:
    """round(perimeter / (shortest edge * f)) (scatterer.cpp:94-112)."""
    x, y, d = _edges(x, y)
    if not x:
        return 0
    tot, mn = 0.0, math.inf
    for v in d:
        tot += v
        mn = min(mn, v)
    if mn == 0.0:
        raise ValueError("polygon has coincident consecutive vertices")
    return int(_cround(tot / (mn * f)))


def _line(a, b) -> ParametrizedLine:
    return ParametrizedLine((float(a[0]), float(a[1])), (float(b[0]), float(b[1])))


def make_N_polygonal_panels(x, y, N: int) -> List[ParametrizedLine]:
    """Exactly N line panels: edge i gets round(N d_i / L) panels, then single
    panels are added to / removed from the edge that keeps the variance of the
    panel lengths d_i / m_i smallest, one at a time, until the total is N
    (scatterer.cpp:114-155).  Each edge is split uniformly with
    ParametrizedLine::split.  With more vertices than panels some edges get no
    panel and the variance is not finite; the first edge is then taken."""
    x, y, d = _edges(x, y)
    n = len(x)
    if n == 0:
        raise ValueError("empty polygon")
    alpha = N / length(x, y)
    m = [_cround(alpha * di) for di in d]
    cs = 0
    for v in m:
        cs += int(v)
    while cs != N:
        k = -1 if cs > N else 1
        best, i_min = math.inf, 0
        for i in range(n):
            m[i] += k
            a = [d[t] / m[t] if m[t] != 0 else math.inf for t in range(n)]
            mean = 0.0
            for v in a:
                mean += v
            mean /= n
            var = 0.0
            for v in a:
                var += (v - mean) * (v - mean)
            var /= n
            if var < best:
                best, i_min = var, i
            m[i] -= k
        m[i_min] += k
        cs += k
    panels: List[ParametrizedLine] = []
    for i in range(n):
        j = i + 1 if i + 1 < n else 0
        panels.extend(ParametrizedLine((x[i], y[i]), (x[j], y[j])).split(int(m[i])))
    return panels


def make_polygonal_panels(x, y, N: int, refinement_factor: float = 0.5) -> List[ParametrizedLine]:
    """About N line panels, graded toward the polygon's corners
    (scatterer.cpp:157-262).

    Edge i gets k_i = N d_i / L panels (the solution of the reference's linear
    system E k = b: k_i / d_i equal for all edges, sum k_i = N), rounded down or
    up edge by edge in order of decreasing k_i by the objective
    std(d/k)/mean(d/k) + |N / sum k - 1|.  An edge with n_i >= 3 panels and
    refinement_factor < 1 is split geometrically: the panel lengths shrink
    toward both ends by a common ratio alpha (Newton's method on the closure
    equation for every split point j, keeping the j whose two end panels are
    closest in length), the end panels scaled by
    f = r + (1 - r)(1 - cos(corner angle))/2.  If the graded panels do not
    close on the edge within 1e-3 of a uniform panel, the edge is split
    uniformly."""
    with np.errstate(all="ignore"):   # coincident vertices give inf / NaN as in C
        return _graded_panels(x, y, N, refinement_factor)


def _graded_panels(x, y, N, refinement_factor):
    x, y, d = _edges(x, y)
    f64 = np.float64
    x, y, d = [f64(v) for v in x], [f64(v) for v in y], [f64(v) for v in d]
    refinement_factor = f64(refinement_factor)
    n = len(x)
    eps = 2.220446049250313e-16
    L = f64(0.0)
    for v in d:
        L += v
    k = [N * di / L for di in d]
    order = sorted(range(n), key=lambda i: (k[i], i), reverse=True)

    def obj():
        q = [d[t] / k[t] for t in range(n)]
        mu = f64(0.0)
        for v in q:
            mu += v
        mu /= n
        ss = f64(0.0)
        for v in q:
            ss += (v - mu) * (v - mu)
        ks = f64(0.0)
        for v in k:
            ks += v
        return np.sqrt(ss / n) / mu + abs(N / ks - 1.0)

    for j in order:
        k[j] = np.floor(k[j])
        f1 = obj()
        k[j] += 1.0
        f2 = obj()
        if f1 == f1 and f1 < f2:
            k[j] -= 1.0

    def sub(a, b):
        return (a[0] - b[0], a[1] - b[1])

    def dot(a, b):
        return a[0] * b[0] + a[1] * b[1]

    def norm(a):
        return np.sqrt(dot(a, a))

    res: List[ParametrizedLine] = []
    for i in range(n):
        p0 = (x[i - 1], y[i - 1]) if i > 0 else (x[n - 1], y[n - 1])
        p1 = (x[i], y[i])
        i2 = i + 1 if i + 1 < n else 0
        i3 = i + 2 if i < n - 2 else i - n + 2
        p2 = (x[i2], y[i2])
        p3 = (x[i3], y[i3])
        D = ((p2[0] - p1[0]) / k[i], (p2[1] - p1[1]) / k[i])
        n_i = int(k[i])
        u, v = sub(p0, p1), sub(p2, p1)
        f1 = (1.0 - dot(u, v) / (norm(u) * norm(v))) / 2.0
        u, v = sub(p1, p2), sub(p3, p2)
        f2 = (1.0 - dot(u, v) / (norm(u) * norm(v))) / 2.0
        f1 = refinement_factor + (1.0 - refinement_factor) * f1
        f2 = refinement_factor + (1.0 - refinement_factor) * f2
        if refinement_factor >= 1.0 or n_i < 3:
            for j in range(n_i):
                a = (p1[0] + j * D[0], p1[1] + j * D[1])
                b = (p1[0] + (j + 1) * D[0], p1[1] + (j + 1) * D[1]) if j + 1 < n_i else p2
                res.append(_line(a, b))
            continue
        j_best, alpha_best, d1_best, d2_best = 0, 0.0, 0.0, 0.0
        for j in range(1, n_i):
            alpha_next, it = refinement_factor, 0
            while True:
                alpha = alpha_next
                fv = f1 * alpha ** (1 - j) + f2 * alpha ** (j + 1 - n_i) - alpha * (f1 + f2) - n_i * (1.0 - alpha)
                fd = n_i - f1 - f2 - f1 * (j - 1.0) * alpha ** (-j) - f2 * (n_i - j - 1.0) * alpha ** (j - n_i)
                alpha_next = alpha - fv / fd
                it += 1
                if not (abs(alpha_next - alpha) > eps and it <= _MAX_ITER_ALPHA):
                    break
            d1 = f1 / alpha ** (j - 1)
            d2 = f2 / alpha ** (n_i - j - 1)
            if not j_best or abs(d1 - d2) < abs(d1_best - d2_best):
                d1_best, d2_best, j_best, alpha_best = d1, d2, j, alpha
        pos = p1
        graded = []
        for j in range(n_i):
            sc = d1_best * alpha_best ** (j_best - 1 - j) if j < j_best else d2_best * alpha_best ** (j - j_best)
            step = (D[0] * sc, D[1] * sc)
            nxt = (pos[0] + step[0], pos[1] + step[1])
            graded.append(_line(pos, nxt if j + 1 < n_i else p2))
            pos = nxt
        if norm(sub(pos, p2)) / norm(D) > 1e-3:
            graded = []
            pos = p1
            for j in range(n_i):
                nxt = (pos[0] + D[0], pos[1] + D[1])
                graded.append(_line(pos, nxt if j + 1 < n_i else p2))
                pos = nxt
        res.extend(graded)
    return res


class SmoothScatterer:
    """A smooth closed curve through the points of a scatterer file
    (scatterer.hpp:140-250, scatterer.cpp:385-533): a periodic Akima spline
    z(t), t in [0, 1], through the points closed by the first one, the
    parameter of point i being the polygon length up to it over the
    perimeter.  ``panelize(n)`` splits it into n ParametrizedPolynomial panels
    of equal arc length."""

    TOL = 1e-8                                   # scatterer.cpp:387
    MAX_ITER = 1000                              # scatterer.cpp:500

    def __init__(self, fname: str, scale: float = 1.0):
        xy = read(fname, scale)
        if xy is None:
            raise RuntimeError("Failed to read data for smooth scatterer")
        x = [float(v) for v in xy[0]] + [float(xy[0][0])]
        y = [float(v) for v in xy[1]] + [float(xy[1][0])]
        n = len(x) - 1
        if n < 3:
            raise RuntimeError("Failed to read data for smooth scatterer")
        L = length(x, y)
        t = [0.0]
        for i in range(n):
            dx, dy = x[i] - x[i + 1], y[i] - y[i + 1]
            t.append(math.sqrt(dx * dx + dy * dy) / L + t[i])
        self.t_vert = np.array(t)
        self.spline = ComplexSpline(self.t_vert, np.array(x) + 1j * np.array(y), periodic=True)
        self._total = -1.0

    def curve_point(self, t: float):
        z = complex(self.spline.eval(t))
        return (z.real, z.imag)

    def total_length(self) -> float:
        """Arc length of the closed curve (scatterer.cpp:484-490)."""
        if self._total < 0.0:
            self._total = self.spline.arc_length(0.0, float(self.t_vert[-1]))
        return self._total

    def panelize(self, npanels: int) -> List[ParametrizedPolynomial]:
        """npanels panels of equal arc length (scatterer.cpp:492-518): the end
        t1 of each panel by Newton steps t1 += (L/n - length(t0, t1)) / |z'(t1)|
        from t0 + (1 - t0) / (panels left), until the relative step is below
        1e-8; the last panel ends at t = 1."""
        total = self.total_length()
        panel_length = total / npanels
        t0 = 0.0
        panels: List[ParametrizedPolynomial] = []
        for i in range(1, npanels):
            t1 = t0 + (1.0 - t0) / (npanels - i + 1)
            L = 0.0
            for _ in range(self.MAX_ITER):
                L = self.spline.arc_length(t0, t1)
                t2 = t1 + (panel_length - L) / abs(complex(self.spline.eval_der(t1)))
                if abs(t2 - t1) / t1 < self.TOL:
                    break
                t1 = t2
            panels.append(ParametrizedPolynomial(self.spline, t0, t1, L))
            t0 = t1
            total -= L
        panels.append(ParametrizedPolynomial(self.spline, t0, 1.0, total))
        return panels


def mesh_from_file(fname: str, N: int, refinement_factor: Optional[float] = None, scale: float = 1.0,
                   smooth: bool = False) -> ParametrizedMesh:
    """A ParametrizedMesh of a scatterer file: exactly N uniform-per-edge panels
    (make_N_polygonal_panels), with a refinement factor graded panels
    (make_polygonal_panels), or with smooth=True -- or a file name ending in
    '@', the convention of examples/plot_solution_polygon.cpp:64-109 -- N
    equal-length panels of the smooth curve through the points."""
    if fname.endswith("@"):
        fname, smooth = fname[:-1], True
    if smooth:
        return ParametrizedMesh(SmoothScatterer(fname, scale).panelize(N))
    xy = read(fname, scale)
    if xy is None:
        raise ValueError("cannot read scatterer file " + fname)
    x, y = xy
    panels = make_N_polygonal_panels(x, y, N) if refinement_factor is None else \
        make_polygonal_panels(x, y, N, refinement_factor)
    return ParametrizedMesh(panels)
