"""Panel geometry and meshes (host side).

Mirrors the reference's curve classes closely enough that meshes built here are
bit-identical to the reference's:

* ``ParametrizedLine`` -- parametrized_line.cpp:20-144 (``split`` evaluates the
  [-1,1] form ``0.5*(t*(x2-x1)+(x2+x1))`` at ``t = -1 + 2i/N``, last point exact).
* ``ParametrizedCircularArc`` -- parametrized_circular_arc.cpp:21-152.
* ``ParametrizedTrigCurve`` -- an analytic curve ``z(t) = sum c_m e^{imt}`` (a
  complex trigonometric polynomial) split into equal-parameter panels.  The
  reference's own smooth-curve class (parametrized_fourier_sum.hpp) is declared
  but has no evaluation methods, so smooth scatterers use this type.
* ``ParametrizedMesh`` -- parametrized_mesh.cpp:13-81.

* ``ParametrizedPolynomial`` -- a piece [tmin, tmax] of a complex Akima spline
  (parametrized_polynomial.cpp:21-173, the panels of smooth scatterers).

Panels are packed as 8 doubles ``{type, a0..a6}``:
line ``{0, x1, y1, x2, y2}``, arc ``{1, cx, cy, r, phi0, phi1}``,
trig-curve segment ``{2, t0, t1, first_coef, n_coefs}`` with a coefficient table of
``{m, re, im}`` triples, spline piece ``{3, tmin, tmax, first_row, n_pieces}`` with
the spline's rows in the same table (helmholtzbem_amd/spline.py).  This is the
layout the C-ABI (include/hbem.h) takes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

PANEL_LINE = 0
PANEL_ARC = 1
PANEL_TRIG = 2
PANEL_SPLINE = 3


@dataclass
class ParametrizedLine:
    """Straight panel from ``start`` to ``end`` (parametrized_line.cpp:20-21)."""
    start: Sequence[float]
    end: Sequence[float]

    def __call__(self, t: float):
        """Point at t in [-1, 1] (parametrized_line.cpp:23-29)."""
        x1, y1 = self.start
        x2, y2 = self.end
        return (0.5 * (t * (x2 - x1) + (x2 + x1)), 0.5 * (t * (y2 - y1) + (y2 + y1)))

    def __getitem__(self, s: float):
        """Point at s in [0, 1] (operator[], parametrized_line.cpp:36-42)."""
        x1, y1 = self.start
        x2, y2 = self.end
        return (s * (x2 - x1) + x1, s * (y2 - y1) + y1)

    def Derivative_01(self, s: float):
        """d/ds of operator[] (parametrized_line.cpp:78-84)."""
        return (self.end[0] - self.start[0], self.end[1] - self.start[1])

    def split(self, n: int) -> List["ParametrizedLine"]:
        """parametrized_line.cpp:116-137."""
        parts = []
        for i in range(n):
            t1 = -1.0 + i * 2.0 / n
            t2 = -1.0 + (i + 1) * 2.0 / n
            if i == n - 1:
                t2 = 1.0
            parts.append(ParametrizedLine(self(t1), self(t2)))
        return parts

    def length(self) -> float:
        a, b = self(-1.0), self(1.0)
        return math.hypot(b[0] - a[0], b[1] - a[1])

    def pack(self, coef_offset: int = 0):
        return [PANEL_LINE, float(self.start[0]), float(self.start[1]), float(self.end[0]), float(self.end[1]), 0.0, 0.0, 0.0]

    is_line = True


@dataclass
class ParametrizedCircularArc:
    """Circular arc (parametrized_circular_arc.cpp:21-28)."""
    center: Sequence[float] = (0.0, 0.0)
    radius: float = 0.25
    phi_start: float = 0.0
    phi_end: float = 2 * math.pi

    def __getitem__(self, s: float):
        """Point at s in [0, 1] (operator[], parametrized_circular_arc.cpp:46-53)."""
        a = s * (self.phi_end - self.phi_start) + self.phi_start
        return (self.center[0] + self.radius * math.cos(a), self.center[1] + self.radius * math.sin(a))

    def Derivative_01(self, s: float):
        """d/ds of operator[] (parametrized_circular_arc.cpp:92-99)."""
        d = self.phi_end - self.phi_start
        a = s * d + self.phi_start
        return (-self.radius * d * math.sin(a), self.radius * d * math.cos(a))

    def split(self, n: int) -> List["ParametrizedCircularArc"]:
        """parametrized_circular_arc.cpp:134-148."""
        parts = []
        for i in range(n):
            a = self.phi_start + i * (self.phi_end - self.phi_start) / n
            b = self.phi_start + (i + 1) * (self.phi_end - self.phi_start) / n
            if i == n - 1:
                b = self.phi_end
            parts.append(ParametrizedCircularArc(tuple(self.center), self.radius, a, b))
        return parts

    def length(self) -> float:
        return self.radius * (self.phi_end - self.phi_start)

    def pack(self, coef_offset: int = 0):
        return [PANEL_ARC, float(self.center[0]), float(self.center[1]), float(self.radius),
                float(self.phi_start), float(self.phi_end), 0.0, 0.0]

    is_line = False


@dataclass
class ParametrizedTrigCurve:
    """Analytic curve segment z(t) = sum_m c_m exp(i m t), t in [t0, t1].

    ``coeffs`` is a list of ``(m, complex c_m)``.
    """
    coeffs: Sequence
    t0: float = 0.0
    t1: float = 2 * math.pi

    def split(self, n: int) -> List["ParametrizedTrigCurve"]:
        parts = []
        for i in range(n):
            a = self.t0 + i * (self.t1 - self.t0) / n
            b = self.t0 + (i + 1) * (self.t1 - self.t0) / n
            if i == n - 1:
                b = self.t1
            parts.append(ParametrizedTrigCurve(self.coeffs, a, b))
        return parts

    def z(self, t):
        return sum(complex(c) * np.exp(1j * m * t) for m, c in self.coeffs)

    def __getitem__(self, s: float):
        """Point at s in [0, 1]: z(t0 + s (t1 - t0))."""
        v = complex(self.z(self.t0 + s * (self.t1 - self.t0)))
        return (v.real, v.imag)

    def Derivative_01(self, s: float):
        """d/ds of the point at s in [0, 1]."""
        t = self.t0 + s * (self.t1 - self.t0)
        v = (self.t1 - self.t0) * sum(1j * m * complex(c) * np.exp(1j * m * t) for m, c in self.coeffs)
        return (float(np.real(v)), float(np.imag(v)))

    def length(self, n=64) -> float:
        t = np.linspace(self.t0, self.t1, n)
        z = self.z(t)
        return float(np.sum(np.abs(np.diff(z))))

    def pack(self, coef_offset: int = 0):
        return [PANEL_TRIG, float(self.t0), float(self.t1), float(coef_offset), float(len(self.coeffs)), 0.0, 0.0, 0.0]

    is_line = False


class ParametrizedPolynomial:
    """Piece [tmin, tmax] of a complex spline z(t) (parametrized_polynomial.cpp:21-173).

    The derivatives carry the chain-rule factor of the parameter map:
    Derivative_01(s) = (tmax - tmin) z'(t).  The reference returns z'(t)
    there (parametrized_polynomial.cpp:91-97) and positions from its
    vectorised Derivative_01_swapped (132-144); both are not what the
    Galerkin integrals need, so the intended derivative is used.
    """

    def __init__(self, spline, tmin: float, tmax: float, length: float = -1.0):
        self.spline, self.tmin, self.tmax = spline, float(tmin), float(tmax)
        self._length = float(length) if length >= 0.0 else spline.arc_length(self.tmin, self.tmax)

    def _t01(self, s):
        return s * (self.tmax - self.tmin) + self.tmin

    def __call__(self, t: float):
        """Point at t in [-1, 1] (parametrized_polynomial.cpp:25-32)."""
        v = complex(self.spline.eval(t * (self.tmax - self.tmin) / 2 + (self.tmax + self.tmin) / 2))
        return (v.real, v.imag)

    def __getitem__(self, s: float):
        v = complex(self.spline.eval(self._t01(s)))
        return (v.real, v.imag)

    def Derivative_01(self, s: float):
        v = (self.tmax - self.tmin) * complex(self.spline.eval_der(self._t01(s)))
        return (v.real, v.imag)

    def split(self, n: int) -> List["ParametrizedPolynomial"]:
        """parametrized_polynomial.cpp:154-169."""
        parts = []
        for i in range(n):
            a = self.tmin + i * (self.tmax - self.tmin) / n
            b = self.tmin + (i + 1) * (self.tmax - self.tmin) / n
            if i == n - 1:
                b = self.tmax
            parts.append(ParametrizedPolynomial(self.spline, a, b))
        return parts

    def length(self) -> float:
        return self._length

    def pack(self, coef_offset: int = 0):
        return [PANEL_SPLINE, self.tmin, self.tmax, float(coef_offset), float(self.spline.pieces), 0.0, 0.0, 0.0]

    is_line = False


class ParametrizedMesh:
    """Ordered panel list (parametrized_mesh.cpp:13-31) packed for the C-ABI."""

    def __init__(self, panels):
        self.panels = list(panels)
        coef_rows = []
        rows = []
        offsets = {}
        for p in self.panels:
            off = 0
            if isinstance(p, ParametrizedTrigCurve):
                key = id(p.coeffs)
                if key not in offsets:
                    offsets[key] = len(coef_rows)
                    for m, c in p.coeffs:
                        c = complex(c)
                        coef_rows.append([float(m), c.real, c.imag])
                off = offsets[key]
            elif isinstance(p, ParametrizedPolynomial):
                key = id(p.spline)
                if key not in offsets:
                    offsets[key] = len(coef_rows)
                    coef_rows.extend(p.spline.rows())
                off = offsets[key]
            rows.append(p.pack(off))
        self.packed = np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(-1, 8))
        self.coefs = np.ascontiguousarray(np.array(coef_rows, dtype=np.float64).reshape(-1, 3))

    def getNumPanels(self) -> int:
        return len(self.panels)

    def getPanels(self):
        return self.panels

    def isPolygonal(self) -> bool:
        """parametrized_mesh.cpp:70-81 (MESH_TYPE_GENERAL)."""
        return all(getattr(p, "is_line", False) for p in self.panels)


# ---------------------------------------------------------------------------
# Workload geometries (seeded; splitmix64-seeded xoshiro256** generator)
# ---------------------------------------------------------------------------
class Xoshiro256ss:
    """xoshiro256** seeded by splitmix64; doubles = (u >> 11) * 2^-53."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        x = seed & self.MASK
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & self.MASK
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
            s.append(z ^ (z >> 31))
        self.s = s

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & Xoshiro256ss.MASK

    def next_u64(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & self.MASK, 7) * 9) & self.MASK
        t = (s[1] << 17) & self.MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self, a=0.0, b=1.0) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo

    def gauss(self) -> float:
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def gaussian_test_matrix(n: int, nc: int, seed: int) -> np.ndarray:
    """Complex Gaussian n x nc test matrix (g1 + i g2)/sqrt(2), column-major fill."""
    rng = Xoshiro256ss(seed)
    R = np.empty((n, nc), dtype=np.complex128)
    for j in range(nc):
        for i in range(n):
            R[i, j] = complex(rng.gauss(), rng.gauss()) * math.sqrt(0.5)
    return np.asfortranarray(R)


def square_mesh(side: float = 0.5, per_side: int = 12, center=(0.0, 0.0)) -> ParametrizedMesh:
    """Square of the given side, CCW from the lower-left corner, each side split
    into ``per_side`` ParametrizedLine panels (roots_brent_square_rsvd.cpp:79-105)."""
    h = 0.5 * side
    cx, cy = center
    c = [(cx - h, cy - h), (cx + h, cy - h), (cx + h, cy + h), (cx - h, cy + h)]
    panels = []
    for a in range(4):
        panels += ParametrizedLine(c[a], c[(a + 1) % 4]).split(per_side)
    return ParametrizedMesh(panels)


def star_curve(amplitude: float = 0.25, lobes: int = 5) -> ParametrizedTrigCurve:
    """r(t) = 1 + a cos(l t):  z(t) = e^{it} + (a/2)(e^{i(l+1)t} + e^{-i(l-1)t})."""
    return ParametrizedTrigCurve([(1, 1.0), (lobes + 1, 0.5 * amplitude), (-(lobes - 1), 0.5 * amplitude)])


def star_mesh(npanels: int = 400, amplitude: float = 0.25, lobes: int = 5) -> ParametrizedMesh:
    return ParametrizedMesh(star_curve(amplitude, lobes).split(npanels))


def circle_mesh(npanels: int, radius: float = 0.25, center=(0.0, 0.0)) -> ParametrizedMesh:
    return ParametrizedMesh(ParametrizedCircularArc(center, radius, 0.0, 2 * math.pi).split(npanels))


def random_polygon_vertices(rng: Xoshiro256ss, nvert: int, rmin=0.5, rmax=1.5):
    ang = sorted(rng.uniform(0.0, 2 * math.pi) for _ in range(nvert))
    rad = [rng.uniform(rmin, rmax) for _ in range(nvert)]
    return [(r * math.cos(a), r * math.sin(a)) for r, a in zip(rad, ang)]


def polygon_mesh(vertices, npanels: int) -> ParametrizedMesh:
    """Split each edge into panels in proportion to its length (floor, then the
    largest remainders; at least one panel per edge)."""
    m = len(vertices)
    lens = [math.hypot(vertices[(a + 1) % m][0] - vertices[a][0], vertices[(a + 1) % m][1] - vertices[a][1])
            for a in range(m)]
    total = sum(lens)
    share = [npanels * l / total for l in lens]
    counts = [max(1, int(math.floor(s))) for s in share]
    rem = npanels - sum(counts)
    order = sorted(range(m), key=lambda a: -(share[a] - math.floor(share[a])))
    k = 0
    while rem > 0:
        counts[order[k % m]] += 1
        rem -= 1
        k += 1
    while rem < 0:
        a = max(range(m), key=lambda a: counts[a])
        counts[a] -= 1
        rem += 1
    panels = []
    for a in range(m):
        panels += ParametrizedLine(vertices[a], vertices[(a + 1) % m]).split(counts[a])
    return ParametrizedMesh(panels)


def random_polygon_mesh(seed: int, nvert: int, npanels: int) -> ParametrizedMesh:
    rng = Xoshiro256ss(seed)
    return polygon_mesh(random_polygon_vertices(rng, nvert), npanels)


def random_polygon_instances(seed: int = 2, count: int = 64, nk: int = 32, k_min: float = 2.0, k_max: float = 20.0,
                             nmin: int = 100, nmax: int = 1000, vmin: int = 6, vmax: int = 16,
                             ci_min: float = 4.0, ci_max: float = 25.0):
    """The many-scatterer workload: `count` random star-shaped polygons from
    one xoshiro256** stream (per instance, in this order: the vertex count m
    in [vmin, vmax], m sorted angles and m radii as random_polygon_vertices
    draws them, the panel count N in [nmin, nmax], c_i uniform in
    [ci_min, ci_max], then nk wavenumbers uniform in [k_min, k_max], sorted).
    Returns a list of (mesh, c_i, ks)."""
    rng = Xoshiro256ss(seed)
    out = []
    for _ in range(count):
        m = rng.randint(vmin, vmax)
        verts = random_polygon_vertices(rng, m)
        N = rng.randint(nmin, nmax)
        c_i = rng.uniform(ci_min, ci_max)
        ks = np.sort(np.array([rng.uniform(k_min, k_max) for _ in range(nk)]))
        out.append((polygon_mesh(verts, N), c_i, ks))
    return out
