"""Smooth scatterers: complex Akima splines (cspline.cpp), spline panels
(parametrized_polynomial.cpp) and SmoothScatterer::panelize (scatterer.cpp:
385-533) on the host, the device and the oracle.

The reference builds these on GSL's Akima splines and boost's Gauss-Kronrod
rule (GSL 2.x and boost are absent here), so the spline is pinned by what
Akima's method fixes: scipy's independent Akima1DInterpolator (same method,
same end extrapolation) for open data and for the periodic continuation,
interpolation of the data, C1 continuity across knots and across the
period, exactness on straight lines; arc lengths against adaptive
quadrature; the panels by their equal lengths and contiguity.  The Galerkin
integrals on spline panels are pinned by the same textbook tensor-Gauss
integrals as the other panel types, on panels that are one cubic piece each,
and the whole path by the analytic dielectric disk through a spline of the
reference's circle data.  The C++ facade header reproduces the Python code.
"""
import math
import os
import subprocess

import numpy as np
import pytest
from scipy import integrate, interpolate

import helmholtzbem_amd as HB
from helmholtzbem_amd import incoming as I
from helmholtzbem_amd import scatterer as S
from helmholtzbem_amd.geometry import ParametrizedPolynomial
from helmholtzbem_amd.spline import ComplexSpline

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "scatterers")


def path(name):
    return os.path.join(DATA, "scatterer_%s.dat" % name)


def test_open_spline_matches_scipy_akima():
    rng = np.random.default_rng(5)
    x = np.cumsum(rng.uniform(0.1, 1.0, 14))
    y = rng.normal(size=14) + 1j * rng.normal(size=14)
    y[5] = y[6] = y[7]                     # two equal secants (weights zero on one side only)
    s = ComplexSpline(x, y)
    t = np.linspace(x[0], x[-1], 997)
    ref = interpolate.Akima1DInterpolator(x, y.real)(t) + 1j * interpolate.Akima1DInterpolator(x, y.imag)(t)
    assert np.abs(s.eval(t) - ref).max() <= 1e-13
    dref = interpolate.Akima1DInterpolator(x, y.real)(t, 1) + 1j * interpolate.Akima1DInterpolator(x, y.imag)(t, 1)
    assert np.abs(s.eval_der(t) - dref).max() <= 1e-12 * np.abs(dref).max()


def test_periodic_spline_is_the_periodic_continuation():
    th = np.sort(np.random.default_rng(2).uniform(0, 2 * np.pi, 11))
    z = np.exp(1j * th) * (1 + 0.2 * np.cos(3 * th))
    t = np.concatenate([th, [th[0] + 2 * np.pi]]) / (2 * np.pi)
    zz = np.concatenate([z, [z[0]]])
    s = ComplexSpline(t, zz, periodic=True)
    # scipy Akima on three periods, evaluated on the middle one
    t3 = np.concatenate([t[:-1] - 1, t[:-1], t[:-1] + 1, [t[0] + 2]])
    z3 = np.concatenate([zz[:-1]] * 3 + [zz[:1]])
    q = np.linspace(t[0], t[-1], 801)
    ref = interpolate.Akima1DInterpolator(t3, z3.real)(q) + 1j * interpolate.Akima1DInterpolator(t3, z3.imag)(q)
    assert np.abs(s.eval(q) - ref).max() <= 1e-13
    # C1 across the period and across every knot
    assert abs(s.eval_der(t[0]) - s.eval_der(t[-1])) <= 1e-12 * abs(s.eval_der(t[0]))
    assert np.abs(s.eval(t) - zz).max() <= 1e-15
    h = 1e-7
    for k in t[1:-1]:
        assert abs(s.eval_der(k + h) - s.eval_der(k - h)) <= 1e-5 * abs(s.eval_der(k))


def test_spline_exact_on_lines_and_arc_length():
    x = np.linspace(0.0, 1.0, 9)
    s = ComplexSpline(x, (2 - 1j) * x + 0.5j)
    t = np.linspace(0, 1, 333)
    assert np.abs(s.eval(t) - ((2 - 1j) * t + 0.5j)).max() <= 1e-15
    assert abs(s.arc_length(0.1, 0.85) - 0.75 * math.sqrt(5)) <= 1e-14
    y = np.exp(2j * x) + 0.3 * x ** 2
    s = ComplexSpline(x, y)
    ref = integrate.quad(lambda u: abs(complex(s.eval_der(u))), 0.13, 0.91, points=list(x[1:-1]), epsabs=1e-14)[0]
    assert abs(s.arc_length(0.13, 0.91) - ref) <= 1e-13


def test_derivative_errors():
    s = ComplexSpline(np.linspace(0, 1, 6), np.arange(6.0))
    with pytest.raises(RuntimeError, match="Order of the spline derivative must be 0, 1 or 2"):
        s.eval_der(0.5, 3)


@pytest.mark.parametrize("name,N", [("circle", 40), ("flower", 100), ("distorted_circle", 64), ("wavy_torus", 90)])
def test_panelize_equal_lengths(name, N):
    sc = S.SmoothScatterer(path(name))
    ps = sc.panelize(N)
    L = sc.total_length()
    assert len(ps) == N and ps[0].tmin == 0.0 and ps[-1].tmax == 1.0
    for a, b in zip(ps[:-1], ps[1:]):
        assert a.tmax == b.tmin
    lens = np.array([p.length() for p in ps])
    # Newton stops at a relative step of 1e-8 in t; the last panel takes the rest
    assert np.abs(lens[:-1] - L / N).max() <= 2e-6 * L / N
    assert abs(lens[-1] - L / N) <= 2e-5 * L / N and abs(lens.sum() - L) <= 1e-13 * L
    for p in ps[::7]:
        assert abs(p.length() - sc.spline.arc_length(p.tmin, p.tmax)) <= 1e-12 * p.length()
    if name == "circle":
        assert abs(L - math.pi) <= 2e-6          # the data are 99 points of the circle of radius 0.5


def test_panel_geometry_and_mesh_packing():
    sc = S.SmoothScatterer(path("distorted_circle"))
    ps = sc.panelize(30)
    mesh = HB.ParametrizedMesh(ps)
    assert not mesh.isPolygonal()
    assert mesh.packed.shape == (30, 8) and np.all(mesh.packed[:, 0] == 3)
    assert mesh.coefs.shape == (3 * sc.spline.pieces + 1, 3)        # one shared table
    p = ps[4]
    a, b = p[0.0], p[1.0]
    assert np.allclose(a, ps[3][1.0], rtol=0, atol=1e-15)              # contiguous panels
    h = 1e-6
    fd = (np.array(p[0.5 + h]) - np.array(p[0.5 - h])) / (2 * h)
    assert np.abs(fd - np.array(p.Derivative_01(0.5))).max() <= 1e-7
    assert np.allclose(p(-1.0), a, atol=1e-15) and np.allclose(p(1.0), b, atol=1e-15)


def knot_mesh(H):
    """A closed spline whose panels are one cubic piece each (smooth integrands
    for the textbook tensor rules)."""
    th = np.linspace(0, 2 * np.pi, 11)
    z = np.exp(1j * th) * (0.6 + 0.15 * np.cos(2 * th))
    z[-1] = z[0]
    t = th / (2 * np.pi)
    s = ComplexSpline(t, z, periodic=True)
    return HB.ParametrizedMesh([ParametrizedPolynomial(s, t[j], t[j + 1]) for j in range(10)])


def test_oracle_general_and_adjacent_pairs_on_spline_panels(O, H):
    import test_oracle_assembly as TA
    mesh = knot_mesh(H)
    meshes = {"spline": mesh}
    orig = TA._meshes
    TA._meshes = lambda _H: meshes
    try:
        for d in (0, 2):
            TA._general_vs_textbook(O, H, "spline", 5, 1, d, TA.K0)
            TA._general_vs_textbook(O, H, "spline", 7, 2, d, 2.0 + 0.5j)
        got = TA.oracle_local(O, mesh, 12, 4, 3, 1)
        exact = TA._adjacent_exact(O, mesh, 4, 3, 1)
        for key in ("V", "W", "Kd", "Kt"):
            assert np.abs(got[key] - exact[key]).max() <= 1e-7 * np.abs(exact[key]).max(), key
    finally:
        TA._meshes = orig


def _disk(R, k, c_o, c_i, th, nmax=40):
    from scipy import special as sp
    ko, ki = math.sqrt(c_o) * k, math.sqrt(c_i) * k
    ud = np.zeros(th.shape, complex)
    un = np.zeros(th.shape, complex)
    for n in range(-nmax, nmax + 1):
        t = 1j ** n * (2j / (math.pi * R)) / (ko * sp.jv(n, ki * R) * sp.h1vp(n, ko * R)
                                            - ki * sp.jvp(n, ki * R) * sp.hankel1(n, ko * R))
        ud += t * sp.jv(n, ki * R) * np.exp(1j * n * th)
        un += t * ki * sp.jvp(n, ki * R) * np.exp(1j * n * th)
    return ud, un


def _disk_case(mesh):
    space = HB.ContinuousSpace1()
    u, g = I.as_functions([I.wavespec_t(I.WaveType.Plane, 1.0)], 2.0)
    U = np.concatenate([space.Interpolate_helmholtz(u, mesh), space.Interpolate_helmholtz_neu(g, mesh)])[:, None]
    th = np.array([math.atan2(p[0.0][1], p[0.0][0]) for p in mesh.getPanels()])
    return U, _disk(0.5, 2.0, 1.0, 4.0, th)


def test_oracle_transmission_on_smooth_circle_converges(O):
    sc = S.SmoothScatterer(path("circle"))
    errs = []
    for N in (16, 32, 64):
        mesh = HB.ParametrizedMesh(sc.panelize(N))
        U, (ud, un) = _disk_case(mesh)
        x = O.tp_solve(mesh, 11, 2.0, 1.0, 4.0, U)[:, 0]
        errs.append((np.linalg.norm(x[:N] - ud) / np.linalg.norm(ud), np.linalg.norm(x[N:] - un) / np.linalg.norm(un)))
    errs = np.array(errs)
    assert errs[-1, 0] <= 1e-5 and errs[-1, 1] <= 6e-5, errs
    assert np.all(errs[:-1, 0] / errs[1:, 0] >= 3.5), errs


def test_cpp_smooth_scatterer_matches_python(tmp_path):
    exe = str(tmp_path / "scatterer_dump")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "scatterer_dump.cpp"), "-o", exe],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    for name, N in (("flower", 64), ("circle", 30)):
        out = subprocess.run([exe, path(name), str(N), "2"], check=True, stdout=subprocess.PIPE,
                             text=True).stdout.strip().split("\n")
        sc = S.SmoothScatterer(path(name))
        ps = sc.panelize(N)
        n, L = out[0].split()
        assert int(n) == N and abs(float(L) - sc.total_length()) <= 1e-14 * sc.total_length()
        got = np.array([[float(v) for v in ln.split()] for ln in out[1:1 + N]])
        want = np.array([[p.tmin, p.tmax, p.length()] for p in ps])
        assert np.abs(got - want).max() <= 1e-12, name
        rows = np.array([[float(v) for v in ln.split()] for ln in out[1 + N:]])
        assert np.array_equal(rows, np.array(sc.spline.rows())), name      # bit for bit


# ---------------------------------------------------------------------------
# Device
# ---------------------------------------------------------------------------
@pytest.mark.gpu
def test_device_assembly_on_smooth_scatterer(H, O):
    mesh = S.mesh_from_file(path("flower") + "@", 60)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 10)
    gb = H.GalerkinBuilder(bd)
    for k in (3.0, 2.0 + 0.5j):
        gb.assembleDense(k, 4.0, H.LayerType.ALL, 1)
        K, W, V = O.assemble(mesh, 10, k, 4.0, d=1)
        for t in (0, 1):
            for got, ref in ((gb.getDoubleLayer(t), K[t]), (gb.getHypersingular(t), W[t]), (gb.getSingleLayer(t), V[t])):
                assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    R = H.gaussian_test_matrix(120, 2, 0)
    ks = np.linspace(2.0, 4.0, 5)
    sig = H.sigma_sweep(bd, ks, 1.0, 9.0, R, 2)
    sig_o = O.SolOp(mesh, 10).sigma_sweep(ks, 1.0, 9.0, R, 2)
    assert np.abs(sig - sig_o).max() <= 1e-10 * np.abs(sig_o).max()


@pytest.mark.gpu
def test_device_transmission_on_smooth_circle(O):
    from helmholtzbem_amd.solvers import tp
    mesh = HB.ParametrizedMesh(S.SmoothScatterer(path("circle")).panelize(64))
    u, g = I.as_functions([I.wavespec_t(I.WaveType.Plane, 1.0)], 2.0)
    x = tp.direct_second_kind.solve(mesh, u, g, 11, 2.0, 1.0, 4.0)
    U, (ud, un) = _disk_case(mesh)
    assert np.linalg.norm(x[:64] - ud) <= 1e-5 * np.linalg.norm(ud)
    A, rhs = O.tp_A_rhs(mesh, 11, 2.0, 1.0, 4.0, U)
    xo = np.linalg.solve(A, rhs)[:, 0]
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.cond(A) * np.linalg.norm(xo)


def test_corner_between_straight_runs_follows_gsl():
    """Both Akima weights vanish at a node where two exactly straight runs
    meet.  GSL's akima (the reference's gsl_interp_akima, cspline.cpp:24-25)
    then makes the piece from that node linear with its secant slope and ends
    the piece before it with that piece's own secant slope, so both runs stay
    exactly straight and the corner stays a corner."""
    x = np.arange(11.0)
    y = np.where(x <= 5.0, 0.0, x - 5.0) + 1j * np.where(x <= 5.0, 2.0 * x, 10.0)
    s = ComplexSpline(x, y)
    t = np.linspace(2.0, 8.0, 601)
    poly = np.where(t <= 5.0, 0.0, t - 5.0) + 1j * np.where(t <= 5.0, 2.0 * t, 10.0)
    assert np.abs(s.eval(t) - poly).max() <= 1e-14
    assert s.eval_der(np.array([4.999]))[0] == 0.0 + 2.0j
    assert s.eval_der(np.array([5.0]))[0] == 1.0 + 0.0j
    # the mean-slope convention (scipy) rounds the corner instead
    ref = interpolate.Akima1DInterpolator(x, y.real)(4.5)
    assert abs(ref) > 1e-3 and s.eval(np.array([4.5]))[0].real == 0.0
