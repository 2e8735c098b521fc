"""Transmission solve tp::direct_second_kind (solvers.cpp:104-152) and the
incoming waves it takes (incoming.cpp).

CPU: the incoming-wave library against the reference's data files, the
loader's rules, the Helmholtz equation, finite-difference gradients, mpmath
Bessel values and an independent quadrature of the Herglotz wave; the P1
interpolation's points and normals; and the oracle's composition of A_rhs /
solve against
  * the analytic dielectric disk (Mie series, derived independently below)
    with c_o = 1 and c_o != 1, converging under mesh refinement, and
  * the medium without a scatterer (c_i = c_o, where the total field is the
    incoming one) on a 5-lobed star, which pins the K_o^T block of A_o: the
    same system with K_o there stalls at a 40% error.
GPU: hbem_tp_A_rhs / hbem_tp_solve against the oracle (operators <= 1e-12
relative Frobenius; traces <= 1e-12 kappa(A) relative, the first-order bound
for systems that agree to 1e-12), several waves per call against single
calls, complex k, and the reference's error for a non-finite right-hand side.
"""
import math
import os

import numpy as np
import pytest
from scipy import integrate
from scipy import special as sp

import helmholtzbem_amd as HB
from helmholtzbem_amd import incoming as I
from helmholtzbem_amd.solvers import tp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "incoming")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------------------
# incoming waves
# ---------------------------------------------------------------------------
def test_reference_wave_files_load():
    w = I.load(os.path.join(GOLD, "incoming_circularJ.dat"))
    assert [p.order for p in w] == [-1, -2, 0, 1, 2]
    assert all(p.type == I.WaveType.CircularJ and p.amplitude == 1.0 and p.x0 == (0.0, 0.0) for p in w)
    (h,) = I.load(os.path.join(GOLD, "incoming_herglotz.dat"))
    assert (h.type, h.amplitude, h.x0, h.angle, h.eps) == (I.WaveType.Herglotz, 1.0, (-0.2, 0.0), 0.85, 0.1)
    (p,) = I.load(os.path.join(GOLD, "incoming_plane_from_right.dat"))
    assert (p.type, p.x0, p.angle) == (I.WaveType.Plane, (1.0, 0.0), 3.14159)
    for f in ("incoming_herglotz_2.dat", "incoming_plane_at_60deg.dat", "incoming_plane_from_below.dat"):
        assert len(I.load(os.path.join(GOLD, f))) == 1


@pytest.mark.parametrize("text,ok", [
    ("Plane, 1.0, 0, 0, 1.047\n", True),
    ("# comment\n\n  Plane , 2 ,0,0 , 0.5 ,\n", True),          # spaces removed, final comma
    ("Plane, 1.0, 0, 0, 1.047,,\n", False),                      # an empty extra field
    ("Plane, 1.0, 0, 0\n", False),                               # missing field
    ("Plane, 1.0, 0, 0, 1, 2\n", False),                         # extra field
    ("Wave, 1.0, 0, 0, 1\n", False),                             # unknown name
    ("Plane, abc, 0, 0, 1\n", False),                            # no number
    ("Plane, 1.5x, 0, 0, 1\n", True),                            # stod takes the prefix
    ("Plane, 1e999, 0, 0, 1\n", False),                          # out of range
    ("CircularJ, 1, 0, 0, 2.7\n", True),                         # stoi takes the prefix
    ("CircularY, 1, 0, 0, 99999999999\n", False),
    ("Herglotz, 1, 0, 0, 0.5\n", False),                         # Herglotz needs eps
])
def test_loader_rules(tmp_path, text, ok):
    f = tmp_path / "w.dat"
    f.write_text(text)
    w = I.load(str(f))
    assert (w is not None) == ok
    if text.startswith("Plane, 1.5x"):
        assert w[0].amplitude == 1.5
    if text.startswith("CircularJ"):
        assert w[0].order == 2


def test_missing_file_loads_empty(tmp_path):
    assert I.load(str(tmp_path / "none.dat")) == []


WAVES = [
    I.wavespec_t(I.WaveType.Plane, 1.3, angle=1.047, x0=(0.2, -0.1)),
    I.wavespec_t(I.WaveType.CircularJ, 0.7, order=-2, x0=(0.1, 0.3)),
    I.wavespec_t(I.WaveType.CircularJ, 1.0, order=1),
    I.wavespec_t(I.WaveType.CircularY, 0.5, order=3, x0=(-0.4, 0.2)),
    I.wavespec_t(I.WaveType.FourierHankel1, 1.0, order=-1, x0=(0.5, 0.5)),
    I.wavespec_t(I.WaveType.FourierHankel2, 1.0, order=2, x0=(-0.5, 0.1)),
    I.wavespec_t(I.WaveType.Herglotz, 1.0, angle=0.85, eps=0.1, x0=(-0.2, 0.0)),
]


@pytest.mark.parametrize("k", [3.0, 2.0 + 0.5j])
@pytest.mark.parametrize("wi", range(len(WAVES)))
def test_waves_solve_helmholtz_with_matching_gradient(k, wi):
    w = [WAVES[wi]]
    pts = np.array([[0.31, -0.22], [-0.13, 0.47], [0.05, 0.06]])
    h = 1e-4
    ex, ey = np.array([h, 0.0]), np.array([0.0, h])
    u = I.compute(w, pts, k)
    g = I.compute_del(w, pts, k)
    fd = np.stack([(I.compute(w, pts + ex, k) - I.compute(w, pts - ex, k)) / (2 * h),
                   (I.compute(w, pts + ey, k) - I.compute(w, pts - ey, k)) / (2 * h)], axis=-1)
    assert np.abs(g - fd).max() <= 1e-6 * max(1.0, np.abs(g).max())
    H2 = 2e-4
    hx, hy = np.array([H2, 0.0]), np.array([0.0, H2])
    lap = (I.compute(w, pts + hx, k) + I.compute(w, pts - hx, k) + I.compute(w, pts + hy, k)
           + I.compute(w, pts - hy, k) - 4 * u) / H2 ** 2
    assert np.abs(lap + k * k * u).max() <= 1e-4 * max(1.0, abs(k) ** 2 * np.abs(u).max())


def test_circular_waves_match_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    for (l, k, x) in [(-2, 2.5, (0.3, 0.4)), (3, 1.5 + 0.7j, (-0.6, 0.2)), (0, 4.0, (0.1, -0.9))]:
        r = math.hypot(*x)
        th = math.atan2(x[1], x[0])
        z = mp.mpc(k) * r
        J = complex(mp.besselj(l, z) * mp.expj(l * th))
        Y = complex(mp.bessely(l, z) * mp.expj(l * th))
        assert abs(I.circular_J(x, (0, 0), l, k) - J) <= 1e-13 * max(1.0, abs(J))
        assert abs(I.circular_Y(x, (0, 0), l, k) - Y) <= 1e-13 * max(1.0, abs(Y))
        assert abs(I.fourier_hankel(x, (0, 0), 1, l, k) - (J + 1j * Y)) <= 1e-13 * max(1.0, abs(J) + abs(Y))


def test_circular_J_gradient_at_its_center():
    """|l| = 1: the limit of the gradient at r = 0 (incoming.cpp:163-165)."""
    k = 2.0
    for l in (-1, 1, 0, 2):
        g0 = I.circular_J_del((0.0, 0.0), (0.0, 0.0), l, k)
        h = 1e-5
        fd = [(I.circular_J((h, 0), (0, 0), l, k) - I.circular_J((-h, 0), (0, 0), l, k)) / (2 * h),
              (I.circular_J((0, h), (0, 0), l, k) - I.circular_J((0, -h), (0, 0), l, k)) / (2 * h)]
        assert np.abs(g0 - np.array(fd)).max() <= 1e-6


def test_herglotz_is_the_mean_of_plane_waves():
    """Independent adaptive quadrature of (1 / 2 eps) int exp(i k p.d(t)) dt."""
    x, x0, a, eps, k = np.array([0.4, -0.3]), (-0.2, 0.0), 0.85, 0.1, 7.0
    p = x - np.array(x0)
    f = lambda t, part: getattr(np.exp(1j * k * (p[0] * np.cos(t) + p[1] * np.sin(t))), part)
    ref = (integrate.quad(f, a - eps, a + eps, args=("real",), epsabs=1e-15)[0]
           + 1j * integrate.quad(f, a - eps, a + eps, args=("imag",), epsabs=1e-15)[0]) / (2 * eps)
    assert abs(I.herglotz(x, x0, a, eps, k) - ref) <= 1e-13


# ---------------------------------------------------------------------------
# P1 interpolation (continuous_space.hpp:67-101)
# ---------------------------------------------------------------------------
def test_interpolation_at_panel_starts_with_outward_normals():
    N, R = 24, 0.7
    mesh = HB.circle_mesh(N, R)
    space = HB.ContinuousSpace1()
    phi = 2 * math.pi * np.arange(N) / N
    d = space.Interpolate_helmholtz(lambda x1, x2: x1 + 2j * x2, mesh)
    assert np.abs(d - R * (np.cos(phi) + 2j * np.sin(phi))).max() <= 1e-15
    # gradient (1, 2i): n = (cos phi, sin phi) on the CCW circle
    nrm = space.Interpolate_helmholtz_neu(lambda x1, x2: np.array([1.0, 2j]), mesh)
    assert np.abs(nrm - (np.cos(phi) + 2j * np.sin(phi))).max() <= 1e-15
    sq = HB.square_mesh(1.0, 2)
    n_sq = space.Interpolate_helmholtz_neu(lambda x1, x2: np.array([1.0, 1j]), sq)
    # bottom, right, top, left edges: outward normals (0,-1), (1,0), (0,1), (-1,0)
    assert np.allclose(n_sq, [-1j, -1j, 1, 1, 1j, 1j, -1, -1], atol=1e-15)


# ---------------------------------------------------------------------------
# The oracle's composition against the physics
# ---------------------------------------------------------------------------
def disk_traces(R, k, c_o, c_i, th, nmax=30):
    """Total-field traces on |x| = R for the plane wave exp(i k_o x_1):
    inside sum_n t_n J_n(k_i r) e^{in theta} with, from continuity of u and
    du/dr at r = R and the Wronskian J_n H_n' - J_n' H_n = 2i / (pi z),
      t_n = i^n (2i / (pi R)) / (k_o J_n(k_i R) H_n'(k_o R) - k_i J_n'(k_i R) H_n(k_o R)),
    k_o = sqrt(c_o) k, k_i = sqrt(c_i) k."""
    ko, ki = math.sqrt(c_o) * k, math.sqrt(c_i) * k
    ud = np.zeros(th.shape, complex)
    un = np.zeros(th.shape, complex)
    for n in range(-nmax, nmax + 1):
        t = 1j ** n * (2j / (math.pi * R)) / (ko * sp.jv(n, ki * R) * sp.h1vp(n, ko * R)
                                            - ki * sp.jvp(n, ki * R) * sp.hankel1(n, ko * R))
        ud += t * sp.jv(n, ki * R) * np.exp(1j * n * th)
        un += t * ki * sp.jvp(n, ki * R) * np.exp(1j * n * th)
    return ud, un


def u_inc_N(mesh, waves, k_o):
    space = HB.ContinuousSpace1()
    cols = []
    for w in waves:
        u, g = I.as_functions(w, k_o)
        cols.append(np.concatenate([space.Interpolate_helmholtz(u, mesh), space.Interpolate_helmholtz_neu(g, mesh)]))
    return np.asfortranarray(np.stack(cols, axis=1))


PLANE_X = [I.wavespec_t(I.WaveType.Plane, 1.0, angle=0.0)]


@pytest.mark.parametrize("k,c_o,c_i", [(2.0, 1.0, 4.0), (3.0, 1.5, 9.0)])
def test_oracle_solve_converges_to_dielectric_disk(O, k, c_o, c_i):
    R = 0.25
    errs = []
    for N in (16, 32, 64):
        mesh = HB.circle_mesh(N, R)
        x = O.tp_solve(mesh, 11, k, c_o, c_i, u_inc_N(mesh, [PLANE_X], math.sqrt(c_o) * k))[:, 0]
        ud, un = disk_traces(R, k, c_o, c_i, 2 * math.pi * np.arange(N) / N)
        errs.append((rel(x[:N], ud), rel(x[N:], un)))
    errs = np.array(errs)
    assert errs[-1, 0] <= 5e-6 and errs[-1, 1] <= 2e-5, errs
    assert np.all(errs[:-1] / errs[1:] >= 6.0), errs          # O(h^3) at the nodes


def test_oracle_solve_without_scatterer_returns_the_incoming_traces(O):
    """c_i = c_o: A = blockdiag(M, M^T) and the solution is the interpolated
    incoming wave up to the discretisation error, which must fall as h^2."""
    errs = []
    w = [I.wavespec_t(I.WaveType.Plane, 1.0, angle=0.7)]
    for N in (64, 128, 256):
        mesh = HB.star_mesh(N, 0.25, 5)
        U = u_inc_N(mesh, [w], 2.0)
        x = O.tp_solve(mesh, 11, 2.0, 1.0, 1.0, U)[:, 0]
        errs.append((rel(x[:N], U[:N, 0]), rel(x[N:], U[N:, 0])))
    errs = np.array(errs)
    assert errs[-1, 0] <= 3e-4 and errs[-1, 1] <= 1.2e-3, errs
    assert np.all(errs[:-1] / errs[1:] >= 3.5), errs


def test_tp_entry_points_validate_before_device():
    import ctypes
    from helmholtzbem_amd import _lib
    L = _lib.lib()
    U = np.zeros(8, dtype=np.complex128)
    rc = L.hbem_tp_solve(None, 1.0, 0.0, 1.0, 4.0, ctypes.c_void_p(U.ctypes.data), 1, ctypes.c_void_p(U.ctypes.data))
    assert rc == _lib.HBEM_EINVAL
    rc = L.hbem_tp_A_rhs(None, 1.0, 0.0, 1.0, 4.0, None, 1, None, None)
    assert rc == _lib.HBEM_EINVAL


# ---------------------------------------------------------------------------
# Device against the oracle
# ---------------------------------------------------------------------------
def _wave_fns(waves, k_o):
    fns = [I.as_functions(w, k_o) for w in waves]
    return [f[0] for f in fns], [f[1] for f in fns]


@pytest.mark.gpu
@pytest.mark.parametrize("proj", [False, True])
def test_A_rhs_parity(O, proj):
    mesh = HB.square_mesh(0.5, 6)
    k, c_o, c_i = 2.0, 1.0, 20.0
    space = HB.ContinuousSpace1()
    bd = HB.BuilderData(mesh, space, space, 11)
    so = HB.SolutionsOperator(bd, proj)
    us, gs = _wave_fns([PLANE_X, WAVES[:2]], k)
    A, rhs = tp.direct_second_kind.A_rhs(mesh, us, gs, k, c_o, c_i, so)
    U = u_inc_N(mesh, [PLANE_X, WAVES[:2]], k)
    Ao, rhso = O.tp_A_rhs(mesh, 11, k, c_o, c_i, U, proj=proj)
    assert rel(A, Ao) <= 1e-12
    assert rel(rhs, rhso) <= 1e-12
    # single wave: a vector, like the reference's Eigen::VectorXcd
    A1, r1 = tp.direct_second_kind.A_rhs(mesh, us[0], gs[0], k, c_o, c_i, so)
    assert r1.shape == (48,) and rel(r1, rhso[:, 0]) <= 1e-12
    # the accessors hold the operators of the last call (gen_sol_op.hpp:120-125)
    K, W, V = O.assemble(mesh, 11, k, c_o)
    assert rel(so.K_ext(), K[0]) <= 1e-12 and rel(so.V_ext(), V[0]) <= 1e-12 and rel(so.W_ext(), W[0]) <= 1e-12
    Ki = O.assemble(mesh, 11, k, c_i)[0]
    assert rel(so.K_int(), Ki[0]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["disk", "star", "complex_k"])
def test_solve_parity(O, case):
    if case == "disk":
        mesh, k, c_o, c_i = HB.circle_mesh(64, 0.25), 2.0, 1.0, 4.0
    elif case == "star":
        mesh, k, c_o, c_i = HB.star_mesh(48, 0.25, 5), 3.0, 1.0, 9.0
    else:
        mesh, k, c_o, c_i = HB.circle_mesh(40, 0.25), 2.0 + 0.4j, 1.0, 4.0
    ko = math.sqrt(c_o) * k
    waves = [PLANE_X, [WAVES[6]]]
    us, gs = _wave_fns(waves, ko)
    x = tp.direct_second_kind.solve(mesh, us, gs, 11, k, c_o, c_i)
    U = u_inc_N(mesh, waves, ko)
    A, rhs = O.tp_A_rhs(mesh, 11, k, c_o, c_i, U)
    xo = np.linalg.solve(A, rhs)
    kappa = np.linalg.cond(A)
    assert rel(x, xo) <= 1e-12 * kappa, (rel(x, xo), kappa)
    if case == "disk":
        N = mesh.getNumPanels()
        ud, un = disk_traces(0.25, k, c_o, c_i, 2 * math.pi * np.arange(N) / N)
        assert rel(x[:N, 0], ud) <= 5e-6 and rel(x[N:, 0], un) <= 2e-5


@pytest.mark.gpu
def test_many_waves_share_one_factorisation(O):
    mesh = HB.star_mesh(40, 0.25, 5)
    k, c_o, c_i = 2.5, 1.0, 6.0
    waves = [[I.wavespec_t(I.WaveType.Plane, 1.0, angle=0.3 * j)] for j in range(9)] + [[w] for w in WAVES[:2]]
    us, gs = _wave_fns(waves, k)
    X = tp.direct_second_kind.solve(mesh, us, gs, 11, k, c_o, c_i)
    assert X.shape == (80, 11)
    for j in (0, 7, 8, 10):
        xj = tp.direct_second_kind.solve(mesh, us[j], gs[j], 11, k, c_o, c_i)
        assert rel(X[:, j], xj) <= 1e-13


@pytest.mark.gpu
def test_solve_rejects_undefined_right_hand_side():
    mesh = HB.circle_mesh(12, 0.25)
    with pytest.raises(HB.HbemError, match="Failed to solve for traces, right-hand side is undefined"):
        tp.direct_second_kind.solve(mesh, lambda x1, x2: complex("nan"), lambda x1, x2: np.zeros(2), 7, 2.0, 1.0,
                                    4.0)


# ---------------------------------------------------------------------------
# C++ facade (include/hbem/incoming.hpp, include/hbem/solvers.hpp)
# ---------------------------------------------------------------------------
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ALL_REAL = """# every wave type
Plane, 1.3, 0.2, -0.1, 1.047
CircularJ, 0.7, 0.1, 0.3, -2
CircularY, 0.5, -0.4, 0.2, 3
FourierHankel1, 1.0, 0.5, 0.5, -1
FourierHankel2, 1.0, -0.5, 0.1, 2
Herglotz, 1.0, -0.2, 0, 0.85, 0.1
"""


def build_cpp_driver(tmp_path):
    import subprocess
    exe = str(tmp_path / "transmission_driver")
    lib = os.path.join(ROOT, "helmholtzbem_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "transmission_driver.cpp"), "-o", exe, "-L", lib, "-lhbem",
                    "-Wl,-rpath," + lib], check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return exe


def run_cpp(exe, *args):
    import subprocess
    out = subprocess.run([exe, *map(str, args)], check=True, stdout=subprocess.PIPE, text=True, timeout=120).stdout
    return np.array([complex(float(a), float(b)) for a, b in (ln.split() for ln in out.strip().split("\n"))])


def test_cpp_incoming_matches_python(tmp_path):
    exe = build_cpp_driver(tmp_path)
    files = [os.path.join(GOLD, f) for f in sorted(os.listdir(GOLD))]
    allf = tmp_path / "all.dat"
    allf.write_text(ALL_REAL)
    pts = np.array([[0.31, -0.22], [-0.13, 0.47], [0.05, 0.06]])
    for f in files + [str(allf)]:
        got = run_cpp(exe, "waves", f, 3.0).reshape(3, 3)
        w = I.load(f)
        want = np.concatenate([I.compute(w, pts, 3.0)[:, None], I.compute_del(w, pts, 3.0)], axis=1)
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max()), f


@pytest.mark.gpu
def test_cpp_facade_solve_matches_python(tmp_path, O):
    exe = build_cpp_driver(tmp_path)
    f = os.path.join(GOLD, "incoming_plane_at_60deg.dat")
    x = run_cpp(exe, "solve", f, 32, 2.0, 1.0, 4.0)
    mesh = HB.circle_mesh(32, 0.25)
    w = I.load(f)
    u, g = I.as_functions(w, 2.0)
    xp = tp.direct_second_kind.solve(mesh, u, g, 11, 2.0, 1.0, 4.0)
    assert rel(x, xp) <= 1e-13
    A, rhs = O.tp_A_rhs(mesh, 11, 2.0, 1.0, 4.0, u_inc_N(mesh, [w], 2.0))
    assert rel(x, np.linalg.solve(A, rhs)[:, 0]) <= 1e-12 * np.linalg.cond(A)
