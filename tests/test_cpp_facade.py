"""The C++ facade (include/hbem/hbem.hpp) compiles against the C-ABI library
with g++ and, on the GPU, reproduces the oracle through a reference-style
driver (the 48-panel square of the reference's small CPU case)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_sigma.cpp")
LIBDIR = os.path.join(ROOT, "helmholtzbem_amd")


def build_driver(tmp_path):
    exe = str(tmp_path / "facade_sigma")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L", LIBDIR, "-lhbem", "-Wl,-rpath," + LIBDIR]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return exe


def test_facade_compiles_and_links(tmp_path):
    exe = build_driver(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_facade_driver_matches_oracle(tmp_path, H, O):
    exe = build_driver(tmp_path)
    out = subprocess.run([exe, "12", "11"], check=True, stdout=subprocess.PIPE, text=True, timeout=120).stdout.split("\n")
    s, der, v01 = (float(x) for x in out[0].split())
    assert "Requested layer is not computed" in out[1] or "Invalid order" in out[1]
    mesh = H.square_mesh(0.5, 12)
    S = O.SolOp(mesh, 11)
    T, Td = S.gen_sol_op(2.0, 1.0, 20.0, d=1)
    R = H.gaussian_test_matrix(96, 2, 0)
    assert abs(s - O.sv(T, R, 2)) <= 1e-10 * s
    assert abs(der - O.sv_der(T, Td, R, 30)[1]) <= 1e-8 * max(1.0, abs(der))
    K, W, V = O.assemble(mesh, 11, 2.0, 20.0, d=0)
    assert abs(v01 - abs(V[0][0, 1])) <= 1e-12 * abs(V[0][0, 1])
    tag, err, nused = out[2].split()
    assert tag == "mixed_far" and float(err) <= 1e-8 and int(nused) >= 0


def test_facade_host_api(tmp_path):
    """The vector randGaussian, assembleSparse's declaration and the
    randomized_svd timers (zero before any timed device work)."""
    exe = str(tmp_path / "facade_api")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "facade_api.cpp"), "-o", exe, "-L", LIBDIR, "-lhbem",
           "-Wl,-rpath," + LIBDIR]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    out = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True, timeout=60).stdout.split("\n")
    name, n, err = out[0].split()
    assert name == "randGaussian_vector" and int(n) == 7 and float(err) <= 1e-15
    assert out[1] == "assembleSparse 1"
    assert out[2] == "timers 0 0 0 0"
