"""Transmission solve (host side): ``tp::direct_second_kind``.

Mirrors the reference's second-kind direct BIE solver for the Helmholtz
transmission problem (solvers.hpp:73-113, solvers.cpp:104-152) over the
C-ABI entry points ``hbem_tp_A_rhs`` / ``hbem_tp_solve``:

* ``A_rhs(mesh, u_inc, u_inc_del, k, c_o, c_i, sol_op)`` returns the operator
  A = sol_op.gen_sol_op(k, c_o, c_i) and rhs = A_o u_inc_N, with
  A_o = [[-K_o + M/2, V_o], [W_o, K_o^T + M/2]] and u_inc_N the P1
  interpolation of the incoming wave's Dirichlet and Neumann traces;
* ``solve(mesh, u_inc, u_inc_del, order, k, c_o, c_i)`` returns the traces
  (Dirichlet, then Neumann coefficients of the total field) solving
  A x = rhs with the unprojected operator, after the reference's finiteness
  checks; the LU solve runs on the device.

``u_inc(x1, x2)`` is the incoming wave and ``u_inc_del(x1, x2)`` its gradient
(a 2-vector), as in the reference's std::function arguments; several waves
may be passed as lists, which share one assembly and one factorisation.
"""
from __future__ import annotations

from typing import Callable, Sequence, Union

import numpy as np

from . import BuilderData, ContinuousSpace1, SolutionsOperator, _vptr
from ._lib import check, lib

Wave = Callable[[float, float], complex]
Grad = Callable[[float, float], Sequence[complex]]


def _u_inc_N(bd: BuilderData, mesh, u_inc, u_inc_del):
    """[Interpolate_helmholtz(u_inc); Interpolate_helmholtz_neu(u_inc_del)]
    per wave, as the columns of a 2N x nwaves matrix (solvers.cpp:124-127)."""
    many = isinstance(u_inc, (list, tuple))
    us = list(u_inc) if many else [u_inc]
    gs = list(u_inc_del) if many else [u_inc_del]
    if len(us) != len(gs) or not us:
        raise ValueError("need one gradient per incoming wave")
    N = bd.getNumPanels()
    out = np.zeros((2 * N, len(us)), dtype=np.complex128, order="F")
    for j, (u, g) in enumerate(zip(us, gs)):
        out[:N, j] = bd.test_space.Interpolate_helmholtz(u, mesh)
        out[N:, j] = bd.trial_space.Interpolate_helmholtz_neu(g, mesh)
    return out, many


class tp:
    """namespace tp (solvers.hpp:65-201)."""

    class direct_second_kind:
        """namespace tp::direct_second_kind (solvers.hpp:73-200)."""

        @staticmethod
        def A_rhs(mesh, u_inc: Union[Wave, Sequence[Wave]], u_inc_del: Union[Grad, Sequence[Grad]], k, c_o: float,
                  c_i: float, sol_op: SolutionsOperator):
            """(A, rhs) of the transmission problem (solvers.cpp:106-131).
            A is projected when sol_op was built with projection, as the
            reference's gen_sol_op call on sol_op decides; rhs has one column
            per wave when lists are passed."""
            bd = sol_op.getBuilderData()
            U, many = _u_inc_N(bd, mesh, u_inc, u_inc_del)
            k = complex(k)
            n = U.shape[0]
            A = np.zeros((n, n), dtype=np.complex128, order="F")
            rhs = np.zeros_like(U, order="F")
            check(lib().hbem_tp_A_rhs(sol_op._g.h, k.real, k.imag, float(c_o), float(c_i), _vptr(U), U.shape[1],
                                      _vptr(A), _vptr(rhs)))
            sol_op._remember(k, c_o, c_i)
            return A, (rhs if many else rhs[:, 0])

        @staticmethod
        def solve(mesh, u_inc: Union[Wave, Sequence[Wave]], u_inc_del: Union[Grad, Sequence[Grad]], order: int, k,
                  c_o: float, c_i: float):
            """Traces of the total field on the boundary (solvers.cpp:133-152):
            2N coefficients (Dirichlet, then Neumann), or 2N x nwaves for
            lists of waves."""
            space = ContinuousSpace1()
            bd = BuilderData(mesh, space, space, order)
            sol_op = SolutionsOperator(bd, False)
            U, many = _u_inc_N(bd, mesh, u_inc, u_inc_del)
            k = complex(k)
            sol = np.zeros_like(U, order="F")
            check(lib().hbem_tp_solve(sol_op._g.h, k.real, k.imag, float(c_o), float(c_i), _vptr(U), U.shape[1],
                                      _vptr(sol)))
            return sol if many else sol[:, 0]
