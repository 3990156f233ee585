"""GPU solvers (u: PCG Newton step; V: bound-constrained active-set Newton)
against the oracle's solvers and the reference's golden outputs.

Tolerances: solutions of the same linear/QP problem agree to 1e-7 relative
(both solved to 1e-10..1e-12); energies of the reference's test2 run
(ValidationTests/results/runtest2.out, 7 printed digits) to 2e-6 relative.
"""
import numpy as np
import pytest

from tests import scenarios as S
from tests.scenarios import relerr

pytestmark = pytest.mark.gpu


def test_u_solve_matches_oracle(oracle):
    s = S.make(n=(11, 10, 9), seed=3, bcU=S.MIXED_U)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    BCU = oracle.vec_set_from_bc(op, "U") * 0.01
    U0 = oracle.vec_apply_dirichlet(op, np.zeros(3 * op.nn), BCU, "U")
    Uo = U0.copy()
    oracle.u_solve(op, F, Uo, BCU, rtol=1e-12)
    Ug, info = P.VF_StepU(U0, BCU, rtol=1e-12)
    assert info.status == 0
    assert relerr(Ug, Uo) < 1e-8
    # residual of the GPU solution through the oracle's VF_UResidual, relative
    # to the initial residual F(U0) (crack pressure and Dirichlet load), the
    # quantity rtol refers to; the slack over 1e-12 covers the difference
    # between the solver's recursive residual and the recomputed one
    r = oracle.u_residual(op, F, Ug, BCU)
    r0 = oracle.u_residual(op, F, U0, BCU)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(r0)
    P.close()


@pytest.mark.parametrize("atnum", [1, 2])
def test_v_solve_matches_oracle(oracle, atnum):
    s = S.make(n=(11, 10, 9), seed=5, atnum=atnum, bcV=S.MIXED_V if atnum == 2 else S.TEST1_V)
    s["U"] = s["U"] * 30.0  # enough elastic energy to activate the bounds
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    ub = s["VIrrev"] if atnum == 1 else None
    Vo = s["V"].copy()
    oracle.v_solve(op, F, s["U"], Vo, ub=ub, tol=1e-12)
    Vg, info = P.VF_StepV(s["U"], s["V"], ub=ub, rtol=1e-12)
    assert info.status == 0
    assert np.max(np.abs(Vg - Vo)) < 1e-7
    P.close()


def test_v_solve_long_chunks_matches_oracle(oracle):
    """The V-solve's CG runs over chunks of node planes that hold free nodes;
    once more than 1024 (tile, plane) pairs hold one, a chunk spans several
    planes.  A tall grid of one tile per plane whose free set covers every
    plane (AT2, strong elastic load) takes that path in its last Newton
    steps: the converged V against the oracle's solve."""
    s = S.make(n=(9, 8, 1100), l=(1.0, 0.9, 120.0), seed=11, atnum=2, bcV=S.MIXED_V)
    s["U"] = s["U"] * 30.0
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    Vo = s["V"].copy()
    oracle.v_solve(op, F, s["U"], Vo, ub=None, tol=1e-12)
    free = (Vo > 1e-12) & (Vo < 1. - 1e-12)
    planes = np.unique(np.arange(Vo.size)[free] // (9 * 8))
    assert planes.size > 1024  # more flagged (tile, plane) pairs than one-plane chunks allow
    Vg, info = P.VF_StepV(s["U"], s["V"], ub=None, rtol=1e-12)
    assert info.status == 0
    assert np.max(np.abs(Vg - Vo)) < 1e-7
    P.close()


def test_reference_test2_goldens(oracle):
    """ValidationTests/test2.c with the runtest2 options (makefile:207):
    -n 41,81,6 -l 1,2,.25 -length .4 -orientation 3 -nu 0 -epsilon 0.05 -eta 1e-8."""
    import vpfhf_amd as V
    from vpfhf_amd import _lib
    nx, ny, nz = 41, 81, 6
    P = V.Problem((nx, ny, nz), (1., 2., .25), E=1., nu=0., alpha=0., beta=0., Gc=1., eta=1e-8,
                  epsilon=.05, atnum=2)
    P.set_bc([V.make_bc({V.X0: V.ZERO}), V.make_bc({V.X1: V.ZERO}), V.make_bc({V.Z0: V.ZERO, V.Z1: V.ZERO})],
             V.make_bc())
    nn = nx * ny * nz
    X = np.array([0.] + [i * 1. / (nx - 1.) for i in range(1, nx)])
    i = np.arange(nn) % nx
    j = (np.arange(nn) // nx) % ny
    VIrrev = np.ones(nn)
    VIrrev[((j == ny // 2) | (j == ny // 2 - 1)) & (X[i] < .4)] = 0.
    zero = np.zeros(nn)
    p = np.full(nn, 1e-3)
    P.set_fields(V=VIrrev, VIrrev=VIrrev, theta=zero, thetaRef=zero, pressure=p)
    U = np.zeros(3 * nn)
    Vv, info = P.VF_StepV(U, VIrrev, rtol=1e-12)
    assert info.status == 0
    P.set_props(hasCrackPressure=1)
    P.set_fields(V=Vv)
    U, info = P.VF_StepU(U, np.zeros(3 * nn), rtol=1e-12)
    assert info.status == 0
    el, _, pw, se = P.energies(U, Vv)
    assert abs(se - 1.294241e-01) / 1.294241e-01 < 2e-6
    assert abs(el - 8.300042e-08) / 8.300042e-08 < 2e-6
    assert abs(pw - 1.660010e-07) / 1.660010e-07 < 2e-6
    P.close()
