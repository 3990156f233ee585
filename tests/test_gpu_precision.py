"""Precision form: FP32 operator applies (Jacobi or Chebyshev(4)-Jacobi
preconditioned CG) inside FP64 iterative refinement, against the plain FP64
solve of the same system.

Bar: the refined solution reaches the FP64 relative residual 1e-10 and agrees
with the FP64 solve within 1e-7 relative (the precision form's bar is 1e-6); the FP64
residual is evaluated by the FP64 operator.
"""
import numpy as np
import pytest

from tests.scenarios import relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pc", ["jacobi", "chebyshev"])
def test_mixed_u_solve_matches_fp64(pc):
    import vpfhf_amd as V
    from vpfhf_amd import _lib, workloads
    spec = workloads.spec_hetero(24, seed=4)
    H, f = workloads.make_host(spec)
    U0, Vf, VI = H.state()
    P = V.Problem((25, 25, 25), nu=.2, alpha=1e-5, beta=1., eta=1e-5, epsilon=spec["eps"], atnum=1,
                  pressurize=True)
    P.set_bc([V.make_bc({q: V.ZERO for q in range(6)})] * 3, V.make_bc())
    P.set_fields(V=Vf, VIrrev=VI, theta=f["theta"], thetaRef=f["thetaRef"], pressure=f["pressure"],
                 E_elem=f["E_elem"])
    BCU = np.zeros(3 * P.nn)
    U64, i64 = P.VF_StepU(U0, BCU, rtol=1e-10)
    Um, im = P.VF_StepU(U0, BCU, rtol=1e-10, precision=_lib.PREC_MIXED,
                        pc=_lib.PC_CHEBYSHEV if pc == "chebyshev" else _lib.PC_JACOBI)
    assert i64.status == 0 and im.status == 0
    assert im.relres <= 1e-10 and im.outer >= 2
    assert relerr(Um, U64) < 1e-7
    # FP64 residual of the refined solution, through VF_UResidual
    F = P.VF_UResidual(Um, BCU)
    assert np.linalg.norm(F) <= 2e-10 * np.linalg.norm(P.VF_UResidual(U0, BCU))
    P.close()
    H.close()
