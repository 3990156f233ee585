"""UNILATERAL_NOCOMPRESSION (-unilateral NOCOMPRESSION) on the GPU against the
oracle's restatement of VFMech.c:115-218, 299-376, 456-524, 819-908,
2125-2153.

Bar: as for UNILATERAL_NONE -- the u-Jacobian at a state U (tension and
compression points mixed), its diagonal, the residual F(U) = K(U) U - b(U),
the V operator with the tension-only coupling and the energies within 1e-12
relative (FP32 apply 2e-6); the Newton u-solve within 1e-8 of the oracle's
Newton solve; batch and engine refuse such problems.
"""
import numpy as np
import pytest

from tests import scenarios as S
from tests.scenarios import relerr

pytestmark = pytest.mark.gpu

SETS = [("clamped", S.CLAMPED, 1), ("test1", S.TEST1_U, 2), ("mixed", S.MIXED_U, 1)]


@pytest.fixture(params=SETS, ids=[q[0] for q in SETS])
def upair(request, oracle):
    name, bu, at = request.param
    s = S.make(n=(9, 8, 7), seed=sum(map(ord, name)) + 1, bcU=bu, atnum=at, unilateral=1)
    s["U"] = s["U"] * 5.  # random state: compressed and stretched Gauss points mixed
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    yield s, op, F, P
    P.close()


def test_nocomp_u_jacobian_and_diag(upair, oracle):
    s, op, F, P = upair
    A = oracle.u_jacobian(op, F, U=s["U"])
    P.set_state(s["U"])
    assert relerr(P.UIJacobian_mult(s["x"]), A @ s["x"]) < 1e-12
    x32 = s["x"].astype(np.float32)
    assert relerr(P.UIJacobian_mult(x32, fp32=True), A @ x32.astype(np.float64)) < 2e-6
    assert relerr(P.u_diag(), A.diagonal()) < 1e-12
    # the state changes the operator: U = 0 gives UNILATERAL_NONE's matrix
    A0 = oracle.u_jacobian(op, F)
    assert relerr(A0 @ s["x"], A @ s["x"]) > 1e-6
    P.set_state(np.zeros_like(s["U"]))
    assert relerr(P.UIJacobian_mult(s["x"]), A0 @ s["x"]) < 1e-12


def test_nocomp_residual_v_operator_energies(upair, oracle):
    s, op, F, P = upair
    BCU = oracle.vec_set_from_bc(op, "U") * 0.3
    assert relerr(P.VF_UResidual(s["U"], BCU), oracle.u_residual(op, F, s["U"], BCU)) < 1e-12
    assert relerr(P.u_rhs(), oracle.u_rhs(op, F)) < 1e-12
    P.VIJacobian_setup(s["U"])
    J = oracle.v_jacobian(op, F, s["U"])
    assert relerr(P.VIJacobian_mult(s["xv"]), J @ s["xv"]) < 1e-12
    assert relerr(P.VF_VResidual(s["V"]), oracle.v_residual(op, F, s["U"], s["V"])) < 1e-12
    el, _, pw, se = P.energies(s["U"], s["V"])
    rel, rpw = oracle.u_energy(op, F, s["U"])
    assert abs(el - rel) <= 1e-12 * abs(rel)
    ef = P.energy_fields(s["U"], s["V"])
    assert relerr(ef[:, 0], oracle.energy_fields(op, F, s["U"], s["V"])[:, 0]) < 1e-12


_NEWTON = {}


def _oracle_newton(oracle, scale):
    """the oracle's Newton solve of one loading (shared by both preconditioners)"""
    if scale not in _NEWTON:
        s = S.make(n=(13, 11, 9), seed=2, bcU=S.TEST1_U, unilateral=1, thermal=False, beta=0.0)
        op, F = S.to_oracle(oracle, s)
        BCU = oracle.vec_set_from_bc(op, "U") * scale
        U0 = oracle.vec_apply_dirichlet(op, np.zeros(3 * op.nn), BCU, "U")
        Uo = U0.copy()
        oracle.u_solve(op, F, Uo, BCU, rtol=1e-12)
        _NEWTON[scale] = (s, op, F, BCU, U0, Uo)
    return _NEWTON[scale]


@pytest.mark.parametrize("pc", ["mg", "jacobi"])
@pytest.mark.parametrize("scale", [-0.02, 0.02])
def test_nocomp_newton_u_solve(oracle, pc, scale):
    """Crack pressure only (no thermal or Biot load, whose branches jump at
    tr eps = 0), compressed or stretched along x.  The multigrid solve of the
    compressed case needs its smoother intervals re-estimated at every Newton
    step (the compression state changes the operator); with intervals from
    an earlier state it took 3e5 CG iterations instead of about 130, which
    the bound on the iteration count below guards against."""
    from vpfhf_amd import _lib
    s, op, F, BCU, U0, Uo = _oracle_newton(oracle, scale)
    P = S.to_gpu(s)
    Ug, info = P.VF_StepU(U0, BCU, rtol=1e-11, pc=_lib.PC_MG if pc == "mg" else _lib.PC_JACOBI)
    assert info.status == 0 and info.outer >= 1
    assert info.iterations < (1000 if pc == "mg" else 5000), info.iterations
    assert relerr(Ug, Uo) < 1e-8, relerr(Ug, Uo)
    r = oracle.u_residual(op, F, Ug, BCU)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(oracle.u_residual(op, F, U0, BCU))
    P.close()


def test_nocomp_refused_by_batch_and_engine():
    import ctypes as C
    import vpfhf_amd as V
    from vpfhf_amd import _lib
    L = _lib.load()
    P = V.Problem((9, 9, 9), unilateral=1)
    hs = (C.c_void_p * 1)(P.h)
    b = C.c_void_p()
    assert L.vfg_batch_create(hs, 1, C.byref(b)) == -1  # VFG_E_ARG
    o = _lib.EngineOpts()
    L.vfg_engine_opts_default(C.byref(o))
    e = C.c_void_p()
    assert L.vfg_engine_create(C.byref(o), C.byref(e)) == 0
    U = V.DeviceBuffer(24 * P.nn)
    Vb = V.DeviceBuffer(8 * P.nn)
    rc = L.vfg_engine_submit(e, P.h, U.ptr, U.ptr, Vb.ptr, Vb.ptr, 1, 1e-3, 0)
    assert rc == -1  # VFG_E_ARG
    L.vfg_engine_destroy(e)
    U.free()
    Vb.free()
    P.close()
