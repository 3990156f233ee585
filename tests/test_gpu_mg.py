"""Multigrid-preconditioned u-solve (VFG_PC_MG): the matrix-free counterpart
of the reference's PCHYPRE / GAMG u-preconditioners (VFCommon.c:1041-1047).

Bar: the same linear system as the Jacobi solve, so the solution agrees with
the oracle's solve within 1e-8 relative (both to 1e-11..1e-12) and the FP64
residual through VF_UResidual is at the solve tolerance, on nested (even) and
non-nested (odd, flat) hierarchies; the
mixed-precision solve with the V-cycle agrees with FP64 within 1e-7; the
alternate-minimisation loop gives the same fields with -U_pc_type hypre as
with Jacobi.  Iteration counts are checked against Jacobi's on the same system.
"""
import numpy as np
import pytest

from tests import scenarios as S
from tests.scenarios import relerr

pytestmark = pytest.mark.gpu

GRIDS = [
    ((17, 17, 17), "clamped"),   # 16^3 -> 8^3 (nested)
    ((17, 9, 13), "mixed"),      # 16x8x12 -> 8x8x6 (y kept)
    ((11, 10, 9), "mixed"),      # 10x9x8 -> 5x5x8 (non-nested, z kept)
    ((26, 26, 26), "clamped"),   # 25^3 -> 13^3 -> 7^3 (non-nested)
    ((41, 21, 6), "test1"),      # 40x20x5 -> 20x10x5 -> 10x5x5 -> 5x5x5 (flat: z not coarsened)
    ((33, 33, 33), "test1"),     # 32^3 -> 16^3 -> 8^3
]
BCS = {"clamped": S.CLAMPED, "mixed": S.MIXED_U, "test1": S.TEST1_U}


@pytest.mark.parametrize("n,bc", GRIDS, ids=["x".join(map(str, g[0])) + "-" + g[1] for g in GRIDS])
def test_mg_u_solve_matches_oracle(oracle, n, bc):
    from vpfhf_amd import _lib
    s = S.make(n=n, seed=7, bcU=BCS[bc])
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    BCU = oracle.vec_set_from_bc(op, "U") * 0.01
    U0 = oracle.vec_apply_dirichlet(op, np.zeros(3 * op.nn), BCU, "U")
    Uo = U0.copy()
    oracle.u_solve(op, F, Uo, BCU, rtol=1e-12)
    Um, im = P.VF_StepU(U0, BCU, rtol=1e-11, pc=_lib.PC_MG)
    Uj, ij = P.VF_StepU(U0, BCU, rtol=1e-11, pc=_lib.PC_JACOBI)
    assert im.status == 0 and ij.status == 0
    assert relerr(Um, Uo) < 1e-8, relerr(Um, Uo)
    r = oracle.u_residual(op, F, Um, BCU)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(oracle.u_residual(op, F, U0, BCU))
    if n[0] == n[1] == n[2] >= 17:
        assert im.iterations * 3 < ij.iterations, (im.iterations, ij.iterations)
    P.close()


def test_mg_repeated_solves_follow_new_coefficients(oracle):
    """The hierarchy is built once per handle; its coefficients follow V and E
    at every solve."""
    from vpfhf_amd import _lib
    s = S.make(n=(17, 17, 17), seed=11)
    P = S.to_gpu(s)
    for k in range(3):
        V = np.clip(s["V"] * (1. - .3 * k) + .1 * k, 0., 1.)
        s2 = dict(s)
        s2["V"] = V
        op, F = S.to_oracle(oracle, s2)
        P.set_fields(V=V)
        BCU = np.zeros(3 * op.nn)
        U0 = np.zeros(3 * op.nn)
        Uo = U0.copy()
        oracle.u_solve(op, F, Uo, BCU, rtol=1e-12)
        Um, im = P.VF_StepU(U0, BCU, rtol=1e-11, pc=_lib.PC_MG)
        assert im.status == 0 and relerr(Um, Uo) < 1e-8, (k, relerr(Um, Uo))
    P.close()


def test_mg_mixed_matches_fp64():
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
    Um, im = P.VF_StepU(U0, BCU, rtol=1e-10, precision=_lib.PREC_MIXED, pc=_lib.PC_MG)
    assert i64.status == 0 and im.status == 0
    assert im.relres <= 1e-10
    assert relerr(Um, U64) < 1e-7
    F = P.VF_UResidual(Um, BCU)
    assert np.linalg.norm(F) <= 2e-10 * np.linalg.norm(P.VF_UResidual(U0, BCU))
    P.close()
    H.close()


def test_mg_altmin_matches_jacobi():
    """Two alternate-minimisation iterations of the 32^3 penny workload
    (crack pressure, beta = 0) with -U_pc_type hypre and with Jacobi."""
    from vpfhf_amd import workloads
    spec = workloads.spec_penny32(beta=0.0)
    out = []
    for extra in ("-vfg_pc jacobi", "-U_pc_type hypre"):
        H, _ = workloads.make_host(spec, u_rtol=1e-11, extra=extra)
        info = H.altmin(maxit=2)
        U, V, _ = H.state()
        out.append((info, U, V, H.energies()))
        H.close()
    (ij, Uj, Vj, ej), (im, Um, Vm, em) = out
    assert ij.status == 0 and im.status == 0
    assert relerr(Um, Uj) < 1e-7 and np.max(np.abs(Vm - Vj)) < 1e-7
    assert abs(em["elastic"] - ej["elastic"]) <= 1e-7 * abs(ej["elastic"])
    assert im.u_its * 5 < ij.u_its, (im.u_its, ij.u_its)


@pytest.mark.parametrize("n,bc", [((17, 17, 17), "clamped"), ((26, 26, 26), "clamped"), ((65, 65, 33), "test1")],
                         ids=["17-nested-small", "26-nonnested", "65x65x33-nested-separable"])
def test_mg_vcycle_is_symmetric(n, bc):
    """The V-cycle as an operator (vfg_u_pc_apply, the PCSHELL apply): CG needs
    M symmetric positive definite, which holds when every restriction is the
    transpose of its prolongation and the pre- and post-smoothing are adjoint.
    r2.M r1 = r1.M r2 to FP32 rounding on grids that take each restriction
    kernel (four lanes per entry onto small grids, the general non-nested
    gather, the separable three-pass form between nested large levels);
    r.M r > 0.  A transfer that is not the other's transpose breaks the first
    by orders of magnitude.  On the Dirichlet rows M is the identity.""" 
    s = S.make(n=n, seed=3, bcU=BCS[bc])
    P = S.to_gpu(s)
    free = P.bc_mask("U") == 0
    rng = np.random.default_rng(9)
    r1 = np.where(free, rng.standard_normal(free.size), 0.)
    r2 = np.where(free, rng.standard_normal(free.size), 0.)
    z1, z2 = P.u_pc_apply(r1), P.u_pc_apply(r2)
    a, b = r2 @ z1, r1 @ z2
    scale = np.linalg.norm(r1) * np.linalg.norm(z2)
    assert abs(a - b) <= 1e-5 * scale, (a, b, scale)
    assert r1 @ z1 > 0 and r2 @ z2 > 0
    # the apply is deterministic and linear
    assert np.array_equal(P.u_pc_apply(r1), z1)
    z12 = P.u_pc_apply(r1 + 2. * r2)
    assert np.linalg.norm(z12 - (z1 + 2. * z2)) <= 1e-5 * np.linalg.norm(z12)
    # Dirichlet rows: identity, and they do not reach the free rows
    rd = rng.standard_normal(free.size)
    zd = P.u_pc_apply(r1 + np.where(free, 0., rd))
    assert np.array_equal(zd[~free], rd[~free])
    assert np.array_equal(zd[free], z1[free])
    assert not free.all() and np.all(z1[~free] == 0.)
    P.close()



def test_pcshell_cg_matches_the_library_solve(oracle):
    """The PETSc binding of the u-KSP with the shells: a host CG (PETSc's
    KSPCG, restated) over the MatShell product (vfg_u_apply) and the PCSHELL
    (vfg_u_pc_apply) solves the Newton system J dU = -F(U0) of
    SNESNEWTONLS's first step (VF_UResidual at U0 carrying its Dirichlet
    values); U0 + dU matches the oracle's solve and the iteration count is
    the library multigrid solve's."""
    from vpfhf_amd import _lib
    s = S.make(n=(33, 33, 33), seed=7, bcU=S.MIXED_U)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    BCU = oracle.vec_set_from_bc(op, "U") * 0.01
    U0 = oracle.vec_apply_dirichlet(op, np.zeros(3 * op.nn), BCU, "U")
    Uo = U0.copy()
    oracle.u_solve(op, F, Uo, BCU, rtol=1e-12)
    b = -P.VF_UResidual(U0, BCU)
    x = np.zeros_like(b)
    r = b.copy()
    z = P.u_pc_apply(r)
    p = z.copy()
    rz, r0, its = r @ z, np.linalg.norm(b), 0
    while np.linalg.norm(r) > 1e-10 * r0 and its < 200:
        q = P.UIJacobian_mult(p)
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = P.u_pc_apply(r)
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
        its += 1
    U = U0 + x
    _, im = P.VF_StepU(U0, BCU, rtol=1e-10, pc=_lib.PC_MG)
    print("PCSHELL CG", its, "iterations; library multigrid solve", im.iterations)
    assert relerr(U, Uo) < 1e-8, relerr(U, Uo)
    assert abs(its - im.iterations) <= 2, (its, im.iterations)
    P.close()
