"""GPU parity of the u-/V-operators against the CPU oracle (FP64).

Bar: Dirichlet masks and BC vectors bit-exact; operator products, diagonals,
load vectors, residuals, energies and crack volume within 1e-12 relative
(2-norm) of the oracle's assembled FP64 operators; FP32 apply within 2e-6.
"""
import numpy as np
import pytest

from tests import scenarios as S
from tests.scenarios import relerr

pytestmark = pytest.mark.gpu

BC_SETS = [
    ("clamped", S.CLAMPED, S.NONE_V, 1),
    ("test1", S.TEST1_U, S.TEST1_V, 1),
    ("test2", S.TEST2_U, S.NONE_V, 2),
    ("mixed", S.MIXED_U, S.MIXED_V, 2),
]


@pytest.fixture(params=BC_SETS, ids=[b[0] for b in BC_SETS])
def pair(request, oracle):
    name, bu, bv, at = request.param
    s = S.make(n=(9, 8, 7), seed=sum(map(ord, name)), bcU=bu, bcV=bv, atnum=at)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    yield s, op, F, P
    P.close()


def test_masks_bit_exact(pair, oracle):
    s, op, F, P = pair
    assert np.array_equal(P.bc_mask("U"), oracle.bc_mask(op, "U"))
    assert np.array_equal(P.bc_mask("V"), oracle.bc_mask(op, "V"))


def test_bc_vectors_exact(pair, oracle):
    s, op, F, P = pair
    base = np.arange(3 * op.nn, dtype=np.float64) * 0.001
    assert np.array_equal(P.VecSetFromBC("U", base.copy()), oracle.vec_set_from_bc(op, "U", base.copy()))
    B = oracle.vec_set_from_bc(op, "U") * 1.7
    got = P.VecApplyDirichletBC(s["U"].copy(), B, "U")
    ref = oracle.vec_apply_dirichlet(op, s["U"].copy(), B, "U")
    assert np.array_equal(got, ref)
    gv = P.VecApplyDirichletBC(s["V"].copy(), s["V"], "V")
    rv = oracle.vec_apply_dirichlet(op, s["V"].copy(), s["V"].copy(), "V")
    assert np.array_equal(gv, rv)
    # in place (vec is its own BC vector), VFCommon.c:1146
    vin = s["V"].copy()
    rv2 = oracle.lib().vfo_vec_apply_dirichlet(op.ref, oracle.C.pointer(op.P.bcV), 1, oracle._dp(vin), oracle._dp(vin))
    assert np.array_equal(P.VecApplyDirichletBC(s["V"].copy(), None, "V"), vin)


def test_u_apply_matches_assembled(pair, oracle):
    s, op, F, P = pair
    A = oracle.u_jacobian(op, F)
    y_ref = A @ s["x"]
    y = P.UIJacobian_mult(s["x"])
    assert relerr(y, y_ref) < 1e-12
    # Dirichlet rows are identity, columns are kept (VFCartFE.c:1293)
    m = oracle.bc_mask(op, "U").astype(bool)
    assert np.array_equal(y[m], s["x"][m])


@pytest.mark.parametrize("n", [(9, 8, 7), (37, 33, 41), (18, 34, 17)])
def test_u_apply_forms_and_refresh(oracle, n):
    """Both forms of the FP64 apply (weights recomputed from V in every apply,
    or stored per element and refreshed after V changes) against the
    assembled operator, on grids of one tile, several tiles with a partial
    last tile, and tiles whose shared sides meet at 16-element boundaries;
    after vfg_set_fields(V) the stored form follows the new V."""
    from vpfhf_amd import _lib
    s = S.make(n=n, seed=31, bcU=S.MIXED_U, bcV=S.MIXED_V, atnum=2)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    try:
        y_ref = oracle.u_jacobian(op, F) @ s["x"]
        for form in (_lib.U_FORM_RECOMPUTE, _lib.U_FORM_STORED):
            P._c(P.L.vfg_u_set_form(P.h, form))
            assert relerr(P.UIJacobian_mult(s["x"]), y_ref) < 1e-12, form
        V2 = np.clip(s["V"] + 0.3 * np.sin(np.arange(op.nn)), 0., 1.)
        P.set_fields(V=V2)
        F2 = op.fields(V=V2, VIrrev=s["VIrrev"], theta=s["theta"], thetaRef=s["thetaRef"], pressure=s["pressure"])
        y2 = oracle.u_jacobian(op, F2) @ s["x"]
        assert relerr(y2, y_ref) > 1e-3
        assert relerr(P.UIJacobian_mult(s["x"]), y2) < 1e-12
        P._c(P.L.vfg_u_set_form(P.h, _lib.U_FORM_RECOMPUTE))
        assert relerr(P.UIJacobian_mult(s["x"]), y2) < 1e-12
    finally:
        P.close()


def test_u_apply_f32(pair, oracle):
    s, op, F, P = pair
    A = oracle.u_jacobian(op, F)
    y = P.UIJacobian_mult(s["x"].astype(np.float32), fp32=True)
    assert relerr(y, A @ s["x"].astype(np.float32).astype(np.float64)) < 2e-6


def test_u_apply_deterministic(pair):
    s, op, F, P = pair
    y1 = P.UIJacobian_mult(s["x"])
    y2 = P.UIJacobian_mult(s["x"])
    assert np.array_equal(y1, y2)


def test_u_diag(pair, oracle):
    s, op, F, P = pair
    A = oracle.u_jacobian(op, F)
    assert relerr(P.u_diag(), A.diagonal()) < 1e-12


def test_u_rhs_and_residual(pair, oracle):
    s, op, F, P = pair
    assert relerr(P.u_rhs(), oracle.u_rhs(op, F)) < 1e-12
    BCU = oracle.vec_set_from_bc(op, "U") * 0.3
    r_ref = oracle.u_residual(op, F, s["U"], BCU)
    assert relerr(P.VF_UResidual(s["U"], BCU), r_ref) < 1e-12


def test_v_operator(pair, oracle):
    s, op, F, P = pair
    P.VIJacobian_setup(s["U"])
    J = oracle.v_jacobian(op, F, s["U"])
    assert relerr(P.VIJacobian_mult(s["xv"]), J @ s["xv"]) < 1e-12
    assert relerr(P.v_diag(), J.diagonal()) < 1e-12
    assert relerr(P.v_rhs(), oracle.v_rhs(op, F, s["U"])) < 1e-12
    assert relerr(P.VF_VResidual(s["V"]), oracle.v_residual(op, F, s["U"], s["V"])) < 1e-12


def test_energies_and_crack_volume(pair, oracle):
    s, op, F, P = pair
    el, ins, pw, se = P.energies(s["U"], s["V"])
    rel, rpw = oracle.u_energy(op, F, s["U"])
    assert ins == 0.0
    assert abs(el - rel) <= 1e-12 * abs(rel)
    assert abs(pw - rpw) <= 1e-12 * max(abs(rpw), 1e-300)
    assert abs(se - oracle.v_energy(op, s["V"])) <= 1e-12 * abs(oracle.v_energy(op, s["V"]))
    vol, cf = P.VolumetricCrackOpening(s["U"], s["V"], cellfield=True)
    rvol, rcf = oracle.crack_volume(op, s["U"], s["V"], cellfield=True)
    hvol = np.prod([s["l"][d] / (s["n"][d] - 1) for d in range(3)])
    assert abs(vol - rvol) <= 1e-12 * np.abs(rcf).sum() * hvol
    assert relerr(cf, rcf) < 1e-12


def test_energy_fields(pair, oracle):
    """Per-cell energy terms against the oracle (1e-12 per column, 2-norm) and
    their sums against the GPU's own whole-domain energies."""
    s, op, F, P = pair
    ef = P.energy_fields(s["U"], s["V"])
    ref = oracle.energy_fields(op, F, s["U"], s["V"])
    assert ef.shape == (op.ne, 3)
    for c in range(3):
        assert relerr(ef[:, c], ref[:, c]) < 1e-12, c
    el, _, pw, se = P.energies(s["U"], s["V"])
    assert abs(ef[:, 0].sum() - el) <= 1e-12 * abs(el)
    assert abs(ef[:, 1].sum() - pw) <= 1e-12 * max(np.abs(ef[:, 1]).sum(), 1e-300)
    assert abs(ef[:, 2].sum() - se) <= 1e-12 * abs(se)
    nodal = P.CellToNodeInterpolation(ef.ravel(), 3)
    assert relerr(nodal, oracle.cell_to_node(op, ref.ravel(), 3)) < 1e-12


@pytest.mark.parametrize("n,l", [((9, 8, 7), (1., 1., 1.)), ((2, 2, 2), (1., 2., .5)), ((17, 3, 11), (3., .2, 1.7)),
                                 ((41, 37, 33), (1., 1.3, .6))])
def test_cell_to_node(oracle, n, l):
    """CellToNodeInterpolation (VFPermfield.c:1156-1327) for 1, 3 and 6
    components against the oracle: same cell order, same products, no
    contraction, so agreement is to the last bit or an ulp of rounding in
    the coordinate divisions."""
    s = S.make(n=n, l=l, seed=3)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    rng = np.random.default_rng(11)
    for dof in (1, 3, 6):
        cell = rng.standard_normal(dof * op.ne) * rng.lognormal(0, 2, dof * op.ne)
        got = P.CellToNodeInterpolation(cell, dof)
        ref = oracle.cell_to_node(op, cell, dof)
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) <= 4e-16 * np.max(np.abs(ref)), dof
    # the nodal crack-opening field of VolumetricCrackOpening (VFPermfield.c:144-146)
    vol, cf, nf = P.VolumetricCrackOpening(s["U"], s["V"], cellfield=True, nodefield=True)
    rvol, rcf = oracle.crack_volume(op, s["U"], s["V"], cellfield=True)
    assert relerr(nf, oracle.cell_to_node(op, rcf)) < 1e-12
    with pytest.raises(Exception):
        P.CellToNodeInterpolation(np.zeros(op.ne), dof=7)
    P.close()


@pytest.mark.parametrize("atnum", [1, 2])
def test_penny_v(oracle, atnum):
    s = S.make(n=(17, 15, 13), atnum=atnum)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    for c in [dict(center=(.5, .5, .5), r=.2, theta=0., phi=0., thickness=.09375),
              dict(center=(.41, .55, .47), r=.23, theta=37., phi=71., thickness=.1)]:
        ref = oracle.penny_v(op, c["center"], c["r"], c["theta"], c["phi"], c["thickness"])
        got = P.penny_v(c["center"], c["r"], c["theta"], c["phi"], c["thickness"])
        assert np.max(np.abs(got - ref)) < 1e-13
    P.close()


def test_larger_grid_apply(oracle):
    """A grid spanning several workgroup tiles and z-chunks in every direction."""
    s = S.make(n=(37, 33, 41), seed=7, bcU=S.MIXED_U)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    A = oracle.u_jacobian(op, F)
    assert relerr(P.UIJacobian_mult(s["x"]), A @ s["x"]) < 1e-12
    P.VIJacobian_setup(s["U"])
    J = oracle.v_jacobian(op, F, s["U"])
    assert relerr(P.VIJacobian_mult(s["xv"]), J @ s["xv"]) < 1e-12
    P.close()


EDGE_GRIDS = [(2, 2, 2), (3, 2, 2), (2, 40, 3), (64, 2, 2), (16, 16, 16), (31, 31, 5), (46, 17, 4), (9, 9, 97),
              (33, 33, 4), (49, 17, 3), (17, 50, 5)]


@pytest.mark.parametrize("n", EDGE_GRIDS, ids=["x".join(map(str, g)) for g in EDGE_GRIDS])
def test_apply_edge_grids(oracle, n):
    """One element, slabs one element thick, exact multiples of the 15-node
    and of the 16-element tiles (full last tiles in x and y, corners shared by
    four tiles), and a tall grid crossing many z-chunks: u- and V-apply
    against the assembled oracle operators (FP64 1e-12, FP32 1e-5), device
    pointers too."""
    import vpfhf_amd as V
    s = S.make(n=n, l=(1.0, 0.7, 1.3), seed=sum(n), bcU=S.MIXED_U, bcV=S.MIXED_V, atnum=2)
    op, F = S.to_oracle(oracle, s)
    P = S.to_gpu(s)
    A = oracle.u_jacobian(op, F)
    ref = A @ s["x"]
    assert relerr(P.UIJacobian_mult(s["x"]), ref) < 1e-12
    assert relerr(P.UIJacobian_mult(s["x"], fp32=True), ref) < 1e-5
    L = V.load()
    dx = V.DeviceBuffer(s["x"].nbytes).upload(s["x"])
    dy = V.DeviceBuffer(s["x"].nbytes)
    V.check(L.vfg_u_apply(P.h, dx.ptr, dy.ptr, V.DEVICE), P.h)
    assert relerr(dy.download(), ref) < 1e-12
    dx.free()
    dy.free()
    P.VIJacobian_setup(s["U"])
    J = oracle.v_jacobian(op, F, s["U"])
    assert relerr(P.VIJacobian_mult(s["xv"]), J @ s["xv"]) < 1e-12
    assert relerr(P.v_diag(), J.diagonal()) < 1e-12
    assert relerr(P.u_diag(), A.diagonal()) < 1e-12
    P.close()
