"""Batching form: an ensemble of uneven-size problems solved together.

Bar: the persistent multi-problem apply equals each member's own apply
bitwise in FP32 and within 1e-14 relative in FP64 (the single-problem FP64
apply runs another element routine with another operation order); the batched
FP64 u-solve agrees with each member's single-problem
solve within 1e-8 relative; the mixed batched solve (FP32 applies in FP64
refinement) within 1e-7 of the FP64 one; a member's batched result is
bitwise identical alone and in the mix.
"""
import numpy as np
import pytest

from tests.scenarios import relerr

pytestmark = pytest.mark.gpu


def _members(sizes, seed=0):
    from vpfhf_amd import workloads
    out = []
    for i, n in enumerate(sizes):
        spec = workloads.spec_hetero(n, seed=seed * 100 + i, ncracks=1 + i % 3, dtheta=0.0)
        H, f = workloads.make_host(spec)
        out.append(H)
    return out


def test_batch_apply_equals_single():
    import vpfhf_amd as V
    from vpfhf_amd.batch import Batch
    hosts = _members([7, 12, 17, 23])
    B = Batch(hosts)
    rng = np.random.default_rng(0)
    xs = [rng.standard_normal(3 * h.nn) for h in hosts]
    ys = B.u_apply(xs)
    L = V.load()
    for h, x, y in zip(hosts, xs, ys):
        yy = np.empty_like(x)
        V.check(L.vfg_u_apply(h.h, x.ctypes.data, yy.ctypes.data, V.HOST), h.h)
        assert relerr(y, yy) < 1e-14, relerr(y, yy)
    y32 = B.u_apply([x.astype(np.float32) for x in xs], fp32=True)
    for h, x, y in zip(hosts, xs, y32):
        yy = np.empty(3 * h.nn, np.float32)
        x32 = x.astype(np.float32)  # keep the array alive across the C call
        V.check(L.vfg_u_apply_f32(h.h, x32.ctypes.data, yy.ctypes.data, V.HOST), h.h)
        assert np.array_equal(y, yy)
    B.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_u_solve(mixed):
    import vpfhf_amd as V
    from vpfhf_amd.batch import Batch
    from vpfhf_amd import _lib
    hosts = _members([8, 13, 19, 25, 11], seed=1)
    states = [h.state() for h in hosts]
    Us = [s[0] for s in states]
    Bs = [np.zeros_like(u) for u in Us]
    B = Batch(hosts)
    B.set_graph(True, 8)
    Ub, infos = B.u_solve(Us, Bs, rtol=1e-10, mixed=mixed)
    assert all(i.status == 0 for i in infos)
    L = V.load()
    for h, U0, Ug in zip(hosts, Us, Ub):
        Ur = np.array(U0)
        info = _lib.SolveInfo()
        o = _lib.SolveOpts(1e-10, 0., 100000, 0, 0, 4)
        Z = np.zeros_like(U0)
        V.check(L.vfg_u_solve(h.h, Ur.ctypes.data, Z.ctypes.data, o, info, V.HOST), h.h)
        assert relerr(Ug, Ur) < (1e-7 if mixed else 1e-8)
    # alone vs in the mix: member 2 batched on its own
    B2 = Batch([hosts[2]])
    B2.set_graph(True, 8)
    Ua, _ = B2.u_solve([Us[2]], [Bs[2]], rtol=1e-10, mixed=mixed)
    assert np.array_equal(Ua[0], Ub[2])
    # graphs on/off give the same iterates
    B.set_graph(False, 8)
    Un, _ = B.u_solve(Us, Bs, rtol=1e-10, mixed=mixed)
    for a, b in zip(Un, Ub):
        assert np.array_equal(a, b)
    B.close()
    B2.close()


def test_batch_u_solve_fp64_multigrid():
    """FP64 batched CG under the members' FP32 V-cycles, batched level by level
    (the single-problem default solve, for the whole ensemble): within 1e-8 of
    each member's own FP64 solve, bitwise the same alone and in the mix and
    with graphs on or off, and a quarter of the Jacobi iterations or fewer."""
    import vpfhf_amd as V
    from vpfhf_amd.batch import Batch
    from vpfhf_amd import _lib
    hosts = _members([8, 13, 19, 25, 11, 32], seed=4)
    Us = [h.state()[0] for h in hosts]
    Bs = [np.zeros_like(u) for u in Us]
    B = Batch(hosts)
    B.set_graph(True, 8)
    Ub, infos = B.u_solve(Us, Bs, rtol=1e-10, pc=_lib.PC_MG)
    assert all(i.status == 0 for i in infos)
    _, ij = B.u_solve(Us, Bs, rtol=1e-10)
    assert sum(i.iterations for i in infos) * 4 < sum(i.iterations for i in ij)
    L = V.load()
    for h, U0, Ug in zip(hosts, Us, Ub):
        Ur = np.array(U0)
        info = _lib.SolveInfo()
        o = _lib.SolveOpts(1e-11, 0., 100000, _lib.PC_MG, _lib.PREC_FP64, 4)
        Z = np.zeros_like(U0)
        V.check(L.vfg_u_solve(h.h, Ur.ctypes.data, Z.ctypes.data, o, info, V.HOST), h.h)
        assert relerr(Ug, Ur) < 1e-8, relerr(Ug, Ur)
    B2 = Batch([hosts[5]])
    B2.set_graph(True, 8)
    Ua, _ = B2.u_solve([Us[5]], [Bs[5]], rtol=1e-10, pc=_lib.PC_MG)
    assert np.array_equal(Ua[0], Ub[5])
    B.set_graph(False, 8)
    Un, _ = B.u_solve(Us, Bs, rtol=1e-10, pc=_lib.PC_MG)
    for a, b in zip(Un, Ub):
        assert np.array_equal(a, b)
    B.close()
    B2.close()


def test_batch_u_solve_multigrid():
    """Batched V-cycles (VFG_PC_MG, FP32 inner CG in FP64 refinement): within
    1e-7 of each member's single-problem FP64 solve, bitwise the same alone and
    in the mix and with graphs on or off, far fewer iterations than Jacobi.
    Members of even and odd sides (nested and non-nested hierarchies)."""
    import vpfhf_amd as V
    from vpfhf_amd.batch import Batch
    from vpfhf_amd import _lib
    hosts = _members([8, 13, 19, 25, 11, 32], seed=2)
    Us = [h.state()[0] for h in hosts]
    Bs = [np.zeros_like(u) for u in Us]
    B = Batch(hosts)
    B.set_graph(True, 8)
    Ub, infos = B.u_solve(Us, Bs, rtol=1e-10, mixed=True, pc=_lib.PC_MG)
    assert all(i.status == 0 for i in infos)
    _, ij = B.u_solve(Us, Bs, rtol=1e-10, mixed=True)
    assert sum(i.iterations for i in infos) * 4 < sum(i.iterations for i in ij)
    L = V.load()
    for h, U0, Ug in zip(hosts, Us, Ub):
        Ur = np.array(U0)
        info = _lib.SolveInfo()
        o = _lib.SolveOpts(1e-10, 0., 100000, _lib.PC_JACOBI, _lib.PREC_FP64, 4)
        Z = np.zeros_like(U0)
        V.check(L.vfg_u_solve(h.h, Ur.ctypes.data, Z.ctypes.data, o, info, V.HOST), h.h)
        assert relerr(Ug, Ur) < 1e-7
    B2 = Batch([hosts[3]])
    B2.set_graph(True, 8)
    Ua, _ = B2.u_solve([Us[3]], [Bs[3]], rtol=1e-10, mixed=True, pc=_lib.PC_MG)
    assert np.array_equal(Ua[0], Ub[3])
    B.set_graph(False, 8)
    Un, _ = B.u_solve(Us, Bs, rtol=1e-10, mixed=True, pc=_lib.PC_MG)
    for a, b in zip(Un, Ub):
        assert np.array_equal(a, b)
    B.close()
    B2.close()
