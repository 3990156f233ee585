"""Serving form: continuous batching of alternate-minimisation requests.

Bars: a request's U and V are bitwise the same alone and in the mix (any
admission order, slot count or neighbours); the mixed-precision engine agrees
with the single-problem FP64 path (vfh_altmin: vfg_u_solve + vfg_v_solve) within
1e-6 relative in U, V, surface energy and crack volume, the FP64 engine within
1e-8 (same algorithm, different reduction order); admission never exceeds the HBM budget and rejects a request larger than
it; graph replays and plain launches give identical results.
"""
import numpy as np
import pytest

from tests.scenarios import relerr

pytestmark = pytest.mark.gpu


def _small(n, seed, am=(2, 3), rate=None, outputs=True):
    from vpfhf_amd import serve
    return serve.make_requests(n, rate=rate, seed=seed, outputs=outputs, lo=8, hi=16, am_range=am)


def test_serve_alone_equals_mix():
    from vpfhf_amd import serve
    R = _small(5, seed=3)
    res, st = serve.serve(R, max_slots=3, realtime=False)
    assert all(r["status"] == 0 for r in res), res
    assert st["completed"] == 5 and st["max_resident"] == 3
    for i in (0, 2, 4):
        A = serve.RequestSet([R.specs[i]], [R.ams[i]], [0.], outputs=True)
        ra, _ = serve.serve(A, max_slots=1, realtime=False)
        assert ra[0]["am_iterations"] == res[i]["am_iterations"]
        assert ra[0]["u_its"] == res[i]["u_its"] and ra[0]["v_its"] == res[i]["v_its"]
        assert np.array_equal(A.U[0], R.U[i]) and np.array_equal(A.V[0], R.V[i])


@pytest.mark.parametrize("mixed", [True, False])
def test_serve_matches_single_problem_path(mixed):
    from vpfhf_amd import serve, workloads
    R = _small(3, seed=5, am=(2, 2))
    res, _ = serve.serve(R, max_slots=4, realtime=False, mixed=mixed)
    tol = 1e-6 if mixed else 1e-8
    for i, s in enumerate(R.specs):
        H, _ = workloads.make_host(s, u_rtol=1e-10)
        info = H.altmin(maxit=R.ams[i], stop_on_tol=True)
        U, V, _ = H.state()
        en = H.energies()
        H.close()
        assert res[i]["status"] == 0 and info.iterations == res[i]["am_iterations"]
        assert relerr(R.U[i], U) < tol, (i, relerr(R.U[i], U))
        assert relerr(R.V[i], V) < tol, (i, relerr(R.V[i], V))
        e = res[i]["energies"]
        assert abs(e[3] - en["surface"]) <= tol * abs(en["surface"])
        assert abs(e[4] - en["crack_volume"]) <= tol * abs(en["crack_volume"]) + 1e-18


def test_serve_runs_every_iteration_at_altmintol_zero():
    """With -altmintol 0 a request runs every drawn iteration (the do-while of
    VFCommon.c:1192-1231 stops only at its cap).  The later iterations start
    u-solves next to their solution, whose rtol * |r0| lies below the FP64
    residual's rounding floor: the refinement stops there (an outer step that
    no longer halves |r|) instead of failing at its step cap.  Every request
    finishes with status 0 after its drawn count, and the mixed engine still
    agrees with the single-problem FP64 path run for the same count."""
    from vpfhf_amd import serve, workloads
    R = serve.make_requests(4, seed=17, outputs=True, lo=8, hi=14, am_range=(6, 9), altmintol=0.0)
    res, _ = serve.serve(R, max_slots=4, realtime=False)
    for i, s in enumerate(R.specs):
        assert res[i]["status"] == 0, (i, res[i])
        assert res[i]["am_iterations"] == R.ams[i], (i, res[i]["am_iterations"], R.ams[i])
    H, _ = workloads.make_host(R.specs[0], u_rtol=1e-10)
    info = H.altmin(maxit=R.ams[0], stop_on_tol=True)
    U, V, _ = H.state()
    H.close()
    assert info.status == 0 and info.iterations == R.ams[0]
    assert relerr(R.U[0], U) < 1e-6 and relerr(R.V[0], V) < 1e-6


def test_serve_graph_and_plain_launches_agree():
    from vpfhf_amd import serve
    R1 = _small(3, seed=9)
    R2 = _small(3, seed=9)
    serve.serve(R1, max_slots=2, realtime=False, use_graph=True, graph_chunk=3)
    serve.serve(R2, max_slots=2, realtime=False, use_graph=False, graph_chunk=3)
    for a, b in zip(R1.U + R1.V, R2.U + R2.V):
        assert np.array_equal(a, b)


def test_serve_budget_admission():
    from vpfhf_amd import serve
    from vpfhf_amd.host import load
    L = load()
    R = _small(4, seed=7, am=(1, 1), outputs=False)
    need = [L.vfh_request_bytes(s["n"] + 1, s["n"] + 1, s["n"] + 1) for s in R.specs]
    two = min(a + b for i, a in enumerate(need) for b in need[i + 1:])
    budget = max(max(need), two - 1)
    res, st = serve.serve(R, max_slots=8, hbm_budget=budget, realtime=False)
    assert all(r["status"] == 0 for r in res)
    assert st["max_resident"] == 1 and st["peak_bytes"] <= budget
    for r, nd in zip(res, need):
        assert 0 < r["device_bytes"] <= nd  # the admission estimate covers what the request holds
    res2, st2 = serve.serve(R, hbm_budget=min(need) * 0.5, realtime=False)
    assert all(r["status"] == -7 for r in res2) and st2["rejected"] == 4 and st2["completed"] == 0


def test_serve_poisson_stream():
    from vpfhf_amd import serve
    R = _small(6, seed=11, am=(1, 2), rate=200.0, outputs=False)
    res, st = serve.serve(R, max_slots=4, realtime=True)
    assert st["completed"] == 6 and st["solves"] == 2 * sum(r["am_iterations"] for r in res)
    for r, t in zip(res, R.arr):
        assert r["status"] == 0
        assert r["t_admit"] >= t.arrival - 1e-9 and r["t_done"] >= r["t_admit"] and r["latency"] > 0
    assert 0 < st["p50"] <= st["p99"] <= st["max_latency"]


def test_serve_failing_request_is_isolated():
    """A request whose pressure field holds a NaN stops with VFG_E_NAN; the
    others in the same mix finish and are bitwise equal to their runs alone."""
    from vpfhf_amd import serve
    R = _small(4, seed=13, am=(2, 2))
    bad = 1
    p = R.keep[5 * bad + 4]  # payload per request: options, E, theta, thetaRef, pressure
    p[len(p) // 2] = float("nan")  # the C loop reads this array in place
    res, st = serve.serve(R, max_slots=4, realtime=False)
    assert res[bad]["status"] == -6
    for i in (0, 2, 3):
        assert res[i]["status"] == 0
        A = serve.RequestSet([R.specs[i]], [R.ams[i]], [0.], outputs=True)
        serve.serve(A, max_slots=1, realtime=False)
        assert np.array_equal(A.U[0], R.U[i]) and np.array_equal(A.V[0], R.V[i])


def test_serve_at2_requests_match_single_problem_path():
    from vpfhf_amd import serve, workloads
    sa = [serve.spec_request(i, seed=17, lo=8, hi=14, am_range=(2, 2)) for i in range(2)]
    for s, _ in sa:
        s["atnum"] = 2
    R = serve.RequestSet([s for s, _ in sa], [a for _, a in sa], [0., 0.], outputs=True)
    res, _ = serve.serve(R, max_slots=2, realtime=False)
    for i, (s, am) in enumerate(sa):
        H, _ = workloads.make_host(s, u_rtol=1e-10)
        info = H.altmin(maxit=am, stop_on_tol=True)
        U, V, _ = H.state()
        H.close()
        assert res[i]["status"] == 0 and info.iterations == res[i]["am_iterations"]
        assert relerr(R.U[i], U) < 1e-6 and relerr(R.V[i], V) < 1e-6


def test_serve_field_outputs():
    """Nodal crack opening (fields->VolCrackOpening, VFPermfield.c:144-146) and
    per-cell energy terms returned with each request equal those of the same
    final state loaded into a single-problem context, and the energy terms sum
    to the request's reported energies."""
    from vpfhf_amd import serve, workloads
    R = serve.make_requests(3, seed=21, outputs=True, fields=True, lo=8, hi=16, am_range=(2, 2))
    res, _ = serve.serve(R, max_slots=3, realtime=False)
    for i, s in enumerate(R.specs):
        assert res[i]["status"] == 0
        H, _ = workloads.make_host(s)
        H.set_state(U=R.U[i], V=R.V[i])
        assert np.array_equal(H.crack_opening_field(), R.opening[i])
        assert np.array_equal(H.energy_fields(), R.energy[i])
        H.close()
        e = res[i]["energies"]
        assert abs(R.energy[i][:, 0].sum() - e[0]) <= 1e-12 * abs(e[0])
        assert abs(R.energy[i][:, 2].sum() - e[3]) <= 1e-12 * abs(e[3])
        assert np.all(np.isfinite(R.opening[i])) and R.opening[i].shape == R.V[i].shape


def test_context_field_writers(tmp_path):
    """FieldsBinaryWrite / FieldsVTKWrite of a solved context: the records in
    the reference's order (U, velocity, V, pmult, theta, pressure,
    VolCrackOpening, VolLeakOffRate; VFCommon.c:1363-1370) hold the context's
    state bit for bit, and the .vts arrays hold the same values by component."""
    from vpfhf_amd import host, serve, workloads
    R = serve.make_requests(1, seed=4, outputs=True, lo=9, hi=12, am_range=(2, 2))
    s = R.specs[0]
    H, f = workloads.make_host(s)
    H.altmin(2)
    U, V, _ = H.state()
    op = H.crack_opening_field()
    E = H.energy_fields()
    nn, ne = H.nn, H.ne
    name = H.fields_binary_write(str(tmp_path / "run"), 3, theta=f["theta"], pressure=f["pressure"])
    assert name.endswith("run.00003.bin")
    r = host.petsc_vec_read(name)
    assert [x.size for x in r] == [3 * nn, 3 * nn, nn, ne, nn, nn, nn, nn]
    assert np.array_equal(r[0], U) and np.array_equal(r[2], V) and np.array_equal(r[6], op)
    assert np.array_equal(r[4], f["theta"]) and np.array_equal(r[5], f["pressure"])
    assert not r[1].any() and not r[3].any() and not r[7].any()
    H.coordinates_write(str(tmp_path / "run.bin"))
    X = host.petsc_vec_read(str(tmp_path / "run.bin"))[0].reshape(-1, 3)
    g = H.opts.grid
    assert X.shape == (nn, 3) and X[-1].tolist() == [g.lx, g.ly, g.lz] and X[1, 0] == g.lx / (g.nx - 1)
    nod, cel = H.fields_vts_write(str(tmp_path / "run"), 3, theta=f["theta"])
    from tests.test_fieldio import _vts
    _, _, pts, a = _vts(nod)
    assert np.array_equal(pts, X)
    for c in range(3):
        assert np.array_equal(a[("PointData", "Displacement_%d" % c)], U[c::3])
    assert np.array_equal(a[("PointData", "Fracture")], V)
    assert np.array_equal(a[("PointData", "Temperature")], f["theta"])
    assert not a[("PointData", "Pressure")].any()  # pressure not given: zeros
    assert np.array_equal(a[("PointData", "Volumetric Crack Opening")], op)
    _, _, _, b = _vts(cel)
    for c, k in enumerate(("Elastic Energy", "Pressure Work", "Surface Energy")):
        assert np.array_equal(b[("CellData", k)], E[:, c])
    H.close()


def test_serve_multigrid_and_jacobi_engines_agree():
    """The mixed engine's u-CG under each request's batched FP32 V-cycle (the
    default) and under Jacobi reach the same results, within the solvers'
    tolerances, in fewer u-iterations; requests of several sizes share the
    batched V-cycle while they join and leave."""
    from vpfhf_amd import serve, _lib
    Rm = serve.make_requests(6, seed=13, outputs=True, lo=10, hi=28, am_range=(2, 3))
    Rj = serve.make_requests(6, seed=13, outputs=True, lo=10, hi=28, am_range=(2, 3))
    rm, stm = serve.serve(Rm, max_slots=4, realtime=False, pc=_lib.PC_MG)
    rj, _ = serve.serve(Rj, max_slots=4, realtime=False, pc=_lib.PC_JACOBI)
    for a, b in zip(rm, rj):
        assert a["status"] == 0 and b["status"] == 0
        assert a["u_its"] < b["u_its"], (a["u_its"], b["u_its"])
    for i in range(6):
        assert relerr(Rm.U[i], Rj.U[i]) < 1e-6, (i, relerr(Rm.U[i], Rj.U[i]))
        assert relerr(Rm.V[i], Rj.V[i]) < 1e-6, (i, relerr(Rm.V[i], Rj.V[i]))
    assert stm["completed"] == 6 and stm["max_resident"] == 4
