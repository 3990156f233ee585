"""Benchmark: GDoF/s of the damaged-elasticity apply (MatShell MatMult of the
u-Jacobian) on one MI355X, plus u/V solve throughput, with the roofline of the
apply kernel and the reference algorithm's CPU path timed on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu] [--no-solves]

Workload (N=1): 128^3 Q1 hex elements (6.44 M displacement dof), lognormal E
(log-std 0.5, correlation 8 cells), 3 random penny cracks, smooth pore
pressure in [0, 2e-3], uniform temperature rise, clamped faces, FP64.  A step
is one apply y = A_u x over the whole grid; inputs are resident in HBM and
the x/y pairs rotate over 4 buffers (>256 MB) so the last-level cache cannot
serve them.  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (dense)
# FP64 operations of the apply's element routine (u_elem_apply_lds), counted
# in the compiled kernel's AMD assembly: 978 FMA + 123 MUL + 261 ADD per element
FP64_FLOPS_PER_ELEMENT = 2 * 978 + 123 + 261


def apply_bytes(nn, ne, fp32=False, hetero=True, stored=False):
    """Algorithmic HBM bytes of one apply.  Recomputing form: read x and V,
    write y (per node), read E (per element).  Stored form: read x, write y
    (per node), read the 27 stored weights W_g s_g E_e (per element)."""
    w = 4 if fp32 else 8
    if stored:
        return 6 * w * nn + 27 * w * ne
    return (3 * w + 3 * w + w) * nn + (8 if hetero else 0) * ne


def read_traffic(name):
    """HBM bytes per launch of the apply kernel from a committed counter
    summary: the line 'hbm_bytes_per_launch: <int>'."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("hbm_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def copy_calibration(L, h, n=1 << 26, reps=4):
    """Plain streaming kernel of known size for counter calibration:
    out = min(a, b) over n doubles reads 16 n and writes 8 n bytes."""
    import vpfhf_amd as V
    bufs = [V.DeviceBuffer(8 * n) for _ in range(3)]
    for _ in range(reps):
        V.check(L.vfg_vec_pointwise_min(h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n), h)
    V.check(L.vfg_synchronize(h), h)
    for b in bufs:
        b.free()


def cpu_model():
    """Model name of the host CPU with its family / model / stepping numbers
    (first processor of /proc/cpuinfo; a virtual machine may report a short
    name)."""
    try:
        info = {}
        for line in open("/proc/cpuinfo"):
            if not line.strip():
                break
            k, _, v = line.partition(":")
            info.setdefault(k.strip(), v.strip())
        if "model name" in info:
            return "%s (family %s, model %s, stepping %s)" % (info["model name"], info.get("cpu family", "?"),
                                                            info.get("model", "?"), info.get("stepping", "?"))
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(n_sample=48, seconds=6.0):
    """The reference's per-iteration kernel on the host: the assembled AIJ
    u-Jacobian (VF_UIJacobian + MatApplyDirichletBC, 81 nnz/row) times a
    vector, from the oracle's C restatement, OpenMP over rows; timed on every
    core the process may use and on one core."""
    from oracle import vforacle as O
    from vpfhf_amd import inputs
    L = O.lib()
    n = n_sample
    E = inputs.lognormal_E((n, n, n), .5, 8., 0)
    bc = [O.make_bc({f: O.ZERO for f in range(6)})] * 3
    op = O.OracleProblem((n + 1,) * 3, E=1., nu=.2, beta=1., atnum=1, epsilon=2. / n, pressurize=True,
                         E_elem=E, bcU=bc)
    V = O.penny_v(op, (.5, .5, .5), .2, 0., 0., 3. / n)
    F = op.fields(V=V, VIrrev=V, theta=np.zeros(op.nn), thetaRef=np.zeros(op.nn), pressure=np.full(op.nn, 1e-3))
    nt = int(L.vfo_num_threads())
    t0 = time.time()
    A = O.u_jacobian(op, F, csr=False)
    t_asm = time.time() - t0
    x = np.random.default_rng(0).standard_normal(3 * op.nn)

    def rate(threads):
        L.vfo_set_num_threads(threads)
        O.ell_matvec(A, x)
        reps, t0 = 0, time.time()
        while time.time() - t0 < seconds:
            O.ell_matvec(A, x)
            reps += 1
        return 3 * op.nn / ((time.time() - t0) / reps) / 1e9, reps
    v_all, r_all = rate(nt)
    v_one, r_one = rate(1)
    L.vfo_set_num_threads(nt)
    L.vfo_ell_free(A)
    return {"value": round(v_all, 4), "unit": "GDoF/s", "cores": nt, "kind": "port",
            "sample": "assembled AIJ MatMult (81 nnz/row, FP64) of the %d^3-element u-Jacobian, %d products on %d "
                      "threads and %d on one, after a %.1f s assembly; oracle C restatement of VF_UIJacobian"
                      % (n, r_all, nt, r_one, t_asm),
            "one_core": {"value": round(v_one, 4), "unit": "GDoF/s", "cores": 1},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def time_apply(L, h, nn, steps, warmup, fp32=False, nbuf=4):
    """Seconds per u-apply from HIP events on the handle's stream over `steps`
    back-to-back launches after `warmup` untimed ones; the (x, y) pairs rotate
    over nbuf buffers (>= 4 x 103 MB at 128^3) so the 256 MB last-level cache
    cannot hold the working set between launches."""
    import vpfhf_amd as V
    from vpfhf_amd import _lib
    w = 4 if fp32 else 8
    rng = np.random.default_rng(1)
    dt = np.float32 if fp32 else np.float64
    xs = [V.DeviceBuffer(3 * w * nn).upload(rng.standard_normal(3 * nn).astype(dt)) for _ in range(nbuf)]
    ys = [V.DeviceBuffer(3 * w * nn) for _ in range(nbuf)]
    xa = (C.c_void_p * nbuf)(*[b.ptr for b in xs])
    ya = (C.c_void_p * nbuf)(*[b.ptr for b in ys])
    ms = C.c_double()
    _lib.check(L.vfg_synchronize(h), h)
    _lib.check(L.vfg_time_u_apply(h, xa, ya, nbuf, max(warmup, 1), int(fp32), C.byref(ms)), h)
    _lib.check(L.vfg_time_u_apply(h, xa, ya, nbuf, steps, int(fp32), C.byref(ms)), h)
    for b in xs + ys:
        b.free()
    return ms.value * 1e-3 / steps


def apply_at(n, steps, warmup, seed=0):
    """The FP64 u-apply of the same workload generators at n^3 elements
    (256^3: the precision form's grid, whose working set is 4 x the last-level
    cache), with its algorithmic HBM fraction."""
    from vpfhf_amd import _lib, workloads
    H, _ = workloads.make_host(workloads.spec_hetero(n, seed=seed), u_rtol=1e-10)
    L = _lib.load()
    try:
        out = {"elements": H.ne, "dof": 3 * H.nn, "forms": time_forms(L, H.h, H.nn, H.ne, steps, warmup)}
        out.update(out["forms"]["recompute"])
        return out
    finally:
        H.close()


def time_forms(L, h, nn, ne, steps, warmup):
    """Both forms of the FP64 apply on one handle (the recomputing form, the
    library default, last, so the handle is left in it)."""
    from vpfhf_amd import _lib
    forms = {}
    for name, form in (("stored", _lib.U_FORM_STORED), ("recompute", _lib.U_FORM_RECOMPUTE)):
        _lib.check(L.vfg_u_set_form(h, form), h)
        t = time_apply(L, h, nn, steps, warmup)
        b = apply_bytes(nn, ne, stored=form == _lib.U_FORM_STORED)
        forms[name] = {"form": name, "ms": round(t * 1e3, 4), "gdofs": round(3 * nn / t / 1e9, 3),
                       "algorithmic_bytes": b, "achieved_GBs": round(b / t / 1e9, 1),
                       "frac": round(b / t / 1e9 / PEAK_HBM_GBS, 4)}
    return forms


def _config1_problem():
    """The 32^3-element penny case with crack pressure only (beta = 0; with
    beta = 1 the pore and crack loads of a uniform p cancel and U = 0)."""
    from oracle import vforacle as O
    O.lib()
    n, h = 32, 1.0 / 32
    bc = [O.make_bc({f: O.ZERO for f in range(6)})] * 3
    op = O.OracleProblem((n + 1,) * 3, E=1., nu=.2, beta=0., atnum=1, Gc=1., epsilon=2 * h, pressurize=True,
                         bcU=bc)
    V = O.penny_v(op, (.5, .5, .5), .2, 0., 0., 3 * h)
    F = op.fields(V=V, VIrrev=V, theta=np.zeros(op.nn), thetaRef=np.zeros(op.nn), pressure=np.full(op.nn, 1e-3))
    return O, op, F


def cpu_solves():
    """The reference's u-solve on the host at the 32^3-element penny case:
    the oracle's C restatement assembles the u-Jacobian (VF_UIJacobian) and
    solves it with Jacobi-CG to rtol 1e-10 (one Newton step of
    SNESNEWTONLS on a linear problem), OpenMP over rows."""
    O, op, F = _config1_problem()
    U = np.zeros(3 * op.nn)
    t0 = time.time()
    its, _ = O.u_solve(op, F, U, np.zeros(3 * op.nn), rtol=1e-10)
    tu = time.time() - t0
    return {"u_solve_s": round(tu, 4), "u_cg_iterations": int(its), "u_solves_per_s": round(1.0 / tu, 3),
            "cores": int(O.lib().vfo_num_threads()), "kind": "port",
            "sample": "32^3 elements, one penny crack, p = 1e-3 on the crack faces; assembly included, as the "
                      "reference assembles the Jacobian at every Newton step"}, U


def gpu_solves_config1(U_cpu=None):
    """The same u-solve on the GPU (vfg_u_solve, multigrid-PCG to 1e-10, the
    hierarchy set up inside the timed call), after one untimed solve that
    loads the kernels; and one whole alternate-minimisation iteration of the
    case through the C host."""
    from vpfhf_amd import workloads
    spec = workloads.spec_penny32(beta=0.0)
    P, _ = workloads.make_problem(spec)
    Z = np.zeros(3 * P.nn)
    P.VF_StepU(Z, Z, rtol=1e-10)
    P.close()
    P, _ = workloads.make_problem(spec)
    t0 = time.time()
    Ug, info = P.VF_StepU(Z, Z, rtol=1e-10)
    tu = time.time() - t0
    P.close()
    out = {"u_solve_s": round(tu, 4), "u_cg_iterations": info.iterations, "u_solves_per_s": round(1.0 / tu, 2)}
    if U_cpu is not None:
        out["rel_diff_U_vs_cpu"] = _relerr(Ug, U_cpu)
    H, _ = workloads.make_host(spec, u_rtol=1e-10)
    t0 = time.time()
    ia = H.altmin(maxit=1)
    out["altmin_iteration_s"] = round(time.time() - t0, 4)
    out["altmin_v_newton"] = ia.v_outer
    H.close()
    return out


def _relerr(a, b):
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (d if d > 0 else 1.0)


def ensemble_bench(L, n_members=64, seed=0):
    """Batching form: 64 independent problems, sides log-uniform in [16, 64]
    elements, own lognormal E, 1-4 cracks, injection pressure.  The plain FP64
    GPU drop-in solves them one after another (vfg_u_solve, Jacobi-PCG to
    1e-10; also timed with the multigrid preconditioner); the batched forms
    solve all at once (persistent applies, graph-captured CG, autotuned
    tiles): FP64 CG with Jacobi, FP32 CG inside FP64 refinement with Jacobi or
    with the members' V-cycles batched, and FP64 CG under the batched FP32
    V-cycles.  Each path runs once untimed, then timed from the same initial
    state."""
    import ctypes as C
    from vpfhf_amd import _lib, workloads
    from vpfhf_amd.batch import Batch
    hosts = [workloads.make_host(workloads.spec_ensemble_member(i, seed), u_rtol=1e-10)[0]
             for i in range(n_members)]
    ptr = [H.device_fields() for H in hosts]
    dof = sum(3 * H.nn for H in hosts)

    def reset():
        for H, (U, _, _, _) in zip(hosts, ptr):
            _lib.check(L.vfg_vec_set(H.h, U, 0., 3 * H.nn), H.h)
        for H in hosts:
            _lib.check(L.vfg_synchronize(H.h), H.h)

    def download():
        out = []
        for H, (U, _, _, _) in zip(hosts, ptr):
            a = np.empty(3 * H.nn)
            _lib.check(L.vfg_memcpy(a.ctypes.data, U, a.nbytes, 1))
            out.append(a)
        return out

    o64 = _lib.SolveOpts(1e-10, 0., 100000, _lib.PC_JACOBI, _lib.PREC_FP64, 4)
    info = _lib.SolveInfo()

    def single():
        for H, (U, B, _, _) in zip(hosts, ptr):
            _lib.check(L.vfg_u_solve(H.h, U, B, C.byref(o64), C.byref(info), _lib.DEVICE), H.h)
    reset(); single(); reset()
    t0 = time.time(); single(); t64 = time.time() - t0
    U64 = download()
    o64.pc = _lib.PC_MG
    reset(); single(); reset()
    t0 = time.time(); single(); tmg = time.time() - t0
    Umg = download()
    o64.pc = _lib.PC_JACOBI

    bt = Batch(hosts)
    apply_ms = bt.autotune(fp32=True)
    pick, points = bt.tuning(), bt.tune_points()
    tuned = {"schedules": Batch.SCHEDULES, "points": len(points), "pick": pick,
             "best_per_schedule_ms": {str(sc): round(min(p["ms"] for p in points if p["schedule"] == sc), 4)
                                      for sc in sorted({p["schedule"] for p in points})},
             "slowest_ms": round(max(p["ms"] for p in points), 4)}
    bt.set_graph(True, 10)
    Ua = (C.c_void_p * n_members)(*[p[0] for p in ptr])
    Ba = (C.c_void_p * n_members)(*[p[1] for p in ptr])
    infos = (_lib.SolveInfo * n_members)()
    res = {}
    for key, mixed, pc in (("fp64", False, _lib.PC_JACOBI), ("mixed", True, _lib.PC_JACOBI),
                           ("mixed_mg", True, _lib.PC_MG), ("fp64_mg", False, _lib.PC_MG)):
        om = _lib.SolveOpts(1e-10, 0., 100000, pc, _lib.PREC_MIXED if mixed else _lib.PREC_FP64, 4)

        def batched():
            rc = L.vfg_batch_u_solve(bt.b, Ua, Ba, C.byref(om), infos, _lib.DEVICE)
            bt._c(rc)
        reset(); batched(); reset()
        t0 = time.time(); batched(); tb = time.time() - t0
        Ub = download()
        res[key] = (tb, max(_relerr(a, b) for a, b in zip(Ub, U64)),
                    {"inner_iterations_max": max(infos[i].iterations for i in range(n_members)),
                     "outer_max": max(infos[i].outer for i in range(n_members))})
    bt.close()
    for H in hosts:
        H.close()
    return {"members": n_members, "dof": dof, "elements": sum(s ** 3 for s in
                                                            [workloads.spec_ensemble_member(i, seed)["n"]
                                                             for i in range(n_members)]),
            "fp64_single_solves_per_s": round(n_members / t64, 2),
            "fp64_single_multigrid_solves_per_s": round(n_members / tmg, 2),
            "max_rel_err_multigrid_vs_jacobi": max(_relerr(a, b) for a, b in zip(Umg, U64)),
            "batched_fp64_solves_per_s": round(n_members / res["fp64"][0], 2),
            "batched_mixed_solves_per_s": round(n_members / res["mixed"][0], 2),
            "batched_mixed_multigrid_solves_per_s": round(n_members / res["mixed_mg"][0], 2),
            "speedup_mixed_batched_vs_fp64_single": round(t64 / res["mixed"][0], 2),
            "speedup_mixed_multigrid_batched_vs_fp64_single": round(t64 / res["mixed_mg"][0], 2),
            "speedup_mixed_multigrid_batched_vs_multigrid_single": round(tmg / res["mixed_mg"][0], 2),
            "batched_fp64_multigrid_solves_per_s": round(n_members / res["fp64_mg"][0], 2),
            "speedup_fp64_multigrid_batched_vs_fp64_single": round(t64 / res["fp64_mg"][0], 2),
            "speedup_fp64_multigrid_batched_vs_multigrid_single": round(tmg / res["fp64_mg"][0], 2),
            "max_rel_err_fp64_multigrid_batched_vs_fp64": res["fp64_mg"][1],
            "max_rel_err_mixed_vs_fp64": res["mixed"][1], "max_rel_err_batched_fp64_vs_single": res["fp64"][1],
            "max_rel_err_mixed_multigrid_vs_fp64": res["mixed_mg"][1],
            "iterations": {k: v[2] for k, v in res.items()},
            "persistent_fp32_apply_ms": round(apply_ms, 4), "autotune": tuned}


def precision_bench(n=256, seed=0):
    """Precision form at n^3 elements: one alternate-minimisation iteration
    with (a) the FP64 u-CG under the FP32 multigrid V-cycle (the default),
    (b) FP32 CG under the same V-cycle inside FP64 refinement, (c) FP32 CG with
    Chebyshev(4)-Jacobi inside FP64 refinement (the method as first specified
    for this form); all to 1e-10.  Relative errors of (b) and (c) in U, V,
    crack volume and surface energy against (a)."""
    from vpfhf_amd import workloads
    spec = workloads.spec_hetero(n, seed=seed)
    out = {}
    for name, extra in (("fp64", ""), ("mixed", "-vfg_precision mixed"),
                        ("mixed_chebyshev", "-vfg_precision mixed -vfg_pc chebyshev")):
        H, _ = workloads.make_host(spec, u_rtol=1e-10, extra=extra)
        t0 = time.time()
        info = H.altmin(maxit=1)
        dt = time.time() - t0
        U, V, _ = H.state()
        out[name] = (dt, info.u_its, U, V, H.energies())
        H.close()
    d64 = out["fp64"]
    e64 = d64[4]
    res = {"elements": n ** 3, "dof": 3 * (n + 1) ** 3, "fp64_seconds": round(d64[0], 3),
           "fp64_u_cg_iterations": d64[1]}
    for name in ("mixed", "mixed_chebyshev"):
        dm = out[name]
        em = dm[4]
        res.update({name + "_seconds": round(dm[0], 3), name + "_u_inner_iterations": dm[1],
                    name + "_rel_err_U": _relerr(dm[2], d64[2]), name + "_rel_err_V": _relerr(dm[3], d64[3]),
                    name + "_rel_err_crack_volume":
                        abs(em["crack_volume"] - e64["crack_volume"]) / abs(e64["crack_volume"]),
                    name + "_rel_err_surface_energy": abs(em["surface"] - e64["surface"]) / abs(e64["surface"])})
    return res


def serving_bench(rates=(10., 20., 50., 100., 200., 500.), n_req=200, seed=0, n_check=32, max_slots=64):
    """Serving form: a Poisson stream of n_req alternate-minimisation requests
    (sides 24-96 elements, 1-3 cracks, 1-20 iterations each, every drawn
    iteration run: altmintol 0) at each rate, max_slots slots, 200 GB HBM
    budget, FP32 u-CG inside FP64 refinement.  The rates of the workload are
    50-500 requests/s; 10 and 20 are below the device's throughput for this
    mix and show latency without a queue.  The first n_check requests re-run
    on the FP64 engine give the error of the mixed-precision results."""
    from vpfhf_amd import serve
    R = serve.make_requests(n_req, rate=None, seed=seed, outputs=True, altmintol=0.0)
    base_t = serve.arrival_times(n_req, 1.0, seed)  # unit-rate arrivals, scaled per rate
    rows = []
    for k, rate in enumerate(rates):
        for i in range(n_req):
            R.arr[i].arrival = float(base_t[i] / rate)
        t0 = time.time()
        res, st = serve.serve(R, max_slots=max_slots, hbm_budget=200e9, realtime=True)
        if k == 0:
            Um = [u.copy() for u in R.U[:n_check]]
            Vm = [v.copy() for v in R.V[:n_check]]
            ams = [r["am_iterations"] for r in res]
            bad = sum(r["status"] != 0 for r in res)
            statuses = {}
            for r in res:
                statuses[str(r["status"])] = statuses.get(str(r["status"]), 0) + 1
            per_req = {"u_its_mean": round(float(np.mean([r["u_its"] for r in res])), 1),
                       "v_outer_mean": round(float(np.mean([r["v_outer"] for r in res])), 1),
                       "v_its_mean": round(float(np.mean([r["v_its"] for r in res])), 1)}
        rows.append({"rate_per_s": rate, "p50_s": round(st["p50"], 4), "p99_s": round(st["p99"], 4),
                     "mean_s": round(st["mean_latency"], 4), "solves_per_s": round(st["solves_per_s"], 2),
                     "wall_s": round(st["wall"], 3), "max_resident": st["max_resident"], "solves": st["solves"],
                     "completed": st["completed"]})
        print("bench: serving at %g req/s done in %.1f s" % (rate, time.time() - t0), file=sys.stderr, flush=True)
    C = serve.RequestSet(R.specs[:n_check], R.ams[:n_check], [0.] * n_check, outputs=True)
    serve.serve(C, max_slots=64, mixed=False, realtime=False)
    err_u = max(_relerr(a, b) for a, b in zip(Um, C.U))
    err_v = max(_relerr(a, b) for a, b in zip(Vm, C.V))
    sides = [s["n"] for s in R.specs]
    return {"requests": n_req, "max_slots": max_slots, "sides": {"min": min(sides), "max": max(sides), "mean": round(float(np.mean(sides)), 1)},
            "am_iterations": {"min": min(ams), "max": max(ams), "mean": round(float(np.mean(ams)), 2),
                              "total": int(sum(ams)),
                              "histogram": [int(sum(1 for a in ams if a == q)) for q in range(1, 21)]},
            "failed": bad, "status_counts": statuses, "per_request": per_req, "rates": rows,
            "checked_vs_fp64": n_check,
            "max_rel_err_U_vs_fp64": err_u, "max_rel_err_V_vs_fp64": err_v}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--n3", type=int, default=256, help="also time the FP64 apply at this grid (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solves", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--calib", action="store_true", help="also run a plain streaming kernel of known size")
    ap.add_argument("--no-forms", action="store_true", help="skip the ensemble, precision and serving forms")
    ap.add_argument("--forms", default="ensemble,precision,serving", help="comma-separated subset of the forms")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-slots", type=int, default=64, help="resident requests of the serving form")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this path runs on one MI355X and has no multi-GPU decomposition "
              "to run (a large problem would split into z-slabs, ensembles and request streams by problem). "
              "Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import vpfhf_amd as V
    from vpfhf_amd import _lib, workloads

    spec = workloads.spec_hetero(a.n, seed=a.seed)
    H, fields = workloads.make_host(spec, u_rtol=1e-10)
    L = _lib.load()
    h = H.h
    nn, ne = H.nn, H.ne
    # both forms of the FP64 apply; the library default (recompute) is the headline
    forms = time_forms(L, h, nn, ne, a.steps, a.warmup)
    t_step = forms["recompute"]["ms"] * 1e-3
    gdofs = 3 * nn / t_step / 1e9
    byt = apply_bytes(nn, ne)
    achieved = byt / t_step / 1e9
    traffic = read_traffic("u_apply_hetero%d_counters.txt" % a.n)
    if a.calib:
        copy_calibration(L, h)
    # FP32 apply (precision form) on the same operator
    t32 = time_apply(L, h, nn, a.steps, a.warmup, fp32=True)

    extra = {}
    if not a.no_solves:
        # one alternate-minimisation iteration (u-solve to rtol 1e-10, then the
        # bound-constrained V-solve) of the same problem
        t0 = time.time()
        info = H.altmin(maxit=1)
        dt = time.time() - t0
        en = H.energies()
        extra["solves"] = {"u_pc": "multigrid", "u_cg_iterations": info.u_its, "v_newton": info.v_outer,
                           "v_cg": info.v_its, "seconds_per_altmin_iteration": round(dt, 3),
                           "solves_per_s": round(2.0 / dt, 3),
                           "surface_energy": en["surface"], "crack_volume": en["crack_volume"]}
        Um, _, _ = H.state()
        # the rest of the workload's 10 alternate-minimisation iterations
        t0 = time.time()
        info9 = H.altmin(maxit=9)
        dt9 = time.time() - t0
        extra["solves"]["ten_iterations"] = {
            "seconds": round(dt + dt9, 3), "solves_per_s": round(20.0 / (dt + dt9), 3),
            "u_cg_iterations": info.u_its + info9.u_its, "v_cg": info.v_its + info9.v_its, "errV": info9.errV}
        # the same iteration with the Jacobi-preconditioned u-CG
        Hj, _ = workloads.make_host(spec, u_rtol=1e-10, extra="-vfg_pc jacobi")
        t0 = time.time()
        ij = Hj.altmin(maxit=1)
        dtj = time.time() - t0
        Uj, _, _ = Hj.state()
        Hj.close()
        extra["solves"].update({"jacobi_u_cg_iterations": ij.u_its,
                                "jacobi_seconds_per_altmin_iteration": round(dtj, 3),
                                "rel_diff_U_multigrid_vs_jacobi": _relerr(Um, Uj)})
        # the V-Jacobian apply of the same problem (coupling set by the V-solve)
        nbuf, rng, ms = 4, np.random.default_rng(2), C.c_double()
        vx = [V.DeviceBuffer(8 * nn).upload(rng.standard_normal(nn)) for _ in range(nbuf)]
        vy = [V.DeviceBuffer(8 * nn) for _ in range(nbuf)]
        va = (C.c_void_p * nbuf)(*[b.ptr for b in vx])
        vb = (C.c_void_p * nbuf)(*[b.ptr for b in vy])
        _lib.check(L.vfg_time_v_apply(h, va, vb, nbuf, max(a.warmup, 1), C.byref(ms)), h)
        _lib.check(L.vfg_time_v_apply(h, va, vb, nbuf, a.steps, C.byref(ms)), h)
        tv = ms.value * 1e-3 / a.steps
        vbytes = 16 * nn + 216 * ne
        extra["v_apply"] = {"ms": round(tv * 1e3, 4), "algorithmic_bytes": vbytes,
                            "achieved_GBs": round(vbytes / tv / 1e9, 1),
                            "frac_of_hbm_peak": round(vbytes / tv / 1e9 / PEAK_HBM_GBS, 4)}
        for b in vx + vy:
            b.free()

    out = {
        "metric": "GDoF/s elasticity apply; solves/s and p99 latency at 1e-6 rel err vs FP64",
        "value": round(gdofs, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal E, penny cracks, pressure field)",
        "config": {"workload": "u-Jacobian MatShell apply, %d^3 Q1 hex elements, %d displacement dof, FP64 "
                               "plain GPU drop-in, lognormal E, 3 penny cracks, clamped" % (a.n, 3 * nn),
                   "elements": ne, "dof": 3 * nn},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic},
        "apply_forms": forms,
        # the compute roofline that bounds this kernel: FP64 vector issue
        "valu": {"bound": "fp64-valu", "flops_per_element": FP64_FLOPS_PER_ELEMENT,
                 "achieved": round(FP64_FLOPS_PER_ELEMENT * ne / t_step / 1e12, 2), "peak": PEAK_FP64_TFLOPS,
                 "unit": "TFLOP/s",
                 "frac": round(FP64_FLOPS_PER_ELEMENT * ne / t_step / 1e12 / PEAK_FP64_TFLOPS, 4)},
        "fp32_apply": {"ms": round(t32 * 1e3, 4), "gdofs": round(3 * nn / t32 / 1e9, 3),
                       "achieved_GBs": round(apply_bytes(nn, ne, fp32=True) / t32 / 1e9, 1)},
    }
    if a.n3 > 0:
        out["apply_%d" % a.n3] = apply_at(a.n3, a.steps, a.warmup, a.seed)
    out.update(extra)
    H.close()
    if not a.no_forms:
        for key, fn in (("ensemble", lambda: ensemble_bench(L)), ("precision", precision_bench),
                        ("serving", lambda: serving_bench(n_req=a.serve_requests, max_slots=a.serve_slots))):
            if key not in a.forms.split(","):
                continue
            t0 = time.time()
            out[key] = fn()
            print("bench: %s form done in %.1f s" % (key, time.time() - t0), file=sys.stderr, flush=True)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    if not a.no_solves:
        # u-solves of the 32^3-element penny case: GPU drop-in against the
        # reference's assemble-and-CG path on the host
        c1 = {}
        U_cpu = None
        if not a.no_cpu:
            c1["cpu"], U_cpu = cpu_solves()
        c1["gpu"] = gpu_solves_config1(U_cpu)
        if "cpu" in c1:
            c1["speedup_u_solve"] = round(c1["cpu"]["u_solve_s"] / c1["gpu"]["u_solve_s"], 1)
        out["config1_solves"] = c1
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
