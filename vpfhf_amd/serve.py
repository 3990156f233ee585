"""Serving form: a seeded Poisson stream of alternate-minimisation requests run
by the continuous-batching engine (``vfh_serve`` in libvfhost.so, over
``vfg_engine_*`` in libvfg.so).

A request is a complete problem -- the reference's option text (grid,
material, penny cracks, BCfiles clamped faces) plus its input fields
(per-element E, temperature, pressure) -- and a number of alternate-
minimisation iterations.  ``make_requests`` draws the stream: side
log-uniform in [lo, hi] elements, lognormal E with log-std in [0.2, 0.8],
1-3 cracks, uniform injection pressure in [5e-4, 5e-3] on the crack faces
(Biot beta = 0), 1-20 iterations, exponential
inter-arrival gaps at a given rate; all from the counter RNG, so a request is
the same bytes in every stream it appears in.
"""
import ctypes as C
import math

import numpy as np

from . import _lib, inputs, workloads
from .host import Request, RequestResult, ServeOpts, ServeStats, load as load_host


def spec_request(i, seed=0, lo=24, hi=96, am_range=(1, 20), max_cracks=3, altmintol=None):
    """Spec and iteration count of request i of stream `seed`.  altmintol
    (the reference's -altmintol, default 1e-4) ends a request's loop early
    once ||V - Vold||_inf falls below it; 0 runs every drawn iteration (the
    do-while of VFCommon.c:1192-1231 then stops at its iteration cap)."""
    u = inputs.uniform(seed, 5000 + i, 5)
    n = int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u[0])))
    sigma = 0.2 + 0.6 * u[1]
    nc = min(max_cracks, 1 + int(max_cracks * u[2]))
    am = min(am_range[1], am_range[0] + int((am_range[1] - am_range[0] + 1) * u[3]))
    p = 5e-4 + (5e-3 - 5e-4) * u[4]
    s = workloads.spec_hetero(n, seed=seed * 100000 + i, sigma=sigma, ncracks=nc, dtheta=0.0)
    s["pressure"] = ("const", p)
    s["beta"] = 0.0  # injection pressure on the crack faces (a uniform Biot load would cancel it)
    s["name"] = "request%d" % i
    if altmintol is not None:
        s["altmintol"] = float(altmintol)
    return s, am


def arrival_times(n, rate, seed=0):
    """Poisson process: exponential gaps of mean 1/rate from the counter RNG."""
    u = inputs.uniform(seed, 4999, n)
    return np.cumsum(-np.log1p(-u) / rate)


class RequestSet:
    """Host payloads of a request stream (kept alive for the C loop)."""

    def __init__(self, specs, ams, arrivals, outputs=False, fields=False):
        self.specs, self.ams = list(specs), list(ams)
        n = len(self.specs)
        self.arr = (Request * n)()
        self.keep = []
        self.U = [None] * n
        self.V = [None] * n
        self.opening = [None] * n
        self.energy = [None] * n
        for i, (s, am, t) in enumerate(zip(self.specs, self.ams, arrivals)):
            f = workloads.build_fields(s)
            opt = workloads.options_text(s).encode()
            arrays = [np.ascontiguousarray(f[k], np.float64) if f.get(k) is not None else None
                      for k in ("E_elem", "theta", "thetaRef", "pressure")]
            self.keep += [opt] + arrays
            r = self.arr[i]
            r.options = opt
            r.E_elem, r.theta, r.thetaRef, r.pressure = [None if a is None else a.ctypes.data for a in arrays]
            r.arrival = float(t)
            r.am_maxit = int(am)
            if outputs:
                nn = (s["n"] + 1) ** 3
                self.U[i] = np.zeros(3 * nn)
                self.V[i] = np.zeros(nn)
                r.U_out = self.U[i].ctypes.data
                r.V_out = self.V[i].ctypes.data
            if fields:
                self.opening[i] = np.zeros((s["n"] + 1) ** 3)
                self.energy[i] = np.zeros((s["n"] ** 3, 3))
                r.opening_out = self.opening[i].ctypes.data
                r.energy_out = self.energy[i].ctypes.data

    def __len__(self):
        return len(self.specs)


def make_requests(n, rate=None, seed=0, outputs=False, fields=False, **kw):
    """n requests of stream `seed`; rate None: all arrive at t = 0."""
    sa = [spec_request(i, seed, **kw) for i in range(n)]
    t = arrival_times(n, rate, seed) if rate else np.zeros(n)
    return RequestSet([s for s, _ in sa], [a for _, a in sa], t, outputs, fields)


def serve(reqs, max_slots=64, hbm_budget=200e9, mixed=True, realtime=True, u_rtol=1e-10, v_rtol=1e-8,
          graph_chunk=1, use_graph=True, pc=None):
    """Runs the stream; returns (per-request result dicts, stats dict)."""
    L = load_host()
    o = ServeOpts()
    L.vfh_serve_opts_default(C.byref(o))
    o.engine.max_slots = max_slots
    o.engine.precision = _lib.PREC_MIXED if mixed else _lib.PREC_FP64
    o.engine.u_rtol = u_rtol
    o.engine.v_rtol = v_rtol
    o.engine.graph_chunk = graph_chunk
    o.engine.use_graph = int(use_graph)
    if pc is not None:
        o.engine.pc = pc
    o.hbm_budget = hbm_budget
    o.realtime = int(realtime)
    n = len(reqs)
    res = (RequestResult * n)()
    st = ServeStats()
    rc = L.vfh_serve(C.byref(o), reqs.arr, n, res, C.byref(st))
    if rc:
        raise _lib.VfgError(rc, L.vfh_last_error(None).decode())
    out = []
    for i in range(n):
        r = res[i]
        out.append(dict(status=r.status, am_iterations=r.am_iterations, u_its=r.u_its, v_outer=r.v_outer,
                        v_its=r.v_its, errV=r.errV, t_admit=r.t_admit, t_done=r.t_done, latency=r.latency,
                        energies=list(r.energies), device_bytes=r.device_bytes))
    stats = {k: getattr(st, k) for k, _ in ServeStats._fields_}
    return out, stats
