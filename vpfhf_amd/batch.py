"""Ensembles of independent problems (vfg_batch, include/vfg.h).

``Batch(problems)`` takes ``vpfhf_amd.Problem`` or ``vpfhf_amd.host.Host``
members; ``u_solve`` runs VF_StepU for every member together (FP64 or FP32
inside FP64 refinement) and returns the solutions and per-member infos.
"""
import ctypes as C

import numpy as np

from . import _lib

class Batch:
    def __init__(self, members):
        self.L = _lib.load()
        self.members = list(members)
        hs = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        b = C.c_void_p()
        rc = self.L.vfg_batch_create(hs, len(self.members), C.byref(b))
        if rc:
            raise _lib.VfgError(rc, "vfg_batch_create failed")
        self.b = b.value
        self.sizes = [3 * m.nn for m in self.members]

    def _c(self, rc):
        if rc:
            raise _lib.VfgError(rc, self.L.vfg_batch_last_error(self.b).decode())

    def close(self):
        if getattr(self, "b", None):
            self.L.vfg_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, fp32=True):
        ms = C.c_double()
        self._c(self.L.vfg_batch_autotune(self.b, int(fp32), C.byref(ms)))
        return ms.value

    SCHEDULES = {0: "16x16 tiles, register march", 1: "8x8 tiles, register march", 2: "16x16 tiles, LDS-ring march"}

    def tune_points(self):
        """The (schedule, z-chunk scale, grid, ms) points the last autotune timed."""
        n = self.L.vfg_batch_tune_points(self.b, None, 0)
        arr = (_lib.TunePoint * max(n, 1))()
        self.L.vfg_batch_tune_points(self.b, arr, n)
        return [dict(schedule=p.schedule, oz_scale=p.oz_scale, grid=p.grid, ms=p.ms) for p in arr[:n]]

    def tuning(self):
        """The configuration in use: schedule, z-chunk scale, grid, its time."""
        p = _lib.TunePoint()
        self._c(self.L.vfg_batch_tuning(self.b, C.byref(p)))
        return dict(schedule=p.schedule, oz_scale=p.oz_scale, grid=p.grid, ms=p.ms)

    def set_graph(self, enable=True, chunk=10):
        self._c(self.L.vfg_batch_set_graph(self.b, int(enable), chunk))

    def time_apply(self, fp32=False, nrep=20):
        ms = C.c_double()
        self._c(self.L.vfg_batch_time_u_apply(self.b, int(fp32), nrep, C.byref(ms)))
        return ms.value / nrep

    def u_apply(self, xs, fp32=False):
        dt = np.float32 if fp32 else np.float64
        xs = [np.ascontiguousarray(x, dt) for x in xs]
        ys = [np.empty_like(x) for x in xs]
        from . import DeviceBuffer
        dx = [DeviceBuffer(x.nbytes).upload(x) for x in xs]
        dy = [DeviceBuffer(x.nbytes) for x in xs]
        xa = (C.c_void_p * len(xs))(*[d.ptr for d in dx])
        ya = (C.c_void_p * len(xs))(*[d.ptr for d in dy])
        self._c((self.L.vfg_batch_u_apply_f32 if fp32 else self.L.vfg_batch_u_apply)(self.b, xa, ya))
        return [d.download(dt, len(x)) for d, x in zip(dy, xs)]

    def u_solve(self, Us, BCUs, rtol=1e-10, maxit=100000, mixed=False, pc=_lib.PC_JACOBI):
        """Batched u-solve; pc = PC_MG runs the members' FP32 V-cycles batched (under
        the FP64 CG, or the FP32 CG of the mixed refinement)."""
        Us = [np.array(u, np.float64, copy=True) for u in Us]
        Bs = [np.ascontiguousarray(b, np.float64) for b in BCUs]
        ua = (C.c_void_p * len(Us))(*[u.ctypes.data for u in Us])
        ba = (C.c_void_p * len(Bs))(*[b.ctypes.data for b in Bs])
        o = _lib.SolveOpts(rtol, 0.0, maxit, pc, _lib.PREC_MIXED if mixed else _lib.PREC_FP64, 4)
        infos = (_lib.SolveInfo * len(Us))()
        rc = self.L.vfg_batch_u_solve(self.b, ua, ba, C.byref(o), infos, _lib.HOST)
        self._c(rc)
        return Us, [infos[i] for i in range(len(Us))]
