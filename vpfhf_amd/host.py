"""ctypes binding of the C host layer (``libvfhost.so``, include/vfhost.h).

``Host`` wraps one ``vfh_ctx``: the option subset of the reference's
context (same option names: ``-n``, ``-l``, ``-nu``, ``-atnum``,
``-U_X0_BC_0 ZERO``, ``-pc0_center`` ...), device-resident U, BCU, V, VIrrev,
VFTimeStepPrepare and the alternate-minimisation loop.
"""
import ctypes as C
import os

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libvfhost.so")
MAX_PENNY = 64


class Penny(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("r", C.c_double), ("theta", C.c_double), ("phi", C.c_double),
                ("thickness", C.c_double)]


class Options(C.Structure):
    _fields_ = [("grid", _lib.Grid), ("props", _lib.Props), ("bcU", _lib.BC * 3), ("bcV", _lib.BC),
                ("altmintol", C.c_double), ("altminmaxit", C.c_int), ("u_rtol", C.c_double),
                ("v_rtol", C.c_double), ("npc", C.c_int), ("pc", Penny * MAX_PENNY),
                ("precision", C.c_int), ("pc_type", C.c_int)]


class AltminInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("u_its", C.c_int), ("v_outer", C.c_int), ("v_its", C.c_int),
                ("errV", C.c_double), ("status", C.c_int)]


class Request(C.Structure):
    _fields_ = [("options", C.c_char_p), ("E_elem", C.c_void_p), ("theta", C.c_void_p), ("thetaRef", C.c_void_p),
                ("pressure", C.c_void_p), ("arrival", C.c_double), ("am_maxit", C.c_int),
                ("U_out", C.c_void_p), ("V_out", C.c_void_p), ("opening_out", C.c_void_p),
                ("energy_out", C.c_void_p)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("am_iterations", C.c_int), ("u_its", C.c_int), ("v_outer", C.c_int),
                ("v_its", C.c_int), ("errV", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("latency", C.c_double), ("energies", C.c_double * 5), ("device_bytes", C.c_int64)]


class ServeOpts(C.Structure):
    _fields_ = [("engine", _lib.EngineOpts), ("hbm_budget", C.c_double), ("realtime", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("wall", C.c_double), ("solves", C.c_double), ("solves_per_s", C.c_double), ("p50", C.c_double),
                ("p99", C.c_double), ("mean_latency", C.c_double), ("max_latency", C.c_double),
                ("completed", C.c_int), ("rejected", C.c_int), ("max_resident", C.c_int),
                ("peak_bytes", C.c_double), ("ticks", C.c_int64), ("replays", C.c_int64)]


_hl = None


def load():
    global _hl
    if _hl is None:
        _lib.load()  # libvfg.so first (libvfhost links it)
        if not os.path.exists(LIBPATH):
            raise ImportError("libvfhost.so not built: run `python -m vpfhf_amd.build`")
        L = C.CDLL(LIBPATH)
        vp, dp, O = C.c_void_p, C.POINTER(C.c_double), C.POINTER(Options)
        protos = {
            "vfh_options_default": (None, [O]),
            "vfh_options_parse": (C.c_int, [O, C.c_char_p]),
            "vfh_last_error": (C.c_char_p, [vp]),
            "vfh_create": (C.c_int, [O, C.POINTER(vp)]),
            "vfh_destroy": (C.c_int, [vp]),
            "vfh_handle": (vp, [vp]),
            "vfh_device_fields": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
            "vfh_set_inputs": (C.c_int, [vp, vp, vp, vp, vp]),
            "vfh_fields_initialize": (C.c_int, [vp]),
            "vfh_set_bcu": (C.c_int, [vp, C.c_double]),
            "vfh_timestep_prepare": (C.c_int, [vp]),
            "vfh_altmin": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(AltminInfo)]),
            "vfh_energies": (C.c_int, [vp, dp]),
            "vfh_get_state": (C.c_int, [vp, vp, vp, vp]),
            "vfh_crack_opening_field": (C.c_int, [vp, vp]),
            "vfh_energy_fields": (C.c_int, [vp, vp]),
            "vfh_set_state": (C.c_int, [vp, vp, vp, vp]),
            "vfh_petsc_vec_write": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp), C.POINTER(C.c_int64), C.c_int]),
            "vfh_fields_binary_write": (C.c_int, [vp, C.c_char_p, C.c_int, vp, vp]),
            "vfh_coordinates_write": (C.c_int, [vp, C.c_char_p]),
            "vfh_vts_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_int, C.POINTER(C.c_char_p), C.POINTER(vp), C.c_int,
                                        C.POINTER(C.c_char_p), C.POINTER(vp)]),
            "vfh_fields_vts_write": (C.c_int, [vp, C.c_char_p, C.c_int, vp, vp]),
            "vfh_serve_opts_default": (None, [C.POINTER(ServeOpts)]),
            "vfh_request_bytes": (C.c_int64, [C.c_int, C.c_int, C.c_int]),
            "vfh_serve": (C.c_int, [C.POINTER(ServeOpts), C.POINTER(Request), C.c_int, C.POINTER(RequestResult),
                                    C.POINTER(ServeStats)]),
        }
        for name, (res, args) in protos.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _hl = L
    return _hl


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, np.float64)


def petsc_vec_write(file, vecs, append=False):
    """Writes the arrays of ``vecs`` as PETSc binary Vec records (VecView with
    a binary viewer: int32 1211214, int32 length, big-endian doubles)."""
    keep = [_f64(v).ravel() for v in vecs]
    n = len(keep)
    ptrs = (C.c_void_p * max(n, 1))(*[v.ctypes.data for v in keep])
    lens = (C.c_int64 * max(n, 1))(*[v.size for v in keep])
    rc = load().vfh_petsc_vec_write(os.fsencode(file), n, ptrs, lens, int(bool(append)))
    if rc:
        raise _lib.VfgError(rc, "vfh_petsc_vec_write(%s)" % file)


def petsc_vec_read(file):
    """The Vec records of a PETSc binary file, in order (VecLoad,
    Utils/vtkexport.c:94-110)."""
    raw = open(file, "rb").read()
    out, q = [], 0
    while q < len(raw):
        cid, n = np.frombuffer(raw, ">i4", 2, q)
        if cid != 1211214:
            raise ValueError("%s: record at byte %d is not a Vec (class id %d)" % (file, q, cid))
        out.append(np.frombuffer(raw, ">f8", int(n), q + 8).astype(np.float64))
        q += 8 + 8 * int(n)
    return out


def vts_write(file, grid_n, grid_l, point_data=None, cell_data=None):
    """A VTK XML structured grid (.vts, ASCII) over the reference's node
    coordinates with the named nodal and cell arrays of the two dicts."""
    pd, cd = list((point_data or {}).items()), list((cell_data or {}).items())
    keep = [_f64(a).ravel() for _, a in pd + cd]
    nx, ny, nz = map(int, grid_n)
    nn, ne = nx * ny * nz, (nx - 1) * (ny - 1) * (nz - 1)
    for i, (k, _) in enumerate(pd + cd):
        want = nn if i < len(pd) else ne
        if keep[i].size != want:
            raise ValueError("vts_write: %r has %d values, the grid needs %d" % (k, keep[i].size, want))

    def arrs(items, off):
        names = (C.c_char_p * max(len(items), 1))(*[k.encode() for k, _ in items])
        ptrs = (C.c_void_p * max(len(items), 1))(*[keep[off + i].ctypes.data for i in range(len(items))])
        return names, ptrs
    pn, pp = arrs(pd, 0)
    cn, cp = arrs(cd, len(pd))
    rc = load().vfh_vts_write(os.fsencode(file), *map(int, grid_n), *map(float, grid_l), len(pd), pn, pp, len(cd),
                              cn, cp)
    if rc:
        raise _lib.VfgError(rc, "vfh_vts_write(%s)" % file)


def options(text=""):
    o = Options()
    L = load()
    L.vfh_options_default(C.byref(o))
    if text:
        rc = L.vfh_options_parse(C.byref(o), text.encode())
        if rc:
            raise _lib.VfgError(rc, L.vfh_last_error(None).decode())
    return o


class Host:
    """One problem driven by the C host (device-resident state)."""

    def __init__(self, opts):
        self.L = load()
        self.opts = opts
        ctx = C.c_void_p()
        self._c(self.L.vfh_create(C.byref(opts), C.byref(ctx)), None)
        self.ctx = ctx.value
        g = opts.grid
        self.nn = g.nx * g.ny * g.nz
        self.ne = (g.nx - 1) * (g.ny - 1) * (g.nz - 1)
        self.h = self.L.vfh_handle(self.ctx)

    def _c(self, rc, ctx="self"):
        if rc:
            c = self.ctx if ctx == "self" else ctx
            raise _lib.VfgError(rc, self.L.vfh_last_error(c).decode())

    def close(self):
        if getattr(self, "ctx", None):
            self.L.vfh_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_fields(self):
        ps = [C.c_void_p() for _ in range(4)]
        self._c(self.L.vfh_device_fields(self.ctx, *[C.byref(p) for p in ps]))
        return [p.value for p in ps]

    def set_inputs(self, E_elem=None, theta=None, thetaRef=None, pressure=None):
        keep = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (E_elem, theta, thetaRef, pressure)]
        self._c(self.L.vfh_set_inputs(self.ctx, *[None if a is None else a.ctypes.data for a in keep]))

    def fields_initialize(self):
        self._c(self.L.vfh_fields_initialize(self.ctx))

    def set_bcu(self, scale=1.0):
        self._c(self.L.vfh_set_bcu(self.ctx, scale))

    def timestep_prepare(self):
        self._c(self.L.vfh_timestep_prepare(self.ctx))

    def altmin(self, maxit=0, stop_on_tol=False):
        info = AltminInfo()
        rc = self.L.vfh_altmin(self.ctx, maxit, int(stop_on_tol), C.byref(info))
        self._c(rc)
        return info

    def energies(self):
        out = (C.c_double * 5)()
        self._c(self.L.vfh_energies(self.ctx, out))
        return dict(elastic=out[0], insitu=out[1], pressure_work=out[2], surface=out[3], crack_volume=out[4])

    def crack_opening_field(self):
        """Nodal crack opening, fields->VolCrackOpening (VFPermfield.c:75-146)."""
        out = np.empty(self.nn)
        self._c(self.L.vfh_crack_opening_field(self.ctx, out.ctypes.data))
        return out

    def energy_fields(self):
        """Per-cell (elastic, pressure work, surface) energy terms, shape (N_e, 3)."""
        out = np.empty(3 * self.ne)
        self._c(self.L.vfh_energy_fields(self.ctx, out.ctypes.data))
        return out.reshape(-1, 3)

    def fields_binary_write(self, prefix, step, theta=None, pressure=None):
        """FieldsBinaryWrite (VFCommon.c:1352-1373): ``<prefix>.<step>.bin``."""
        keep = [_f64(theta), _f64(pressure)]
        for a in keep:
            if a is not None and a.size != self.nn:
                raise ValueError("a nodal field needs %d values, got %d" % (self.nn, a.size))
        self._c(self.L.vfh_fields_binary_write(self.ctx, os.fsencode(prefix), step,
                                               *[None if a is None else a.ctypes.data for a in keep]))
        return "%s.%05d.bin" % (prefix, step)

    def coordinates_write(self, file):
        """The Coordinates Vec of the reference's <prefix>.bin (VFCommon.c:383-424)."""
        self._c(self.L.vfh_coordinates_write(self.ctx, os.fsencode(file)))

    def fields_vts_write(self, prefix, step, theta=None, pressure=None):
        """FieldsVTKWrite (VFCommon.c:1415-1460): the nodal and cell .vts files."""
        keep = [_f64(theta), _f64(pressure)]
        for a in keep:
            if a is not None and a.size != self.nn:
                raise ValueError("a nodal field needs %d values, got %d" % (self.nn, a.size))
        self._c(self.L.vfh_fields_vts_write(self.ctx, os.fsencode(prefix), step,
                                            *[None if a is None else a.ctypes.data for a in keep]))
        return "%s_nodal.%05d.vts" % (prefix, step), "%s_cell.%05d.vts" % (prefix, step)

    def state(self):
        U = np.empty(3 * self.nn)
        V = np.empty(self.nn)
        VI = np.empty(self.nn)
        self._c(self.L.vfh_get_state(self.ctx, U.ctypes.data, V.ctypes.data, VI.ctypes.data))
        return U, V, VI

    def set_state(self, U=None, V=None, VIrrev=None):
        keep = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (U, V, VIrrev)]
        self._c(self.L.vfh_set_state(self.ctx, *[None if a is None else a.ctypes.data for a in keep]))
