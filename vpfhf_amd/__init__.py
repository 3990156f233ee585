"""vpfhf_amd -- MI355X-native u-/V-subproblem operators of VPFHF.

Python mirror of the C-ABI in ``include/vfg.h``.  The class :class:`Problem`
holds one ``vfg_handle`` and exposes the reference's operator callbacks under
their reference names (``VF_UResidual``, ``VF_UIJacobian`` as a MatShell
product, ``VF_VResidual``, ``VF_VIJacobian``, ``VF_UEnergy3D``,
``VF_VEnergy3D``, ``VolumetricCrackOpening``, ``VF_StepU``, ``VF_StepV``).
Arrays are numpy float64 in the reference's DMDA natural order.  Every result
is computed by the HIP kernels in ``libvfg.so``; nothing falls back to the CPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (HOST, DEVICE, NONE, ZERO, ONE, FIXED, X0, X1, Y0, Y1, Z0, Z1,  # noqa: F401
                   VfgError, check, load)

__all__ = ["Problem", "DeviceBuffer", "make_bc", "load", "VfgError"]


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceBuffer):
        return a.ptr
    assert a.flags.c_contiguous
    return a.ctypes.data


def make_bc(faces=None, values=None, vertices=None, vertex_values=None):
    """VFBC for one component: {face: type}, {face: value}, {vertex: type}."""
    b = _lib.BC()
    for f, t in (faces or {}).items():
        b.face[f] = t
    for f, v in (values or {}).items():
        b.faceValue[f] = v
    for v, t in (vertices or {}).items():
        b.vertex[v] = t
    for v, val in (vertex_values or {}).items():
        b.vertexValue[v] = val
    return b


class DeviceBuffer:
    """A raw device allocation (used by the bench to keep data in HBM)."""

    def __init__(self, nbytes):
        L = load()
        p = C.c_void_p()
        check(L.vfg_device_alloc(nbytes, C.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(load().vfg_memcpy(self.ptr, arr.ctypes.data, arr.nbytes, 0))
        return self

    def download(self, dtype=np.float64, count=None):
        count = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(count, dtype)
        check(load().vfg_memcpy(out.ctypes.data, self.ptr, out.nbytes, 1))
        return out

    def free(self):
        if self.ptr:
            load().vfg_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Problem:
    """One VPFHF u/V problem on the GPU (a ``vfg_handle``).

    n: nodes per direction (DMDA ``-n``), l: domain lengths (``-l``).
    Property defaults are VFMatPropGet/VFPropGet's (VFCommon.c:460-588).
    """

    def __init__(self, n, l=(1.0, 1.0, 1.0), E=1.0, nu=0.0, alpha=1e-5, beta=1e-4, Gc=1.0,
                 eta=1e-5, epsilon=0.2, irrevtol=5e-2, atnum=2, pressurize=False, unilateral=0):
        self.L = load()
        g = _lib.Grid(n[0], n[1], n[2], l[0], l[1], l[2])
        self.props = _lib.Props(E, nu, alpha, beta, Gc, eta, epsilon, irrevtol, atnum, int(bool(pressurize)),
                                int(unilateral))
        h = C.c_void_p()
        check(self.L.vfg_create(C.byref(g), C.byref(self.props), C.byref(h)))
        self.h = h.value
        self._owned = True
        self.n = tuple(n)
        self.nn = n[0] * n[1] * n[2]
        self.ne = (n[0] - 1) * (n[1] - 1) * (n[2] - 1)

    @classmethod
    def borrow(cls, handle, n):
        """A Problem view of a handle owned elsewhere (e.g. by a C-host
        context): every operator call works, close() leaves the handle alone."""
        P = cls.__new__(cls)
        P.L = load()
        P.h = handle
        P._owned = False
        P.n = tuple(n)
        P.nn = n[0] * n[1] * n[2]
        P.ne = (n[0] - 1) * (n[1] - 1) * (n[2] - 1)
        return P

    # -- lifecycle ----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.L.vfg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return check(rc, self.h)

    def set_props(self, **kw):
        for k, v in kw.items():
            setattr(self.props, {"pressurize": "hasCrackPressure"}.get(k, k), v)
        self._c(self.L.vfg_set_props(self.h, C.byref(self.props)))

    def set_bc(self, bcU=None, bcV=None):
        """VFBCSetFromOptions results: bcU = 3 BC structs, bcV = 1."""
        u = (_lib.BC * 3)(*bcU) if bcU is not None else None
        self._c(self.L.vfg_set_bc(self.h, u, C.byref(bcV) if bcV is not None else None))

    def set_fields(self, V=None, VIrrev=None, theta=None, thetaRef=None, pressure=None, E_elem=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64)
                for a in (V, VIrrev, theta, thetaRef, pressure, E_elem)]
        self._c(self.L.vfg_set_fields(self.h, *[_ptr(a) for a in arrs], HOST))

    def get_fields(self):
        V = np.empty(self.nn)
        VI = np.empty(self.nn)
        self._c(self.L.vfg_get_fields(self.h, _ptr(V), _ptr(VI), HOST))
        return V, VI

    # -- u operator ----------------------------------------------------------
    def UIJacobian_mult(self, x, fp32=False):
        """MatMult of VF_UIJacobian's matrix (Dirichlet rows = identity)."""
        if fp32:
            x = np.ascontiguousarray(x, np.float32)
            y = np.empty_like(x)
            self._c(self.L.vfg_u_apply_f32(self.h, _ptr(x), _ptr(y), HOST))
            return y
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._c(self.L.vfg_u_apply(self.h, _ptr(x), _ptr(y), HOST))
        return y

    def u_diag(self):
        d = np.empty(3 * self.nn)
        self._c(self.L.vfg_u_diag(self.h, _ptr(d), HOST))
        return d

    def u_pc_apply(self, r):
        """z = M r, M the multigrid V-cycle of the u-solve (PCApply of a PCSHELL)."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        self._c(self.L.vfg_u_pc_apply(self.h, _ptr(r), _ptr(z), HOST))
        return z

    def u_rhs(self):
        b = np.empty(3 * self.nn)
        self._c(self.L.vfg_u_rhs(self.h, _ptr(b), HOST))
        return b

    def VF_UResidual(self, U, BCU):
        F = np.empty(3 * self.nn)
        U, BCU = np.ascontiguousarray(U, np.float64), np.ascontiguousarray(BCU, np.float64)  # alive for the call
        self._c(self.L.vfg_u_residual(self.h, _ptr(U), _ptr(BCU), _ptr(F), HOST))
        return F

    # -- V operator ----------------------------------------------------------
    def VIJacobian_setup(self, U):
        U = np.ascontiguousarray(U, np.float64)
        self._c(self.L.vfg_v_set_u(self.h, _ptr(U), HOST))

    def VIJacobian_mult(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._c(self.L.vfg_v_apply(self.h, _ptr(x), _ptr(y), HOST))
        return y

    def v_diag(self):
        d = np.empty(self.nn)
        self._c(self.L.vfg_v_diag(self.h, _ptr(d), HOST))
        return d

    def v_rhs(self):
        b = np.empty(self.nn)
        self._c(self.L.vfg_v_rhs(self.h, _ptr(b), HOST))
        return b

    def VF_VResidual(self, V):
        F = np.empty(self.nn)
        V = np.ascontiguousarray(V, np.float64)
        self._c(self.L.vfg_v_residual(self.h, _ptr(V), _ptr(F), HOST))
        return F

    # -- BC kit --------------------------------------------------------------
    def bc_mask(self, which="U"):
        w = 0 if which == "U" else 1
        m = np.empty((3 if w == 0 else 1) * self.nn, np.uint8)
        self._c(self.L.vfg_bc_mask(self.h, w, _ptr(m), HOST))
        return m

    def VecSetFromBC(self, which="U", B=None):
        w = 0 if which == "U" else 1
        B = np.zeros((3 if w == 0 else 1) * self.nn) if B is None else np.ascontiguousarray(B, np.float64)
        self._c(self.L.vfg_vec_set_from_bc(self.h, w, _ptr(B), HOST))
        return B

    def VecApplyDirichletBC(self, vec, B=None, which="U"):
        """B=None applies vec's own values as the BC vector, in place
        (VecApplyDirichletBC(V, V, bcV) of VFTimeStepPrepare)."""
        w = 0 if which == "U" else 1
        vec = np.ascontiguousarray(vec, np.float64)
        B = vec if B is None else np.ascontiguousarray(B, np.float64)  # alive for the call
        self._c(self.L.vfg_vec_apply_dirichlet(self.h, w, _ptr(vec), _ptr(B), HOST))
        return vec

    # -- diagnostics ---------------------------------------------------------
    def energies(self, U, V):
        """(elastic, insitu work, pressure work, surface) -- VF_UEnergy3D + VF_VEnergy3D."""
        out = (C.c_double * 4)()
        U, V = np.ascontiguousarray(U, np.float64), np.ascontiguousarray(V, np.float64)
        self._c(self.L.vfg_energies(self.h, _ptr(U), _ptr(V), out, HOST))
        return tuple(out)

    def VolumetricCrackOpening(self, U, V, cellfield=False, nodefield=False):
        """Crack volume (VFPermfield.c:75-150); with cellfield the per-cell
        opening, with nodefield also the nodal field the reference stores in
        fields->VolCrackOpening (VFPermfield.c:144-146)."""
        vol = C.c_double()
        cf = np.empty(self.ne) if (cellfield or nodefield) else None
        U, V = np.ascontiguousarray(U, np.float64), np.ascontiguousarray(V, np.float64)
        self._c(self.L.vfg_crack_volume(self.h, _ptr(U), _ptr(V), C.byref(vol), _ptr(cf), HOST))
        out = (vol.value,)
        if cellfield:
            out += (cf,)
        if nodefield:
            out += (self.CellToNodeInterpolation(cf),)
        return out if len(out) > 1 else vol.value

    def set_state(self, U):
        """The u-Jacobian's state under UNILATERAL_NOCOMPRESSION (VF_UIJacobian(U))."""
        U = np.ascontiguousarray(U, np.float64)
        self._c(self.L.vfg_u_set_state(self.h, _ptr(U), HOST))

    def energy_fields(self, U, V):
        """Per-cell (elastic, pressure work, surface) energy terms, shape (N_e, 3):
        the element terms of VF_UEnergy3D / VF_VEnergy3D (VFMech.c:1111-1427,
        2474-2526)."""
        out = np.empty(3 * self.ne)
        U, V = np.ascontiguousarray(U, np.float64), np.ascontiguousarray(V, np.float64)
        self._c(self.L.vfg_energy_fields(self.h, _ptr(U), _ptr(V), _ptr(out), HOST))
        return out.reshape(self.ne, 3)

    def CellToNodeInterpolation(self, cell, dof=1):
        """Volume-weighted cell -> node averaging (VFPermfield.c:1156-1327)."""
        cell = np.ascontiguousarray(cell, np.float64)
        if cell.size != dof * self.ne:
            raise ValueError(f"cell field has {cell.size} values, expected {dof * self.ne}")
        node = np.empty(dof * self.nn)
        self._c(self.L.vfg_cell_to_node(self.h, _ptr(cell), _ptr(node), dof, HOST))
        return node

    def penny_v(self, center, r, theta, phi, thickness):
        V = np.empty(self.nn)
        c = (C.c_double * 3)(*center)
        self._c(self.L.vfg_penny_build_v(self.h, _ptr(V), c, r, theta, phi, thickness, HOST))
        return V

    # -- solvers -------------------------------------------------------------
    @staticmethod
    def opts(rtol=1e-10, atol=0.0, maxit=100000, pc=_lib.PC_MG, precision=_lib.PREC_FP64, cheb_degree=4):
        return _lib.SolveOpts(rtol, atol, maxit, pc, precision, cheb_degree)

    def VF_StepU(self, U, BCU, **kw):
        """One Newton step of the u-problem; returns (U, info)."""
        U = np.array(U, np.float64, copy=True)
        info = _lib.SolveInfo()
        o = self.opts(**kw)
        BCU = np.ascontiguousarray(BCU, np.float64)
        self._c(self.L.vfg_u_solve(self.h, _ptr(U), _ptr(BCU), C.byref(o), C.byref(info), HOST))
        return U, info

    def VF_StepV(self, U, V, ub=None, **kw):
        """Bound-constrained V-solve; returns (V, info)."""
        V = np.array(V, np.float64, copy=True)
        info = _lib.SolveInfo()
        o = self.opts(**kw)
        ubp = None if ub is None else np.ascontiguousarray(ub, np.float64)
        U = np.ascontiguousarray(U, np.float64)
        self._c(self.L.vfg_v_solve(self.h, _ptr(U), _ptr(V), _ptr(ubp), C.byref(o), C.byref(info), HOST))
        return V, info
