"""ctypes binding of the CPU oracle (``libvforacle.so``).

TEST INFRASTRUCTURE ONLY: the oracle is the checker.  It may be imported by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` -- never by the product package ``vpfhf_amd``.

The C functions restate the reference element loops (see ``vforacle.c`` for
the file:line of each).  This module only marshals numpy arrays.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvforacle.so")

NONE, ZERO, ONE, FIXED = 0, 1, 2, 3
X0, X1, Y0, Y1, Z0, Z1 = range(6)


class BC(C.Structure):
    _fields_ = [("face", C.c_int * 6), ("faceValue", C.c_double * 6),
                ("vertex", C.c_int * 8), ("vertexValue", C.c_double * 8)]


class Elem(C.Structure):
    _fields_ = [("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("weight", C.c_double * 27),
                ("phi", C.c_double * (8 * 27)),
                ("dphi", C.c_double * (8 * 3 * 27))]


class Problem(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("E", C.c_double), ("nu", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("Gc", C.c_double),
                ("eta", C.c_double), ("epsilon", C.c_double), ("irrevtol", C.c_double),
                ("atnum", C.c_int), ("hasCrackPressure", C.c_int),
                ("E_elem", C.POINTER(C.c_double)),
                ("bcU", BC * 3), ("bcV", BC), ("unilateral", C.c_int)]


class Fields(C.Structure):
    _fields_ = [("V", C.POINTER(C.c_double)), ("VIrrev", C.POINTER(C.c_double)),
                ("theta", C.POINTER(C.c_double)), ("thetaRef", C.POINTER(C.c_double)),
                ("pressure", C.POINTER(C.c_double))]


class Ell(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("width", C.c_int),
                ("col", C.POINTER(C.c_int64)), ("val", C.POINTER(C.c_double))]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        P = C.POINTER(Problem)
        Fp = C.POINTER(Fields)
        Ep = C.POINTER(Ell)
        L.vfo_elem3d_init.argtypes = [C.POINTER(Elem), C.c_double, C.c_double, C.c_double]
        L.vfo_bilinear_u.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.POINTER(Elem)]
        L.vfo_u_jacobian.argtypes = [P, Fp]
        L.vfo_u_jacobian.restype = Ep
        L.vfo_v_jacobian.argtypes = [P, Fp, dp]
        L.vfo_v_jacobian.restype = Ep
        L.vfo_ell_free.argtypes = [Ep]
        L.vfo_ell_matvec.argtypes = [Ep, dp, dp]
        L.vfo_ell_to_csr.argtypes = [Ep, C.POINTER(C.c_int64), C.POINTER(C.c_int64), dp]
        L.vfo_ell_to_csr.restype = C.c_int64
        L.vfo_u_residual.argtypes = [P, Fp, dp, dp, dp]
        L.vfo_u_rhs.argtypes = [P, Fp, dp]
        L.vfo_v_residual.argtypes = [P, Fp, dp, dp, dp]
        L.vfo_v_rhs.argtypes = [P, Fp, dp, dp]
        L.vfo_u_energy.argtypes = [P, Fp, dp, dp, dp]
        L.vfo_v_energy.argtypes = [P, dp]
        L.vfo_v_energy.restype = C.c_double
        L.vfo_crack_volume.argtypes = [P, dp, dp, dp]
        L.vfo_crack_volume.restype = C.c_double
        L.vfo_u_jacobian_at.argtypes = [P, Fp, dp]
        L.vfo_u_jacobian_at.restype = Ep
        El = C.POINTER(Elem)
        L.vfo_bilinear_u_nocomp.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, El]
        L.vfo_grad_u_thermoporo_nocomp.argtypes = [dp, dp, dp, dp, dp] + [C.c_double] * 5 + [El]
        L.vfo_energy_density_sd_nocomp.argtypes = [dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, El]
        L.vfo_bilinear_v_coupling_nocomp.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, El]
        L.vfo_elastic_energy_nocomp_local.argtypes = [dp, dp, dp, dp] + [C.c_double] * 5 + [El]
        L.vfo_elastic_energy_nocomp_local.restype = C.c_double
        L.vfo_energy_fields.argtypes = [P, Fp, dp, dp, dp]
        L.vfo_energy_fields.restype = None
        L.vfo_cell_to_node.argtypes = [P, dp, dp, C.c_int]
        L.vfo_cell_to_node.restype = None
        L.vfo_bc_mask.argtypes = [P, C.POINTER(BC), C.c_int, C.POINTER(C.c_uint8)]
        L.vfo_vec_set_from_bc.argtypes = [P, C.POINTER(BC), C.c_int, dp]
        L.vfo_vec_apply_dirichlet.argtypes = [P, C.POINTER(BC), C.c_int, dp, dp]
        L.vfo_residual_apply_dirichlet.argtypes = [P, C.POINTER(BC), C.c_int, dp, dp, dp]
        L.vfo_bc_parse_options.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(BC)]
        L.vfo_penny_build_v.argtypes = [dp, P, dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.vfo_distance_to_penny.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double]
        L.vfo_distance_to_penny.restype = C.c_double
        L.vfo_u_solve.argtypes = [P, Fp, dp, dp, C.c_double, C.c_int, dp]
        L.vfo_u_solve.restype = C.c_int
        L.vfo_v_solve.argtypes = [P, Fp, dp, dp, dp, C.c_double, C.c_int]
        L.vfo_v_solve.restype = C.c_int
        L.vfo_pcg.argtypes = [Ep, dp, dp, C.c_double, C.c_double, C.c_int, dp]
        L.vfo_pcg.restype = C.c_int
        L.vfo_num_threads.restype = C.c_int
        L.vfo_set_num_threads.argtypes = [C.c_int]
        L.vfo_elastic_energy_density.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, C.POINTER(Elem)]
        L.vfo_bilinear_v_at.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Elem)]
        L.vfo_bilinear_v_coupling.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, C.POINTER(Elem)]
        L.vfo_elastic_energy_local.argtypes = [dp, dp, dp, dp] + [C.c_double] * 5 + [C.POINTER(Elem)]
        L.vfo_elastic_energy_local.restype = C.c_double
        L.vfo_surface_energy_local.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Elem)]
        L.vfo_surface_energy_local.restype = C.c_double
        L.vfo_crack_volume_local.argtypes = [dp, dp, C.POINTER(Elem), dp]
        L.vfo_crack_volume_local.restype = C.c_double
        L.vfo_pressure_work_local.argtypes = [dp, dp, dp, C.POINTER(Elem)]
        L.vfo_pressure_work_local.restype = C.c_double
        L.vfo_grad_u_thermoporo.argtypes = [dp, dp, dp, dp] + [C.c_double] * 5 + [C.POINTER(Elem)]
        L.vfo_grad_u_crackpressure.argtypes = [dp, dp, dp, C.POINTER(Elem)]
        L.vfo_residual_v_at.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Elem)]
        L.vfo_residual_v_crackpressure.argtypes = [dp, dp, dp, C.POINTER(Elem)]
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_bc(faces=None, values=None, vertices=None):
    b = BC()
    for f, t in (faces or {}).items():
        b.face[f] = t
    for f, v in (values or {}).items():
        b.faceValue[f] = v
    for v, t in (vertices or {}).items():
        b.vertex[v] = t
    return b


class OracleProblem:
    """Holds a ``Problem`` struct plus the arrays it points to."""

    def __init__(self, n, l=(1.0, 1.0, 1.0), E=1.0, nu=0.0, alpha=1e-5, beta=1e-4, Gc=1.0,
                 eta=1e-5, epsilon=0.2, irrevtol=5e-2, atnum=2, pressurize=False,
                 E_elem=None, bcU=None, bcV=None, unilateral=0):
        self.P = Problem()
        self.P.nx, self.P.ny, self.P.nz = n
        self.P.lx, self.P.ly, self.P.lz = l
        for k, v in dict(E=E, nu=nu, alpha=alpha, beta=beta, Gc=Gc, eta=eta,
                         epsilon=epsilon, irrevtol=irrevtol).items():
            setattr(self.P, k, v)
        self.P.atnum = atnum
        self.P.hasCrackPressure = int(bool(pressurize))
        self.P.unilateral = int(unilateral)
        self.E_elem = None if E_elem is None else np.ascontiguousarray(E_elem, np.float64)
        self.P.E_elem = _dp(self.E_elem) if self.E_elem is not None else C.POINTER(C.c_double)()
        for c in range(3):
            self.P.bcU[c] = bcU[c] if bcU is not None else BC()
        self.P.bcV = bcV if bcV is not None else BC()
        self.nn = n[0] * n[1] * n[2]
        self.ne = (n[0] - 1) * (n[1] - 1) * (n[2] - 1)
        self._keep = []

    @property
    def ref(self):
        return C.byref(self.P)

    def fields(self, V=None, VIrrev=None, theta=None, thetaRef=None, pressure=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64)
                for a in (V, VIrrev, theta, thetaRef, pressure)]
        self._keep.append(arrs)
        F = Fields(*[_dp(a) if a is not None else C.POINTER(C.c_double)() for a in arrs])
        self._keep.append(F)
        return F


def ell_to_csr(E):
    L = lib()
    n = E.contents.nrows
    nnz = L.vfo_ell_to_csr(E, None, None, None)
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(nnz, np.int64)
    va = np.zeros(nnz, np.float64)
    L.vfo_ell_to_csr(E, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                     ci.ctypes.data_as(C.POINTER(C.c_int64)), _dp(va))
    return sp.csr_matrix((va, ci, rp), shape=(n, n))


def u_jacobian(op, F, csr=True, U=None):
    """VF_UIJacobian; U: the state (NOCOMPRESSION), None for U = 0."""
    L = lib()
    E = L.vfo_u_jacobian_at(op.ref, C.byref(F), _dp(U)) if U is not None else L.vfo_u_jacobian(op.ref, C.byref(F))
    if not csr:
        return E
    A = ell_to_csr(E)
    L.vfo_ell_free(E)
    return A


def v_jacobian(op, F, U):
    L = lib()
    E = L.vfo_v_jacobian(op.ref, C.byref(F), _dp(U))
    A = ell_to_csr(E)
    L.vfo_ell_free(E)
    return A


def ell_matvec(E, x):
    y = np.zeros(E.contents.nrows)
    lib().vfo_ell_matvec(E, _dp(x), _dp(y))
    return y


def u_residual(op, F, U, BCU):
    r = np.zeros(3 * op.nn)
    lib().vfo_u_residual(op.ref, C.byref(F), _dp(U), _dp(BCU), _dp(r))
    return r


def u_rhs(op, F):
    b = np.zeros(3 * op.nn)
    lib().vfo_u_rhs(op.ref, C.byref(F), _dp(b))
    return b


def v_residual(op, F, U, V):
    r = np.zeros(op.nn)
    lib().vfo_v_residual(op.ref, C.byref(F), _dp(U), _dp(V), _dp(r))
    return r


def v_rhs(op, F, U):
    b = np.zeros(op.nn)
    lib().vfo_v_rhs(op.ref, C.byref(F), _dp(U), _dp(b))
    return b


def u_energy(op, F, U):
    e = C.c_double()
    p = C.c_double()
    lib().vfo_u_energy(op.ref, C.byref(F), _dp(U), C.byref(e), C.byref(p))
    return e.value, p.value


def v_energy(op, V):
    return lib().vfo_v_energy(op.ref, _dp(V))


def crack_volume(op, U, V, cellfield=False):
    cf = np.zeros(op.ne) if cellfield else None
    vol = lib().vfo_crack_volume(op.ref, _dp(U), _dp(V), _dp(cf))
    return (vol, cf) if cellfield else vol


def energy_fields(op, F, U, V):
    """Per-cell (elastic, pressure work, surface) energy terms, shape (N_e, 3)."""
    out = np.zeros(3 * op.ne)
    lib().vfo_energy_fields(op.ref, C.byref(F), _dp(U), _dp(V), _dp(out))
    return out.reshape(op.ne, 3)


def cell_to_node(op, cell, dof=1):
    """CellToNodeInterpolation (VFPermfield.c:1156-1327)."""
    cell = np.ascontiguousarray(cell, np.float64)
    node = np.zeros(dof * op.nn)
    lib().vfo_cell_to_node(op.ref, _dp(cell), _dp(node), dof)
    return node


def bc_mask(op, which="U"):
    dof = 3 if which == "U" else 1
    m = np.zeros(dof * op.nn, np.uint8)
    bcs = op.P.bcU if which == "U" else C.pointer(op.P.bcV)
    lib().vfo_bc_mask(op.ref, C.cast(bcs, C.POINTER(BC)), dof,
                      m.ctypes.data_as(C.POINTER(C.c_uint8)))
    return m


def vec_set_from_bc(op, which="U", out=None):
    dof = 3 if which == "U" else 1
    out = np.zeros(dof * op.nn) if out is None else out
    bcs = op.P.bcU if which == "U" else C.pointer(op.P.bcV)
    lib().vfo_vec_set_from_bc(op.ref, C.cast(bcs, C.POINTER(BC)), dof, _dp(out))
    return out


def vec_apply_dirichlet(op, vec, B, which="U"):
    dof = 3 if which == "U" else 1
    bcs = op.P.bcU if which == "U" else C.pointer(op.P.bcV)
    lib().vfo_vec_apply_dirichlet(op.ref, C.cast(bcs, C.POINTER(BC)), dof, _dp(vec), _dp(B))
    return vec


def parse_bc_options(text, prefix, dof):
    arr = (BC * dof)()
    n = lib().vfo_bc_parse_options(text.encode(), prefix.encode(), dof, arr)
    if n < 0:
        raise ValueError("bad BC option value")
    return [arr[c] for c in range(dof)]


def penny_v(op, center, r, theta, phi, thickness):
    V = np.zeros(op.nn)
    c = np.ascontiguousarray(center, np.float64)
    lib().vfo_penny_build_v(_dp(V), op.ref, _dp(c), r, theta, phi, thickness)
    return V


def u_solve(op, F, U, BCU, rtol=1e-10, maxit=100000):
    rr = C.c_double()
    it = lib().vfo_u_solve(op.ref, C.byref(F), _dp(U), _dp(BCU), rtol, maxit, C.byref(rr))
    return it, rr.value


def v_solve(op, F, U, V, ub=None, tol=1e-10, maxit=200):
    return lib().vfo_v_solve(op.ref, C.byref(F), _dp(U), _dp(V), _dp(ub), tol, maxit)


def elem(hx, hy, hz):
    e = Elem()
    lib().vfo_elem3d_init(C.byref(e), hx, hy, hz)
    return e
