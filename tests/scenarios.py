"""Matched inputs for the oracle and the GPU library.

A scenario is a plain dict of numpy arrays and BC specs; ``to_oracle`` and
``to_gpu`` turn it into an oracle problem / a ``vpfhf_amd.Problem`` holding the
same data.  BC specs are {component: ({face: type}, {face: value}, {vertex: type})}.
"""
import numpy as np

CLAMPED = {c: ({f: 1 for f in range(6)}, {}, {}) for c in range(3)}   # BCfiles/3DClamped.txt
TEST1_U = {0: ({0: 3, 1: 3}, {0: -1.0, 1: 1.0}, {}), 1: ({0: 1, 1: 1}, {}, {}), 2: ({0: 1, 1: 1}, {}, {})}
TEST1_V = ({0: 2, 1: 2}, {}, {})
TEST2_U = {0: ({0: 1}, {}, {}), 1: ({1: 1}, {}, {}), 2: ({4: 1, 5: 1}, {}, {})}
# a mixed set exercising every type, overlapping faces and vertices
MIXED_U = {0: ({0: 3, 3: 1, 4: 2}, {0: 0.25, 3: 9.0}, {1: 2, 7: 3}),
           1: ({2: 1, 5: 3}, {5: -0.5}, {0: 1, 6: 2}),
           2: ({1: 2}, {}, {4: 3, 3: 1})}
MIXED_V = ({0: 2, 5: 1, 2: 3}, {2: 0.7}, {5: 2, 2: 1})
NONE_V = ({}, {}, {})


def _bc(mod, spec):
    faces, values, verts = spec
    b = mod.BC()
    for f, t in faces.items():
        b.face[f] = t
    for f, v in values.items():
        b.faceValue[f] = v
    for v, t in verts.items():
        b.vertex[v] = t
        if t == 3:
            b.vertexValue[v] = 0.125 * (v + 1)
    return b


def make(n=(9, 8, 7), l=(1.0, 0.9, 0.8), seed=0, nu=0.2, atnum=1, pressurize=True, hetero=True,
         bcU=CLAMPED, bcV=NONE_V, E=1.0, alpha=1e-5, beta=1.0, Gc=1.0, eta=1e-5, eps=None,
         irrevtol=5e-2, thermal=True, unilateral=0):
    rng = np.random.default_rng(seed)
    nn = n[0] * n[1] * n[2]
    ne = (n[0] - 1) * (n[1] - 1) * (n[2] - 1)
    h = l[0] / (n[0] - 1)
    s = dict(n=n, l=l, nu=nu, atnum=atnum, pressurize=pressurize, E=E, alpha=alpha, beta=beta, Gc=Gc,
             eta=eta, eps=eps if eps is not None else 2 * h, irrevtol=irrevtol, bcU=bcU, bcV=bcV,
             unilateral=unilateral)
    s["E_elem"] = np.exp(0.5 * rng.standard_normal(ne)) if hetero else None
    V = np.clip(rng.uniform(-0.2, 1.2, nn), 0.0, 1.0)
    s["V"] = V
    s["VIrrev"] = np.clip(V + rng.uniform(-0.05, 0.2, nn), 0.0, 1.0)
    s["theta"] = rng.uniform(0, 10, nn) if thermal else np.zeros(nn)
    s["thetaRef"] = rng.uniform(0, 1, nn) if thermal else np.zeros(nn)
    s["pressure"] = rng.uniform(0, 2e-3, nn)
    s["U"] = 1e-3 * rng.standard_normal(3 * nn)
    s["x"] = rng.standard_normal(3 * nn)
    s["xv"] = rng.standard_normal(nn)
    return s


def to_oracle(O, s):
    op = O.OracleProblem(s["n"], s["l"], E=s["E"], nu=s["nu"], alpha=s["alpha"], beta=s["beta"], Gc=s["Gc"],
                         eta=s["eta"], epsilon=s["eps"], irrevtol=s["irrevtol"], atnum=s["atnum"],
                         pressurize=s["pressurize"], E_elem=s["E_elem"],
                         bcU=[_bc(O, s["bcU"].get(c, ({}, {}, {}))) for c in range(3)], bcV=_bc(O, s["bcV"]),
                         unilateral=s.get("unilateral", 0))
    F = op.fields(V=s["V"], VIrrev=s["VIrrev"], theta=s["theta"], thetaRef=s["thetaRef"], pressure=s["pressure"])
    return op, F


def to_gpu(s):
    import vpfhf_amd as V
    from vpfhf_amd import _lib
    P = V.Problem(s["n"], s["l"], E=s["E"], nu=s["nu"], alpha=s["alpha"], beta=s["beta"], Gc=s["Gc"],
                  eta=s["eta"], epsilon=s["eps"], irrevtol=s["irrevtol"], atnum=s["atnum"],
                  pressurize=s["pressurize"], unilateral=s.get("unilateral", 0))
    P.set_bc([_bc(_lib, s["bcU"].get(c, ({}, {}, {}))) for c in range(3)], _bc(_lib, s["bcV"]))
    P.set_fields(V=s["V"], VIrrev=s["VIrrev"], theta=s["theta"], thetaRef=s["thetaRef"],
                 pressure=s["pressure"], E_elem=s["E_elem"])
    return P


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
