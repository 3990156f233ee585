"""vfg_energies_static, the energies entry point that allocates nothing (its
reduction partials live in a static device array), against vfg_energies on
two calls in a row, the second on a smaller grid, with device inputs.  Both
entry points run the same kernels in the same order, so the results are
bitwise equal."""
import ctypes as C

import numpy as np
import pytest

from tests import scenarios as S

pytestmark = pytest.mark.gpu


def _energies_both(n, seed):
    from vpfhf_amd import DeviceBuffer, _lib
    L = _lib.load()
    s = S.make(n=n, seed=seed, atnum=1 + seed % 2)
    P = S.to_gpu(s)
    try:
        dU = DeviceBuffer(8 * s["U"].size).upload(s["U"])
        dV = DeviceBuffer(8 * s["V"].size).upload(s["V"])
        ref = (C.c_double * 4)()
        got = (C.c_double * 4)()
        _lib.check(L.vfg_energies(P.h, dU.ptr, dV.ptr, ref, _lib.DEVICE), P.h)
        _lib.check(L.vfg_energies_static(P.h, dU.ptr, dV.ptr, got), P.h)
        dU.free()
        dV.free()
        return list(ref), list(got)
    finally:
        P.close()


def test_energies_static_matches_energies():
    # first call: 400k cells (1563 reduction blocks); second: a smaller grid
    for n, seed in (((81, 77, 71), 1), ((13, 11, 9), 2)):
        ref, got = _energies_both(n, seed)
        assert ref[0] != 0. and ref[2] != 0. and ref[3] != 0.
        assert got == ref, (n, ref, got)
