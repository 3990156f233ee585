"""The C-ABI libraries load and export every function their headers declare;
the C host parses the reference's option syntax.  No GPU needed: nothing here
launches a kernel or creates a handle."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vf[gh]_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header,lib", [("vfg.h", "libvfg.so"), ("vfhost.h", "libvfhost.so")])
def test_library_exports_header(header, lib):
    from vpfhf_amd import _lib, host
    _lib.load()
    L = host.load() if lib == "libvfhost.so" else _lib.load()
    names = declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    if header == "vfg.h":
        assert set(names) <= set(_lib.PROTOS), set(names) - set(_lib.PROTOS)


def test_missing_library_fails_loudly(tmp_path):
    from vpfhf_amd import _lib
    with pytest.raises(ImportError):
        _lib._lib, saved = None, _lib._lib
        try:
            _lib.load(str(tmp_path / "libvfg.so"))
        finally:
            _lib._lib = saved


def test_host_options_parser():
    from vpfhf_amd.host import options
    gold = os.path.join(ROOT, "tests", "golden", "bcfiles")
    o = options(open(os.path.join(gold, "3DClamped.txt")).read())
    for c in range(3):
        assert list(o.bcU[c].face) == [1] * 6
    # the runtest1 command line (ValidationTests/makefile:198)
    o = options("-n 101,5,5 -l 10,1,1  -nu 0 -epsilon 0.5 -eta 1e-8 -gc .1 -altmintol 1.e-3 -maxtimestep 11 "
                "-maxtimevalue 4 -atnum 1 -V_X0_BC ONE -U_X0_BC_0 FIXED -U_X0_0 -1. -U_X0_BC_1 ZERO "
                "-U_X0_BC_2 ZERO -V_X1_BC ONE -U_X1_BC_0 FIXED -U_X1_0 1. -U_X1_BC_1 ZERO -U_X1_BC_2 ZERO "
                "-verbose 1 -alpha 0 -beta 0 -pressure 0")
    assert (o.grid.nx, o.grid.ny, o.grid.nz) == (101, 5, 5)
    assert (o.grid.lx, o.grid.ly, o.grid.lz) == (10., 1., 1.)
    assert o.props.nu == 0 and o.props.epsilon == .5 and o.props.eta == 1e-8 and o.props.Gc == .1
    assert o.props.atnum == 1 and o.props.alpha == 0 and o.props.beta == 0 and o.altmintol == 1e-3
    assert o.bcU[0].face[0] == 3 and o.bcU[0].faceValue[0] == -1. and o.bcU[0].faceValue[1] == 1.
    assert o.bcU[1].face[0] == 1 and o.bcU[2].face[1] == 1 and o.bcV.face[0] == 2 and o.bcV.face[1] == 2
    o = options(open(os.path.join(gold, "3DRigidMotion.txt")).read())
    assert o.bcU[0].vertex[0] == 1 and o.bcU[2].vertex[1] == 1 and o.bcU[1].vertex[4] == 1
    o = options("-npc 2 -pc1_center .1,.2,.3 -pc1_r .25 -pc1_theta 30 -pc1_phi 45 -pc1_thickness 0.01 -pressurize")
    assert o.npc == 2 and tuple(o.pc[1].center) == (.1, .2, .3) and o.pc[1].r == .25
    assert o.pc[1].theta == 30 and o.pc[1].phi == 45 and o.pc[1].thickness == .01 and o.props.hasCrackPressure == 1
