"""Field output of the C host (vfio.c): PETSc binary Vec records as VecView
writes them (FieldsBinaryWrite, VFCommon.c:1352-1373) and VTK XML structured
grids (FieldsVTKWrite, VFCommon.c:1415-1460).  The writers run on the host
only; the GPU tests of the context writers are in test_gpu_serve.py.

PETSc's binary format is pinned by its definition rather than by a reference
file (the reference ships none): each record is the class id 1211214
(VEC_FILE_CLASSID) and the length as big-endian int32, then the values as
big-endian IEEE doubles.  The bytes are checked field by field against
``struct`` packing, the reader against the writer, and the .vts files by
parsing them as XML."""
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest


def test_petsc_record_bytes(tmp_path):
    from vpfhf_amd import host
    f = tmp_path / "a.bin"
    host.petsc_vec_write(f, [np.array([1.0, -2.5, 1e300]), np.array([])])
    raw = f.read_bytes()
    want = struct.pack(">ii", 1211214, 3) + struct.pack(">ddd", 1.0, -2.5, 1e300) + struct.pack(">ii", 1211214, 0)
    assert raw == want


def test_petsc_round_trip_and_append(tmp_path):
    from vpfhf_amd import host
    rng = np.random.default_rng(5)
    a, b, c = rng.standard_normal(1000), rng.standard_normal(7) * 1e-300, np.array([np.inf, -0.0, np.pi])
    f = tmp_path / "r.bin"
    host.petsc_vec_write(f, [a, b])
    host.petsc_vec_write(f, [c], append=True)
    r = host.petsc_vec_read(f)
    assert len(r) == 3
    for x, y in zip(r, (a, b, c)):
        assert x.tobytes() == y.tobytes()  # bit-exact, signed zero and inf included
    host.petsc_vec_write(f, [c])  # overwrite
    assert len(host.petsc_vec_read(f)) == 1


def test_petsc_null_vector_is_zero(tmp_path):
    """vec[v] = NULL writes len[v] zeros (the flow fields of FieldsBinaryWrite)."""
    import ctypes as C
    from vpfhf_amd import host
    L = host.load()
    f = tmp_path / "z.bin"
    x = np.arange(4.0)
    ptrs = (C.c_void_p * 2)(None, x.ctypes.data)
    lens = (C.c_int64 * 2)(5, 4)
    assert L.vfh_petsc_vec_write(str(f).encode(), 2, ptrs, lens, 0) == 0
    r = host.petsc_vec_read(f)
    assert np.array_equal(r[0], np.zeros(5)) and np.array_equal(r[1], x)
    lens[0] = -1
    assert L.vfh_petsc_vec_write(str(f).encode(), 2, ptrs, lens, 0) == -1  # VFG_E_ARG
    assert L.vfh_petsc_vec_write(str(tmp_path / "no" / "dir.bin").encode(), 2, ptrs, lens, 0) == -1


def _vts(file):
    root = ET.parse(file).getroot()
    assert root.get("type") == "StructuredGrid"
    piece = root.find("StructuredGrid/Piece")
    arrays = {}
    for sec in ("PointData", "CellData"):
        for d in piece.find(sec).findall("DataArray"):
            arrays[(sec, d.get("Name"))] = np.array(d.text.split(), dtype=np.float64)
    pts = np.array(piece.find("Points/DataArray").text.split(), dtype=np.float64).reshape(-1, 3)
    return root, piece, pts, arrays


def test_vts_structure_and_values(tmp_path):
    from vpfhf_amd import host
    n, l = (4, 3, 5), (2.0, 0.3, 7.0)
    nn, ne = 4 * 3 * 5, 3 * 2 * 4
    rng = np.random.default_rng(1)
    p1, p2, c1 = rng.standard_normal(nn), rng.standard_normal(nn) * 1e-200, rng.standard_normal(ne)
    f = tmp_path / "g.vts"
    host.vts_write(f, n, l, {"Fracture": p1, "Displacement_0": p2}, {"Elastic Energy": c1})
    root, piece, pts, arr = _vts(f)
    assert root.find("StructuredGrid").get("WholeExtent") == "0 3 0 2 0 4"
    assert piece.get("Extent") == "0 3 0 2 0 4"
    # %.17g round-trips every double exactly
    assert arr[("PointData", "Fracture")].tobytes() == p1.tobytes()
    assert arr[("PointData", "Displacement_0")].tobytes() == p2.tobytes()
    assert arr[("CellData", "Elastic Energy")].tobytes() == c1.tobytes()
    # points in natural order (x fastest), X_i = i l / (n - 1) (VFCommon.c:393-397)
    k, j, i = np.meshgrid(np.arange(5), np.arange(3), np.arange(4), indexing="ij")
    want = np.stack([i.ravel() * 2.0 / 3, j.ravel() * 0.3 / 2, k.ravel() * 7.0 / 4], 1)
    assert np.array_equal(pts, want)
    assert pts[-1].tolist() == [2.0, 0.3, 7.0]


def test_vts_rejects_bad_sizes(tmp_path):
    from vpfhf_amd import host
    with pytest.raises(ValueError):
        host.vts_write(tmp_path / "b.vts", (3, 3, 3), (1, 1, 1), {"a": np.zeros(26)})
    with pytest.raises(ValueError):
        host.vts_write(tmp_path / "b.vts", (3, 3, 3), (1, 1, 1), None, {"c": np.zeros(27)})
    from vpfhf_amd import _lib
    with pytest.raises(_lib.VfgError):
        host.vts_write(tmp_path / "b.vts", (1, 3, 3), (1, 1, 1))
