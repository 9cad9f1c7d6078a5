"""The compact per-voxel resource (NEC2C-DIELECTRIC 4, one VOXELS record with
a binary side file; nec2c_amd/csrc/voxels.cpp) on the CPU: its cells are
the cells the reference's own mesher (oracle/_ref/libmeshref.so, built from
dielectric_mesh.c) emits for the same model written as one BOX per voxel,
its materials follow the reference's MATERIAL / SAMPLE syntax, and malformed
resources are refused with the record, value and line."""
import os

import numpy as np
import pytest

from nec2c_amd import vie
from oracle import mesh_check as M


def write_resource(tmp_path, body, name="model.diel"):
    p = tmp_path / name
    p.write_text(body)
    return str(p)


def lattice_ids(shape, seed=1, holes=True):
    rng = np.random.default_rng(seed)
    ids = rng.integers(1, 4, size=shape).astype("<u4")
    if holes:
        ids[rng.random(shape) < 0.15] = 0
    return ids


MATERIALS = """MATERIAL 1 ER 4.0
MATERIAL 2 ER 6.0 SIGMA 0.02
MATERIAL 3 TABLE
SAMPLE 3 FREQ_MHZ 100 ER 5.0 SIGMA 0.02
SAMPLE 3 FREQ_MHZ 300 ER 3.0 EI 0.1
"""


def test_cells_equal_the_reference_mesher_on_one_box_per_voxel(tmp_path):
    if not M.ref_available():
        pytest.skip("oracle/_ref/libmeshref.so not built")
    shape, h, org = (7, 5, 4), 0.0078125, (0.1, -0.05, 0.02)
    ids = lattice_ids(shape)
    ids.tofile(tmp_path / "ids.u32")
    path = write_resource(tmp_path, "NEC2C-DIELECTRIC 4\n" + MATERIALS +
                          f"VOXELS {shape[0]} {shape[1]} {shape[2]} PITCH {h!r} ORIGIN {org[0]!r} {org[1]!r} "
                          f"{org[2]!r} MATERIALS ids.u32\nEND\n")
    v = vie.Voxels(path)
    assert v.shape == shape and v.pitch == h and v.cells == int((ids > 0).sum())
    centers, sizes, mids = v.cell_arrays()
    # the same model as one BOX record per dielectric voxel
    idx = np.argwhere(ids > 0)
    mn = np.asarray(org) + h * idx
    mx = np.asarray(org) + h * (idx + 1)
    ok, rec, val, rc, rs, rb = M.reference_cells(mn, mx, frequency_mhz=100.0, max_kh=0.5)
    assert ok, (rec, val)
    assert len(rc) == v.cells
    # the same cells (the mesher orders boxes canonically; compare as sets)
    ours = np.round(centers / h, 6)
    theirs = np.round(rc / h, 6)
    assert sorted(map(tuple, ours)) == sorted(map(tuple, theirs))
    np.testing.assert_allclose(np.sort(rc, axis=0), np.sort(centers, axis=0), rtol=0, atol=1e-15)
    assert np.all(sizes == h) and np.allclose(rs, h, rtol=1e-14)
    assert np.array_equal(mids, ids[ids > 0])
    v.close()


def test_permittivity_side_file_and_vacuum_voxels(tmp_path):
    shape = (3, 4, 2)
    eps = np.full(shape, 1.0 + 0j)
    eps[0] = 4 - 0.2j
    eps[2, 1:3] = 7.5 - 1e-3j
    raw = np.stack([eps.real, eps.imag], -1).astype("<f8")
    raw.tofile(tmp_path / "eps.c128")
    path = write_resource(tmp_path, "# a comment\nnec2c-dielectric 4\nvoxels 3 4 2 pitch 0.01 origin 0 0 0.005 "
                                    "permittivity \"eps.c128\"  # trailing comment\nend\n\n# after END\n")
    v = vie.Voxels(path)
    assert v.cells == int((eps != 1).sum()) == 12
    centers, _, mids = v.cell_arrays()
    assert np.array_equal(mids, np.arange(1, 13))
    np.testing.assert_allclose(centers[0], [0.005, 0.005, 0.01])


@pytest.mark.parametrize("body,record,line,message", [
    ("NEC2C-DIELECTRIC 3\nEND\n", "NEC2C-DIELECTRIC", 1, "version header"),
    ("NEC2C-DIELECTRIC 4\nMATERIAL 1 ER 0.5\n", "MATERIAL", 2, "ER must be"),
    ("NEC2C-DIELECTRIC 4\nMATERIAL 1 ER 2\nMATERIAL 1 ER 3\n", "MATERIAL", 3, "duplicate"),
    ("NEC2C-DIELECTRIC 4\nMATERIAL 1 ER 2 EI -1\n", "MATERIAL", 2, "EI or SIGMA"),
    ("NEC2C-DIELECTRIC 4\nBOX 1 0 0 0 1 1 1\n", "BOX", 2, "unknown record"),
    ("NEC2C-DIELECTRIC 4\nVOXELS 2 2 0 PITCH 1 ORIGIN 0 0 0 MATERIALS f\n", "VOXELS", 2, "voxel counts"),
    ("NEC2C-DIELECTRIC 4\nVOXELS 2 2 2 PITCH -1 ORIGIN 0 0 0 MATERIALS f\n", "VOXELS", 2, "PITCH"),
    ("NEC2C-DIELECTRIC 4\nVOXELS 2 2 2 PITCH 1 ORIGIN 0 0 0 MATERIALS missing.u32\nEND\n", "VOXELS", 2,
     "voxel file"),
    ("NEC2C-DIELECTRIC 4\nMATERIAL 1 ER 2\nVOXELS 2 1 1 PITCH 1 ORIGIN 0 0 0 MATERIALS ids.u32\nEND\n", "VOXELS",
     3, "undeclared material"),
    ("NEC2C-DIELECTRIC 4\nMATERIAL 1 TABLE\nSAMPLE 1 FREQ_MHZ 10 ER 2 EI 0\nVOXELS 2 1 1 PITCH 1 ORIGIN 0 0 0 "
     "MATERIALS ids.u32\nEND\n", "SAMPLE", 2, "at least two samples"),
    ("NEC2C-DIELECTRIC 4\nMATERIAL 1 ER 2\nVOXELS 2 1 1 PITCH 1 ORIGIN 0 0 0 MATERIALS ids.u32\n", "END", 3,
     "missing END"),
    ("NEC2C-DIELECTRIC 4\nMATERIAL 1 ER 2\nVOXELS 2 1 1 PITCH 1 ORIGIN 0 0 0 MATERIALS ids.u32\nEND\nMATERIAL 2 "
     "ER 3\n", "MATERIAL", 5, "follow END"),
])
def test_malformed_resources_are_refused(tmp_path, body, record, line, message):
    np.array([1, 2], dtype="<u4").tofile(tmp_path / "ids.u32")       # id 2 is undeclared where used
    with pytest.raises(ValueError) as e:
        vie.Voxels(write_resource(tmp_path, body))
    text = str(e.value)
    assert f":{line}: {record}" in text and message in text, text


def test_side_file_of_the_wrong_length_and_nonpassive_permittivity(tmp_path):
    np.zeros(5, dtype="<u4").tofile(tmp_path / "short.u32")
    body = "NEC2C-DIELECTRIC 4\nMATERIAL 1 ER 2\nVOXELS 2 2 2 PITCH 1 ORIGIN 0 0 0 MATERIALS short.u32\nEND\n"
    with pytest.raises(ValueError, match="nx\\*ny\\*nz"):
        vie.Voxels(write_resource(tmp_path, body))
    np.array([[2.0, 0.1]], dtype="<f8").tofile(tmp_path / "gain.c128")        # Im > 0: active medium
    body = "NEC2C-DIELECTRIC 4\nVOXELS 1 1 1 PITCH 1 ORIGIN 0 0 0 PERMITTIVITY gain.c128\nEND\n"
    with pytest.raises(ValueError, match="passive"):
        vie.Voxels(write_resource(tmp_path, body))
