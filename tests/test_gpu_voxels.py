"""Systems prepared from the compact per-voxel resource on the device: the
same operator as the cell list of the reference's own prepare
(oracle/_ref) for that model, with materials by id (conductive and
tabulated) and with a permittivity per voxel, free space and PEC image."""
import numpy as np
import pytest

from nec2c_amd import vie, workloads as W

pytestmark = pytest.mark.gpu

MATERIALS = """MATERIAL 1 ER 4.0
MATERIAL 2 ER 6.0 SIGMA 0.02
MATERIAL 3 TABLE
SAMPLE 3 FREQ_MHZ 100 ER 5.0 SIGMA 0.02
SAMPLE 3 FREQ_MHZ 300 ER 3.0 EI 0.1
"""
SPECS = [vie.MaterialSpec(1, 4.0), vie.MaterialSpec(2, 6.0, 0.0, 0.02),
         vie.MaterialSpec(3, 1.0, samples=((100.0, 5.0, 0.02, 1), (300.0, 3.0, 0.1, 0)))]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("pec", [False, True])
def test_voxel_materials_system_matches_the_reference_prepare(tmp_path, ref, pec):
    shape, h = (8, 7, 6), 0.02
    org = (0.0, 0.0, h if pec else -0.03)
    rng = np.random.default_rng(3)
    ids = rng.integers(1, 4, size=shape).astype("<u4")
    ids[rng.random(shape) < 0.1] = 0
    ids.tofile(tmp_path / "ids.u32")
    (tmp_path / "m.diel").write_text("NEC2C-DIELECTRIC 4\n" + MATERIALS +
                                     f"VOXELS 8 7 6 PITCH {h!r} ORIGIN {org[0]!r} {org[1]!r} {org[2]!r} "
                                     "MATERIALS ids.u32\nEND\n")
    v = vie.Voxels(str(tmp_path / "m.diel"))
    k0 = W.k0_of(200e6)
    s = vie.System.from_voxels(v, 2, k0, pec=pec)
    centers, sizes, mids = v.cell_arrays()
    cells = [vie.Cell(tuple(c), tuple(z), int(m)) for c, z, m in zip(centers, sizes, mids)]
    rs = ref.system(cells, SPECS, 2, k0, pec=pec)
    c, vv = 2, 3 * v.cells
    C = np.array([[30 - 2j, 1.0], [1.0, 25 + 1j]])
    B = 0.01 * (rng.standard_normal((c, vv)) + 1j * rng.standard_normal((c, vv)))
    D = 0.01 * (rng.standard_normal((vv, c)) + 1j * rng.standard_normal((vv, c)))
    s.set_blocks(C=C, B=B, D=D)
    rs.set_blocks(C, B, D)
    x = rng.standard_normal(c + vv) + 1j * rng.standard_normal(c + vv)
    err = rel(s.apply_ex(x, vie.FP64), rs.apply(x))
    print(f"voxel system (PEC {pec}) vs reference prepare: {err:.2e}", flush=True)
    assert err < 1e-11
    s.close()
    v.close()


def test_voxel_permittivity_system_equals_the_cell_system(tmp_path):
    n = 12
    pitch = 0.5 / 64
    eps = W.random_permittivity((n, n, n), seed=9)
    raw = np.stack([eps.real, eps.imag], -1).astype("<f8")
    raw.tofile(tmp_path / "eps.c128")
    (tmp_path / "m.diel").write_text(f"NEC2C-DIELECTRIC 4\nVOXELS {n} {n} {n} PITCH {pitch!r} ORIGIN 0 0 0 "
                                     "PERMITTIVITY eps.c128\nEND\n")
    v = vie.Voxels(str(tmp_path / "m.diel"))
    k0 = W.k0_of(300e6)
    sv = vie.System.from_voxels(v, 0, k0)
    sc = vie.System(W.cube_cells(n, pitch), [vie.MaterialSpec(1, 1.5)], 0, k0, pitch=pitch)
    sc.set_permittivities(eps.reshape(-1))
    x = np.random.default_rng(1).standard_normal(sv.order) + 0.5j
    assert np.array_equal(sv.apply_ex(x, vie.FP64), sc.apply_ex(x, vie.FP64))
    sv.close()
    sc.close()
    v.close()
