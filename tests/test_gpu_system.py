"""GPU parity of the hybrid [C B; D V] system (dielectric_fft.c:41-88): FP64
and FP32 operator products, the fused Jacobi product, PEC image embedding,
and FP64 / mixed device solves against the reference's own build."""
import numpy as np
import pytest

from nec2c_amd import vie, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def coupled_case(n=6, freq=300e6, seed=2, conductors=True):
    pitch = 0.5 / 64
    eps = W.random_permittivity((n, n, n), seed).reshape(-1)
    centers = W.cube_cells(n, pitch)
    k0 = W.k0_of(freq)
    dip = W.Dipole(center=(-0.1, 0.25, 0.25))
    c = dip.segments if conductors else 0
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], c, k0, pitch=pitch)
    s.set_permittivities(eps)
    blocks = None
    if conductors:
        C = dip.impedance_matrix(k0)
        B, D = W.coupling_blocks(dip, centers, pitch, eps, k0)
        s.set_blocks(C=C, B=B, D=D)
        blocks = (C, B, D)
    return s, eps, pitch, k0, blocks, dip


def ref_coupled(ref, n, pitch, k0, eps, blocks):
    xyz = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    p = ref.plan((n, n, n), xyz, pitch, k0, eps[0])
    p.set_permittivities(eps)
    return p, (p.coupled(*blocks) if blocks else None)


def test_volume_only_fp32_and_jacobi(ref):
    n = 8
    s, eps, pitch, k0, _, _ = coupled_case(n, conductors=False)
    rp, _ = ref_coupled(ref, n, pitch, k0, eps, None)
    x = np.random.default_rng(1).standard_normal(s.order) + 1j * np.random.default_rng(2).standard_normal(s.order)
    y64 = s.apply_ex(x, vie.FP64)
    yr = rp.apply(x)
    assert rel(y64, yr) < 1e-11
    y32 = s.apply_ex(x, vie.MIXED)
    print("fp32 rel", rel(y32, yr), flush=True)
    assert rel(y32, yr) < 2e-6
    yj = s.apply_ex(x, vie.FP64, jacobi=True)
    assert rel(yj, rp.apply(rp.precondition(x))) < 1e-11
    yj32 = s.apply_ex(x, vie.MIXED, jacobi=True)
    assert rel(yj32, yj) < 2e-6


def test_coupled_apply_matches_reference(ref):
    n = 6
    s, eps, pitch, k0, blocks, dip = coupled_case(n)
    rp, rc = ref_coupled(ref, n, pitch, k0, eps, blocks)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(s.order) + 1j * rng.standard_normal(s.order)
    y = s.apply(x)
    yr = rc.apply(x)
    print("coupled fp64 rel", rel(y, yr), flush=True)
    assert rel(y, yr) < 1e-11
    y32 = s.apply_ex(x, vie.MIXED)
    print("coupled fp32 rel", rel(y32, yr), flush=True)
    assert rel(y32, yr) < 2e-6
    # Jacobi: conductor rows divide by C_ii, volume rows by eps - (eps-1) self
    pz = s.precondition(x)
    assert rel(pz[:21], x[:21] / np.diag(blocks[0])) < 1e-14
    assert rel(s.apply_ex(x, vie.FP64, jacobi=True), rc.apply(pz)) < 1e-11


@pytest.mark.parametrize("c", [3, 300, 330])
def test_many_conductors_coupling(ref, c):
    """Conductor counts below, at and above one B-reduction launch (320 rows):
    the products with random C/B/D against the reference's block loops."""
    n = 4
    pitch = 0.5 / 64
    eps = W.random_permittivity((n, n, n), 7).reshape(-1)
    centers = W.cube_cells(n, pitch)
    k0 = W.k0_of(200e6)
    rng = np.random.default_rng(c)
    nv = 3 * n ** 3
    C = rng.standard_normal((c, c)) + 1j * rng.standard_normal((c, c)) + 50 * np.eye(c)
    B = 1e-3 * (rng.standard_normal((c, nv)) + 1j * rng.standard_normal((c, nv)))
    D = 1e-3 * (rng.standard_normal((nv, c)) + 1j * rng.standard_normal((nv, c)))
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], c, k0, pitch=pitch)
    s.set_permittivities(eps)
    s.set_blocks(C=C, B=B, D=D)
    rp, rc = ref_coupled(ref, n, pitch, k0, eps, (C, B, D))
    x = rng.standard_normal(s.order) + 1j * rng.standard_normal(s.order)
    yr = rc.apply(x)
    assert rel(s.apply(x), yr) < 1e-11
    assert rel(s.apply_ex(x, vie.MIXED), yr) < 2e-6
    pz = s.precondition(x)
    assert rel(s.apply_ex(x, vie.FP64, jacobi=True), rc.apply(pz)) < 1e-11
    s.close()


@pytest.mark.parametrize("precision", [vie.FP64, vie.MIXED])
def test_coupled_solve(ref, precision):
    n = 6
    s, eps, pitch, k0, blocks, dip = coupled_case(n)
    rp, rc = ref_coupled(ref, n, pitch, k0, eps, blocks)
    b = np.zeros(s.order, complex)
    b[:21] = dip.rhs()
    okr, xr, repr_ = rc.gmres(b, rtol=1e-10)
    assert okr
    ok, x, rep = s.solve(b, precision=precision, relative_tolerance=1e-10)
    print(precision, rep.as_dict(), repr_, flush=True)
    assert ok and rep.relative_residual <= 1e-10
    fi = dip.feed_index()
    zerr = abs(1 / x[fi] - 1 / xr[fi]) / abs(1 / xr[fi])
    assert zerr < 1e-6, zerr
    assert rel(x[:21], xr[:21]) < 1e-6


def test_pec_image_matches_reference(ref):
    # two cells above a PEC ground with a half-cell gap (dielectric_fft_test.c:8-46)
    mats = [vie.MaterialSpec(1, 3, .2), vie.MaterialSpec(2, 2, .4)]
    for gap in (0.0, 0.125):
        cells = [vie.Cell((.25 * i, -.25 * i, .125 + gap + .25 * i), (.25, .25, .25), i + 1) for i in range(2)]
        s = vie.System(cells, mats, 0, .5, pec=True)
        rcells = [O.Cell(c.center, c.size, c.material_id) for c in cells]
        rmats = [O.Material(m.id, m.epsilon_r, m.epsilon_loss) for m in mats]
        rs = ref.system(rcells, rmats, 0, .5, pec=True)
        x = np.sin(1 + np.arange(6)) + 1j * np.cos(.3 * np.arange(6))
        assert rel(s.apply(x), rs.apply(x)) < 1e-11
        assert rel(s.precondition(x), rs.precondition(x)) < 1e-13
        ok, sol, rep = s.gmres(s.apply(x))
        assert ok and np.max(np.abs(sol - x)) < 1e-10


def test_pec_larger_block(ref):
    # 4x3x3 block on a gap of one cell above the PEC plane, lossy
    pitch = 0.05
    cells = [vie.Cell((pitch * i, pitch * j, pitch * (k + 1.5)), (pitch,) * 3, 1)
             for i in range(4) for j in range(3) for k in range(3)]
    mats = [vie.MaterialSpec(1, 5, .7)]
    k0 = 6.0
    s = vie.System(cells, mats, 0, k0, pec=True)
    rs = ref.system([O.Cell(c.center, c.size, 1) for c in cells], [O.Material(1, 5, .7)], 0, k0, pec=True)
    x = np.random.default_rng(5).standard_normal(s.order) + 0.5j
    assert rel(s.apply(x), rs.apply(x)) < 1e-11
    assert rel(s.apply_ex(x, vie.MIXED), rs.apply(x)) < 2e-6


def test_mixed_escalates_to_fp64_when_fp32_stalls(ref):
    """A system whose volume response dwarfs the excitation (D scaled by 1e7,
    so |x_v| / |b| ~ 1e7): FP32 products cannot lower the true residual at
    all, and a MIXED solve must hand over to FP64 cycles and still converge
    against the reference operator."""
    n, c = 6, 5
    pitch = 0.5 / 64
    eps = W.random_permittivity((n, n, n), 9).reshape(-1)
    centers = W.cube_cells(n, pitch)
    k0 = W.k0_of(200e6)
    rng = np.random.default_rng(5)
    nv = 3 * n ** 3
    C = rng.standard_normal((c, c)) + 1j * rng.standard_normal((c, c)) + 20 * np.eye(c)
    B = 1e-9 * (rng.standard_normal((c, nv)) + 1j * rng.standard_normal((c, nv)))
    D = 1e7 * (rng.standard_normal((nv, c)) + 1j * rng.standard_normal((nv, c)))
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], c, k0, pitch=pitch)
    s.set_permittivities(eps)
    s.set_blocks(C=C, B=B, D=D)
    rp, rc = ref_coupled(ref, n, pitch, k0, eps, (C, B, D))
    b = np.zeros(s.order, complex)
    b[0] = 1.0
    ok, x, rep = s.solve(b, precision=vie.MIXED, relative_tolerance=1e-6, restart=30, maximum_iterations=600)
    print("mixed", rep.as_dict(), "escalated", s.escalated, flush=True)
    assert ok and s.escalated
    r = b - rc.apply(x)
    assert np.linalg.norm(r) <= 1.05e-6 * np.linalg.norm(b)
    ok64, x64, rep64 = s.solve(b, precision=vie.FP64, relative_tolerance=1e-6, restart=30, maximum_iterations=600)
    assert ok64 and not s.escalated
    s.close()


@pytest.mark.parametrize("precision", [vie.FP64, vie.MIXED])
def test_recycled_cycles_converge_to_reference(ref, precision):
    """GMRES cycles over a recycled subspace (GCRO with the subspace of the
    first cycle's harmonic Ritz vectors kept fixed): restart 16 forces several
    cycles on the coupled 6^3 system.  The recycled solve must reach the
    reference operator's residual target and the reference GMRES solution, and
    plain cycles of the same restart must still do so too."""
    n = 6
    s, eps, pitch, k0, blocks, dip = coupled_case(n)
    rp, rc = ref_coupled(ref, n, pitch, k0, eps, blocks)
    b = np.zeros(s.order, complex)
    b[:21] = dip.rhs()
    okr, xr, _ = rc.gmres(b, rtol=1e-10)
    assert okr
    opts = dict(relative_tolerance=1e-10, restart=16, maximum_iterations=3000)
    ok0, x0, rep0 = s.solve(b, precision=precision, recycle=-1, **opts)
    assert ok0 and s.recycled == 0
    ok, x, rep = s.solve(b, precision=precision, recycle=6, **opts)
    print(precision, "plain", rep0.as_dict(), "recycled", rep.as_dict(), "k", s.recycled, flush=True)
    assert ok and rep.restarts >= 1 and s.recycled == 6
    r = b - rc.apply(x)
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)
    fi = dip.feed_index()
    for sol in (x0, x):
        zerr = abs(1 / sol[fi] - 1 / xr[fi]) / abs(1 / xr[fi])
        assert zerr < 1e-6, zerr
        assert rel(sol[:21], xr[:21]) < 1e-6
    s.close()


def test_pec_with_conductors_matches_reference(ref):
    """Conductor coupling on a PEC-imaged system: the fused B z_v of the first
    volume pass must skip the image cells and the fused D z_c of the last pass
    must land on the physical ones (dielectric_fft.c:44-64)."""
    pitch = 0.05
    cells = [vie.Cell((pitch * i, pitch * j, pitch * (k + 1.5)), (pitch,) * 3, 1)
             for i in range(4) for j in range(3) for k in range(3)]
    mats = [vie.MaterialSpec(1, 5, .7)]
    k0, c = 6.0, 5
    nv = 3 * len(cells)
    rng = np.random.default_rng(11)
    C = rng.standard_normal((c, c)) + 1j * rng.standard_normal((c, c)) + 20 * np.eye(c)
    B = 0.1 * (rng.standard_normal((c, nv)) + 1j * rng.standard_normal((c, nv)))
    D = 0.1 * (rng.standard_normal((nv, c)) + 1j * rng.standard_normal((nv, c)))
    s = vie.System(cells, mats, c, k0, pec=True)
    s.set_blocks(C=C, B=B, D=D)
    rs = ref.system([O.Cell(q.center, q.size, 1) for q in cells], [O.Material(1, 5, .7)], c, k0, pec=True)
    for r in range(c):
        for q in range(c):
            assert rs.set(r, q, C[r, q])
        for u in range(nv):
            assert rs.set(r, c + u, B[r, u])
            assert rs.set(c + u, r, D[u, r])
    x = rng.standard_normal(s.order) + 1j * rng.standard_normal(s.order)
    yr = rs.apply(x)
    assert rel(s.apply(x), yr) < 1e-11
    assert rel(s.apply_ex(x, vie.MIXED), yr) < 2e-6
    assert rel(s.apply_ex(x, vie.FP64, jacobi=True), rs.apply(rs.precondition(x))) < 1e-11
    s.close()
