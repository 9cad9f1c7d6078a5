"""Every launch variant of every volume pass computes the same butterflies and
products: the pass tuner picks among them by time only, so results must not
depend on its choice beyond rounding (the compiler may contract a multiply
and an add into one fused operation differently in differently unrolled
instantiations).  Each variant is forced in turn (vie_fft_set_variants) and
the product compared with variant 0: within 1e-14 relative in FP64 and 1e-6
in FP32, on geometries where the x pass also offers its kz-mirror-paired
variants, its XCD-grouped tile order and (x extent 256) its register-resident
FP32 kernel, the register-resident y passes (y extent 256), with holes and
over a PEC plane."""
import numpy as np
import pytest

from nec2c_amd import vie, workloads as W

pytestmark = pytest.mark.gpu


def system(shape, pec, holes):
    pitch = 0.5 / 64
    centers = W.cube_cells(shape, pitch, origin=(0, 0, 0.5 * pitch if pec else 0))
    eps = W.random_permittivity(shape, 5).reshape(-1)
    if holes:
        keep = np.arange(len(centers)) % 9 != 4
        centers, eps = centers[keep], eps[keep]
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], 0, W.k0_of(250e6), pec=pec, pitch=pitch)
    s.set_permittivities(eps)
    return s


@pytest.mark.parametrize("shape,pec,holes", [((20, 12, 16), False, True), ((9, 17, 12), True, False),
                                             ((64, 64, 64), False, False),
                                             # x extent 256: the register-resident FP32 x pass
                                             ((100, 6, 5), False, True), ((90, 7, 8), True, False),
                                             # y extent 256: the register-resident y passes
                                             ((5, 90, 6), False, True)])
def test_all_pass_variants_give_identical_products(shape, pec, holes):
    s = system(shape, pec, holes)
    x = np.random.default_rng(2).standard_normal(s.order) + 1j * np.random.default_rng(3).standard_normal(s.order)
    base = {}
    assert s.set_variants([0] * 5, [0] * 5)
    for prec in (vie.FP64, vie.MIXED):
        base[prec] = s.apply_ex(x, prec, jacobi=True)
    tried = same = 0
    tol = {vie.FP64: 1e-14, vie.MIXED: 1e-6}
    for pas in range(5):
        for v in range(1, 32):
            var = [0] * 5
            var[pas] = v
            # x-pass variants from 16 (the register-resident kernel) are FP32 only
            if not s.set_variants(var if v < 16 else [0] * 5, var):
                continue
            tried += 1
            for prec in ((vie.FP64, vie.MIXED) if v < 16 else (vie.MIXED,)):
                y = s.apply_ex(x, prec, jacobi=True)
                err = float(np.max(np.abs(y - base[prec])) / np.max(np.abs(base[prec])))
                assert err <= tol[prec], (shape, pas, v, prec, err)
                same += bool(np.array_equal(y, base[prec]))
    print(f"{shape} PEC {pec}: {tried} variants agree, {same} products bit-identical", flush=True)
    assert tried >= 12
    s.close()


@pytest.mark.parametrize("shape,pec,holes", [((20, 12, 16), False, True), ((9, 17, 12), True, False),
                                             ((64, 64, 64), False, False), ((30, 40, 10), False, False),
                                             ((7, 50, 20), True, True)])
def test_fused_z_y_passes_give_the_same_products(shape, pec, holes):
    """The fused z+y forward (pass 0 carries pass 1) and y+z inverse (pass 4
    carries pass 3) for padded y and z extents up to 128: the same product as
    the separate passes, alone and together, FP64 and FP32; a fused pass
    without its skipped partner (or the other way round) is refused."""
    s = system(shape, pec, holes)
    x = np.random.default_rng(2).standard_normal(s.order) + 1j * np.random.default_rng(3).standard_normal(s.order)
    assert s.set_variants([0] * 5, [0] * 5)
    base = {prec: s.apply_ex(x, prec, jacobi=True) for prec in (vie.FP64, vie.MIXED)}
    P = [1 << max(0, int(np.ceil(np.log2(2 * e - 1)))) for e in shape]
    if pec:
        P[2] = 1 << int(np.ceil(np.log2(2 * (2 * shape[2]) - 1)))
    code = 96 + int(np.log2(P[1]))
    tol = {vie.FP64: 1e-14, vie.MIXED: 1e-6}
    tried = 0
    for var in ([code, 120, 0, 0, 0], [0, 0, 0, 120, code], [code, 120, 0, 120, code]):
        if not s.set_variants(var, var):
            continue
        tried += 1
        for prec in (vie.FP64, vie.MIXED):
            y = s.apply_ex(x, prec, jacobi=True)
            err = float(np.max(np.abs(y - base[prec])) / np.max(np.abs(base[prec])))
            print(f"{shape} PEC {pec} variants {var} precision {prec}: {err:.2e}", flush=True)
            assert err <= tol[prec], (shape, var, prec, err)
    assert tried == 3
    assert not s.set_variants([code, 0, 0, 0, 0], [0] * 5)
    assert not s.set_variants([0, 120, 0, 0, 0], [0] * 5)
    assert not s.set_variants([0, 0, 0, 0, code], [0] * 5)
    s.close()
