"""GPU parity of the residual path (DCT, quant, dequant, IDCT) against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from xavs2_amd import lib as X

pytestmark = pytest.mark.gpu


def laplacian_blocks(n, nblk, seed, zero_frac=0.3, b=6.0, lim=255):
    rng = np.random.default_rng(seed)
    v = np.rint(rng.laplace(0.0, b, (nblk, n, n))).clip(-lim, lim).astype(np.int16)
    v[rng.random(nblk) < zero_frac] = 0
    return v


def extremes(n):
    yy, xx = np.mgrid[0:n, 0:n]
    chk = np.where((yy + xx) % 2 == 0, 255, -255).astype(np.int16)
    return np.stack([np.full((n, n), 255, np.int16), np.full((n, n), -255, np.int16), chk,
                     np.zeros((n, n), np.int16)])


def run_residual(n, qp, b_intra, blocks):
    nblk = blocks.shape[0]
    src = X.upload(blocks)
    lv = X.DeviceBuffer(blocks.nbytes)
    rc = X.DeviceBuffer(blocks.nbytes)
    nz = X.DeviceBuffer(4 * nblk)
    X.check(X.lib().xme_residual_batch(n, qp, b_intra, src.at(), lv.at(), rc.at(), nz.at(), nblk, None),
            "xme_residual_batch")
    return (lv.download(blocks.shape, np.int16), rc.download(blocks.shape, np.int16), nz.download(nblk, np.int32))


@pytest.mark.parametrize("n", [4, 8, 16, 32])
@pytest.mark.parametrize("qp", [22, 27, 32, 37, 42])
def test_residual_path_matches_oracle(n, qp):
    blocks = np.concatenate([laplacian_blocks(n, 997, seed=n * 100 + qp), extremes(n)])
    for b_intra in (0, 1):
        level, recon, nnz = run_residual(n, qp, b_intra, blocks)
        e_level, e_recon, e_nnz = O.residual_batch(n, qp, b_intra, blocks)
        assert np.array_equal(level, e_level)
        assert np.array_equal(recon, e_recon)
        assert np.array_equal(nnz, e_nnz)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_dct_idct_quant_entry_points(n):
    rng = np.random.default_rng(n)
    blocks = rng.integers(-32768, 32768, (301, n, n)).astype(np.int16)    # exercises the int16 wraps
    blocks[:150] = laplacian_blocks(n, 150, seed=3, zero_frac=0.1)
    src = X.upload(blocks)
    dst = X.DeviceBuffer(blocks.nbytes)
    X.check(X.lib().xme_dct_batch(n, src.at(), dst.at(), len(blocks), None), "dct")
    got = dst.download(blocks.shape, np.int16)
    exp = np.stack([O.dct(n, b) for b in blocks])
    assert np.array_equal(got, exp)
    coef = rng.integers(-4000, 4000, blocks.shape).astype(np.int16)
    src2 = X.upload(coef)
    X.check(X.lib().xme_idct_batch(n, src2.at(), dst.at(), len(coef), None), "idct")
    got = dst.download(coef.shape, np.int16)
    exp = np.stack([O.idct(n, c) for c in coef])
    assert np.array_equal(got, exp)
    qp = 37
    l = int(np.log2(n))
    scale, shift, add = O.quant_params(l, qp, 0)
    nz = X.DeviceBuffer(4 * len(coef))
    X.check(X.lib().xme_quant_batch(n, src2.at(), len(coef), qp, 0, nz.at(), None), "quant")
    q = src2.download(coef.shape, np.int16)
    eq = coef.copy()
    enz = np.array([O.lib().xo_quant(O.ptr(eq[i]), n * n, scale, shift, add) for i in range(len(coef))])
    assert np.array_equal(q, eq) and np.array_equal(nz.download(len(coef), np.int32), enz)
    X.check(X.lib().xme_dequant_batch(n, src2.at(), len(coef), qp, None), "dequant")
    d = src2.download(coef.shape, np.int16)
    ds, dsh = O.dequant_params(l, qp)
    ed = eq.copy()
    for i in range(len(coef)):
        O.lib().xo_dequant(O.ptr(ed[i]), n * n, ds, dsh)
    assert np.array_equal(d, ed)


def test_partial_workgroup_and_bad_args():
    blocks = laplacian_blocks(32, 5, seed=9)        # fewer blocks than one workgroup holds
    level, recon, nnz = run_residual(32, 27, 0, blocks)
    e_level, e_recon, e_nnz = O.residual_batch(32, 27, 0, blocks)
    assert np.array_equal(level, e_level) and np.array_equal(recon, e_recon)
    src = X.upload(blocks)
    assert X.lib().xme_residual_batch(12, 27, 0, src.at(), src.at(), src.at(), None, 5, None) == -1
    assert X.lib().xme_residual_batch(32, 80, 0, src.at(), src.at(), src.at(), None, 5, None) == -1
