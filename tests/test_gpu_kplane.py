"""The QPI sums on the GPU against the oracle's literal restatements of the
reference loops: the k-plane autocorrelation (wtsurf_kplane_autocorrelation),
also at a 257 x 257 plane against the FFT cyclic correlation of the folded
plane, and the k-path jdos of surfstat_jdos (wtsurf_kpath_jdos)."""
import numpy as np
import pytest

from wannier_tools_amd.surface import WtsurfError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nkx,nky,m1,m2,nfield", [
    (9, 5, 8, 4, 1),      # jdos moduli (nkx - 1, nky - 1)
    (9, 5, 7, 4, 4),      # jsdos with an even Nk1 (m1 = nkx - 2)
    (1, 3, 1, 2, 1),      # one-column plane
    (33, 65, 32, 64, 4),
    (301, 7, 300, 6, 2),  # a row longer than one workgroup
])
def test_autocorrelation_vs_oracle(gpu, oracle_lib, nkx, nky, m1, m2, nfield):
    print("autocorrelation", nkx, nky, m1, m2, nfield, flush=True)
    rng = np.random.default_rng(nkx + 7 * nky + m1)
    a = rng.uniform(0.0, 2.0, (nfield, nkx, nky))
    if nfield > 1:
        a[1:] -= 1.0  # spin fields take both signs
    got = gpu.kplane_autocorrelation(a, m1, m2)
    want = oracle_lib.kplane_autocorrelation(a, m1, m2)
    assert got.shape == (nkx, nky)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_autocorrelation_257_plane_vs_fft(gpu):
    nkx = nky = 257
    m1 = m2 = 256
    rng = np.random.default_rng(5)
    a = rng.uniform(0.0, 1.0, (4, nkx, nky))
    e = np.zeros((4, m1, m2))
    for i1 in range(nkx):
        e[:, i1 % m1, :] += a[:, i1, :m2]
        e[:, i1 % m1, 0] += a[:, i1, m2]
    cyc = np.fft.ifft2(np.conj(np.fft.fft2(e)) * np.fft.fft2(a[:, :m1, :m2])).real.sum(axis=0)
    want = cyc[np.arange(nkx)[:, None] % m1, np.arange(nky)[None, :] % m2]
    got = gpu.kplane_autocorrelation(a, m1, m2)
    assert np.allclose(got, want, rtol=1e-11, atol=0)


def test_autocorrelation_rejects_bad_moduli(gpu):
    a = np.ones((5, 5))
    for m1, m2 in ((0, 4), (6, 4), (4, 6)):
        with pytest.raises(WtsurfError):
            gpu.kplane_autocorrelation(a, m1, m2)


@pytest.mark.parametrize("nk,nomega,nfield", [(1, 3, 1), (2, 2, 2), (22, 21, 4), (513, 7, 4),
                                              (1024, 3, 1)])
def test_kpath_jdos_vs_oracle(gpu, oracle_lib, nk, nomega, nfield):
    print("kpath jdos", nk, nomega, nfield, flush=True)
    rng = np.random.default_rng(nk + nomega)
    a = rng.uniform(0.0, 3.0, (nfield, nk, nomega))
    got = gpu.kpath_jdos(a)
    want = oracle_lib.kpath_jdos(a)
    assert got.shape == a.shape
    # the same terms added in the same order: equal up to contraction into FMA
    assert np.allclose(got, want, rtol=1e-13, atol=0)
