"""The per-instance batch (wtsurf_spectral_batch): the mixed workload of four
models (Num_wann 8, 16, 36, 72 with Np 1, 2, 3, 3: Ndim 8 to 216; R in
[-2, 2]^3) over random (k, omega) with eta log-uniform on [1e-4, 1e-2] eV, one
eta per instance, against the oracle instance by instance."""
import numpy as np
import pytest

from helpers import max_rel
from wannier_tools_amd import model as M

pytestmark = pytest.mark.gpu

MODELS = [(8, 1, 40), (16, 2, 41), (36, 3, 42), (72, 3, 43)]
FLOOR_X = 5.0


def _oracle_instance(oracle_lib, nw, np_, irvec, hr, k, w, eta, top, bot):
    dl, dr, db, its = oracle_lib.spectral(nw, np_, irvec, hr, k[None, :], np.array([w]), eta,
                                          top, bot)
    return np.array([dl.item(), dr.item(), db.item()]), int(its.item())


@pytest.mark.parametrize("nw,np_,seed", MODELS)
def test_mixed_batch_vs_oracle(gpu, oracle_lib, nw, np_, seed):
    print("mixed batch", nw, np_, flush=True)
    mdl = M.synthetic_model(nw, 2, 0.5, M.decay_l2(1.5), seed=seed)
    irvec, hr = mdl.surface_cell()
    gpu.set_model_arrays(nw, np_, irvec, hr)
    n = nw * np_
    ninst = 6 if n <= 108 else 3
    rng = np.random.default_rng(seed)
    kpts = rng.uniform(-0.5, 0.5, (4, 2))
    k_index = rng.integers(0, kpts.shape[0], ninst).astype(np.int32)
    omegas = rng.uniform(-1.0, 1.0, ninst) * M.EV2HARTREE
    etas = 10.0 ** rng.uniform(-4.0, -2.0, ninst) * M.EV2HARTREE
    dl, dr, db, its, st = gpu.spectral_batch(kpts, k_index, omegas, etas)
    assert np.all(st == 0)
    top = M.default_top_orbitals(nw)
    for i in range(ninst):
        k = kpts[k_index[i]]
        ref, oit = _oracle_instance(oracle_lib, nw, np_, irvec, hr, k, omegas[i], etas[i], top, top)
        assert its[i] == oit
        # FP64 floor of this instance: one-ulp perturbations of H(R)
        floor = 0.0
        prng = np.random.default_rng(7)
        for _ in range(2):
            hp = hr * (1 + 2.2e-16 * prng.choice([-1.0, 1.0], size=hr.shape))
            pert, _ = _oracle_instance(oracle_lib, nw, np_, irvec, hp, k, omegas[i], etas[i], top, top)
            floor = max(floor, max_rel(pert, ref))
        err = max_rel(np.array([dl[i], dr[i], db[i]]), ref)
        print(f"  instance {i}: eta {etas[i] / M.EV2HARTREE:.2e} eV, iters {its[i]}, "
              f"max rel err {err:.2e}, floor {floor:.2e}", flush=True)
        assert err < max(FLOOR_X * floor, 1e-12)


def test_batch_matches_grid(gpu):
    """The batch over every (k, omega) of a grid with one eta gives the grid
    run's traces bit for bit (same kernels, different instance order)."""
    mdl = M.model_ndim16()
    irvec, hr = mdl.surface_cell()
    gpu.set_model_arrays(8, 2, irvec, hr)
    kp = np.array([[0.1, 0.0], [0.3, -0.2], [-0.45, 0.1]])
    om = np.array([-0.2, 0.0, 0.15, 0.4]) * M.EV2HARTREE
    eta = 2e-3 * M.EV2HARTREE
    gl, gr, gb, git, _ = gpu.spectral_grid(kp, om, eta)
    ki = np.repeat(np.arange(kp.shape[0]), om.size).astype(np.int32)
    wi = np.tile(om, kp.shape[0])
    bl, br, bb, bit, _ = gpu.spectral_batch(kp, ki, wi, np.full(wi.size, eta))
    # grid outputs are indexed [ik, iw]; batch instance ik * nomega + iw
    shape = (kp.shape[0], om.size)
    assert np.array_equal(bl.reshape(shape), gl)
    assert np.array_equal(br.reshape(shape), gr)
    assert np.array_equal(bb.reshape(shape), gb)
    assert np.array_equal(bit.reshape(shape), git)
    # the batch's stage times are reported like a grid run's
    ms_fourier, ms_decimate = gpu.timing()
    assert ms_fourier > 0.0 and ms_decimate > 0.0


def test_batch_rejects_bad_k_index(gpu):
    from wannier_tools_amd.surface import WtsurfError
    mdl = M.model_ndim16()
    irvec, hr = mdl.surface_cell()
    gpu.set_model_arrays(8, 2, irvec, hr)
    with pytest.raises(WtsurfError):
        gpu.spectral_batch([[0.0, 0.0]], [1], [0.0], [1e-4])
