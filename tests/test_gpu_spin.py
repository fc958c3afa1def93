"""Spin-resolved traces on the GPU (wtsurf_set_spin_traces / wtsurf_spin_download)
against the oracle's full G * sigma products, for surfstat's choice of
orbitals on GRR (the TopOrbitals indices) and SurfaceDOSkk's (the bottom
orbitals of the last layer), on grid and batch runs, through the LDS and the
global-memory Gauss-Jordan kernels."""
import numpy as np
import pytest

from wannier_tools_amd import model as M
from wannier_tools_amd.surface import WtsurfError

pytestmark = pytest.mark.gpu

FLOOR_X = 5.0


def _spin_floor(oracle_lib, nw, np_, irvec, hr, kp, om, eta, top, bot, mode, divisor, base):
    """Largest absolute change of the spin traces under one-ulp perturbations of H(R)."""
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(2):
        hp = hr * (1 + 2.2e-16 * rng.choice([-1.0, 1.0], size=hr.shape))
        p = oracle_lib.spectral_spin(nw, np_, irvec, hp, kp, om, eta, top, bot, spin_mode=mode,
                                     divisor=divisor)
        worst = max(worst, np.abs(p[4] - base[4]).max(), np.abs(p[5] - base[5]).max())
    return worst


CASES = [
    # nw, np, seed, soc, top (1-based), bot, nk, nomega
    (10, 3, 11, 0.05, list(range(6, 11)) + list(range(1, 3)), [1, 2, 3, 6, 7, 8], 3, 4),
    (36, 2, 17, 0.05, list(range(1, 37)), list(range(1, 37)), 2, 2),
    (7, 2, 19, 0.0, [1, 2, 3, 4, 5, 6, 7, 9, 14], [2, 5, 7], 2, 3),   # odd nw: orbital 7 unpaired
    (72, 2, 23, 0.05, list(range(1, 73)), list(range(1, 37)), 1, 2),  # order 144: matrix in global memory
]


@pytest.mark.parametrize("mode,divisor", [("surfstat", 1.0), ("arc", M.REF_PI)])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nw{c[0]}_np{c[1]}")
def test_spin_traces_vs_oracle(gpu, oracle_lib, case, mode, divisor):
    nw, np_, seed, soc, top, bot, nk, nom = case
    print("spin", mode, "nw", nw, "np", np_, flush=True)
    mdl = M.synthetic_model(nw, 1, 0.4, M.decay_l1, seed=seed, soc_sigma=soc)
    irvec, hr = mdl.surface_cell()
    rng = np.random.default_rng(seed)
    kp = rng.uniform(-0.5, 0.5, (nk, 2))
    om = rng.uniform(-0.8, 0.8, nom) * M.EV2HARTREE
    eta = 3e-3 * M.EV2HARTREE
    top = np.array(top, dtype=np.int32)
    bot = np.array(bot, dtype=np.int32)
    gpu.set_model_arrays(nw, np_, irvec, hr)
    gpu.set_orbitals(top, bot)
    gpu.set_spin_traces(mode)
    try:
        dl, dr, db, its, st = gpu.spectral_grid(kp, om, eta, divisor=divisor)
        sl, sr = gpu.spin_download()
    finally:
        gpu.set_spin_traces("off")
    imode = 1 if mode == "surfstat" else 2
    ref = oracle_lib.spectral_spin(nw, np_, irvec, hr, kp, om, eta, top, bot, spin_mode=imode,
                                   divisor=divisor)
    assert np.all(st == 0) and np.array_equal(its, ref[3])
    floor = _spin_floor(oracle_lib, nw, np_, irvec, hr, kp, om, eta, top, bot, imode, divisor, ref)
    for got, want, what in ((sl, ref[4], "left"), (sr, ref[5], "right")):
        assert got.shape == (3, nk, nom)
        err = np.abs(got - want).max()
        scale = np.abs(want).max()
        print(f"  {what}: max abs err {err:.2e}, floor {floor:.2e}, scale {scale:.2e}", flush=True)
        assert err <= max(FLOOR_X * floor, 1e-13 * scale)
    # the dos traces of the same run are untouched by the spin epilogue
    for got, want in ((dl, ref[0]), (dr, ref[1]), (db, ref[2])):
        assert np.max(np.abs(got - want) / np.abs(want)) < 1e-9


def test_spin_traces_batch(gpu, oracle_lib):
    nw, np_ = 10, 2
    mdl = M.synthetic_model(nw, 1, 0.4, M.decay_l1, seed=29, soc_sigma=0.08)
    irvec, hr = mdl.surface_cell()
    gpu.set_model_arrays(nw, np_, irvec, hr)
    rng = np.random.default_rng(3)
    kp = rng.uniform(-0.5, 0.5, (3, 2))
    ki = np.array([2, 0, 1, 0, 2], dtype=np.int32)
    om = rng.uniform(-0.5, 0.5, ki.size) * M.EV2HARTREE
    eta = 10.0 ** rng.uniform(-3.5, -2.0, ki.size) * M.EV2HARTREE
    gpu.set_spin_traces("arc")
    try:
        gpu.spectral_batch(kp, ki, om, eta, divisor=M.REF_PI)
        sl, sr = gpu.spin_download()
    finally:
        gpu.set_spin_traces("off")
    assert sl.shape == (3, ki.size)
    top = M.default_top_orbitals(nw)
    for i in range(ki.size):
        ref = oracle_lib.spectral_spin(nw, np_, irvec, hr, kp[ki[i]][None], [om[i]], eta[i], top, top,
                                       spin_mode=2, divisor=M.REF_PI)
        scale = max(np.abs(ref[4]).max(), np.abs(ref[5]).max())
        assert np.abs(sl[:, i] - ref[4][:, 0, 0]).max() < 1e-10 * scale
        assert np.abs(sr[:, i] - ref[5][:, 0, 0]).max() < 1e-10 * scale


def test_spin_download_needs_a_spin_run(gpu):
    mdl = M.model_ndim16()
    irvec, hr = mdl.surface_cell()
    gpu.set_model_arrays(8, 2, irvec, hr)
    gpu.set_spin_traces("off")
    gpu.spectral_grid([[0.1, 0.2]], [0.0], 1e-4)
    with pytest.raises(WtsurfError):
        gpu.spin_download()
    with pytest.raises(ValueError):
        gpu.set_spin_traces("sideways")
