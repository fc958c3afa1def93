"""Pins the oracle's spin-resolved traces (wto_spectral_spin) by what the
mathematics fixes, beside the reference's own spindos.dat files
(test_host_files.py):

* Zeeman closed form.  H = 1 (x) H0 + b sigma_c (x) 1 at R = 0 (spin-major
  basis, the reference's |up..., down...> order) splits into the spinless
  model at omega -+ b along the two spin eigenstates of sigma_c, so
    s_c(omega) = A0(omega - b) - A0(omega + b)
    the other two components = 0,   dos(omega) = A0(omega - b) + A0(omega + b),
  with A0 the spinless trace over the same orbitals.  A wrong sign, a
  swapped component or a wrong partner orbital fails this.
* Rotation covariance.  H(R) -> U H(R) U^dagger with U = u (x) 1, u in SU(2),
  rotates the spin traces over a list closed under spin partners by the SO(3)
  matrix R_cd = Tr(u^dagger sigma_c u sigma_d) / 2.
"""
import numpy as np
import pytest

from wannier_tools_amd import model as M

PAULI = [np.array([[0, 1], [1, 0]], dtype=np.complex128),
         np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
         np.array([[1, 0], [0, -1]], dtype=np.complex128)]


def zeeman_model(norb, b, comp, seed):
    """Spinless synthetic model (norb orbitals) and its spinful copy with a
    Zeeman term b sigma_comp on R = 0, both as (irvec, hr) in Hartree."""
    m0 = M.synthetic_model(norb, 1, 0.3, M.decay_l1, seed=seed)
    irvec, hr0 = m0.surface_cell()
    hr = np.array([np.kron(np.eye(2), h) for h in hr0])
    zero = np.nonzero(np.all(irvec == 0, axis=1))[0][0]
    hr[zero] += b * np.kron(PAULI[comp], np.eye(norb))
    return irvec, hr0, hr


def spinless_traces(oracle_lib, norb, np_, irvec, hr0, k, omegas, eta, divisor):
    """-sum Im G over: top (layer 1 of GLL), top indices on GRR, last layer of GRR."""
    out = []
    for w in omegas:
        h00, h01 = oracle_lib.ham_qlayer2qlayer(norb, np_, irvec, hr0, k)
        gll, grr, _, _, _ = oracle_lib.surfgreen_1985(w, eta, h00, h01)
        n = norb * np_
        out.append([-np.imag(np.trace(gll[:norb, :norb])) / divisor,
                    -np.imag(np.trace(grr[:norb, :norb])) / divisor,
                    -np.imag(np.trace(grr[n - norb:, n - norb:])) / divisor])
    return np.array(out)  # (nomega, 3)


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("spin_mode,divisor", [(1, 1.0), (2, M.REF_PI)])
def test_zeeman_closed_form(oracle_lib, comp, spin_mode, divisor):
    norb, np_ = 3, 2
    b = 0.15 * M.EV2HARTREE
    irvec, hr0, hr = zeeman_model(norb, b, comp, seed=21 + comp)
    k = np.array([[0.17, -0.31]])
    om = np.array([-0.4, -0.05, 0.2, 0.55]) * M.EV2HARTREE
    eta = 4e-3 * M.EV2HARTREE
    nw = 2 * norb
    top = np.arange(1, nw + 1, dtype=np.int32)
    dl, dr, db, _, sl, sr = oracle_lib.spectral_spin(nw, np_, irvec, hr, k, om, eta, top, top,
                                                     spin_mode=spin_mode, divisor=divisor)
    lo = spinless_traces(oracle_lib, norb, np_, irvec, hr0, k[0], om - b, eta, divisor)
    hi = spinless_traces(oracle_lib, norb, np_, irvec, hr0, k[0], om + b, eta, divisor)
    right = 1 if spin_mode == 1 else 2
    scale = np.abs(np.concatenate([lo, hi])).max()
    tol = 1e-10 * scale
    assert np.allclose(dl[0], lo[:, 0] + hi[:, 0], rtol=0, atol=tol)
    assert np.allclose(dr[0], lo[:, 2] + hi[:, 2], rtol=0, atol=tol)
    for c in range(3):
        want_l = lo[:, 0] - hi[:, 0] if c == comp else 0.0
        want_r = lo[:, right] - hi[:, right] if c == comp else 0.0
        assert np.allclose(sl[c, 0], want_l, rtol=0, atol=tol), (c, sl[c, 0], want_l)
        assert np.allclose(sr[c, 0], want_r, rtol=0, atol=tol), (c, sr[c, 0], want_r)
    # the two right-hand choices differ here (Np = 2): the mode is not ignored
    assert np.abs(lo[:, 1] - hi[:, 1] - (lo[:, 2] - hi[:, 2])).max() > 100 * tol


def test_rotation_covariance(oracle_lib):
    mdl = M.synthetic_model(8, 1, 0.3, M.decay_l1, seed=13, soc_sigma=0.06)
    irvec, hr = mdl.surface_cell()
    rng = np.random.default_rng(5)
    a = rng.normal(size=4)
    a /= np.linalg.norm(a)
    u2 = a[0] * np.eye(2) + 1j * sum(a[i + 1] * PAULI[i] for i in range(3))  # SU(2)
    U = np.kron(u2, np.eye(4))
    hr_u = np.array([U @ h @ U.conj().T for h in hr])
    rot = np.array([[0.5 * np.trace(u2.conj().T @ PAULI[c] @ u2 @ PAULI[d]).real
                     for d in range(3)] for c in range(3)])
    assert np.allclose(rot @ rot.T, np.eye(3)) and np.linalg.det(rot) > 0.999
    k = np.array([[0.05, 0.4], [-0.3, 0.1]])
    om = np.array([-0.3, 0.25]) * M.EV2HARTREE
    eta = 3e-3 * M.EV2HARTREE
    top = np.arange(1, 9, dtype=np.int32)
    base = oracle_lib.spectral_spin(8, 2, irvec, hr, k, om, eta, top, top, spin_mode=2)
    rotd = oracle_lib.spectral_spin(8, 2, irvec, hr_u, k, om, eta, top, top, spin_mode=2)
    for side in (4, 5):
        want = np.einsum("cd,dkw->ckw", rot, base[side])
        scale = np.abs(base[side]).max()
        assert np.abs(rotd[side] - want).max() < 1e-9 * scale
        assert np.abs(rotd[side] - base[side]).max() > 1e-3 * scale  # the rotation is not trivial
