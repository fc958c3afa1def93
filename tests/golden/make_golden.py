"""Write the golden vectors under tests/golden/ from the reference itself.

Runs only where the reference tree exists: it builds oracle/_ref/libwtref.so
(the reference's own Fortran, compiled unchanged by oracle/Makefile) and calls
its ham_qlayer2qlayer and surfgreen_1985 through the C-ABI harness
(oracle/ref_harness.f90).  The outputs are stored as .npz fixtures so the
tests run without the reference (and on machines that never see it).

    python tests/golden/make_golden.py [--only ndim216,ndim288]
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from wannier_tools_amd import model as M  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def model_digest(irvec: np.ndarray, hr: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(irvec, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(hr, dtype=np.complex128).tobytes())
    return h.hexdigest()


def small_np3_model() -> M.TBModel:
    """4 orbitals, R in [-2,2]^3: exercises Np = 3 block assembly (R3 = -2..2)."""
    return M.synthetic_model(4, 2, 0.4, M.decay_l2(1.2), seed=7, name="nw4_r2")


def traces(gll, grr, gb, nw, top, bot, divisor=1.0):
    n = gll.shape[0]
    dl = 0.0
    for io in top:
        dl = dl - gll[io - 1, io - 1].imag / divisor
    dr = 0.0
    for b in bot:
        io = n - nw + b
        dr = dr - grr[io - 1, io - 1].imag / divisor
    db = 0.0
    for i in range(n):
        db = db - gb[i, i].imag / divisor
    return dl, dr, db


def run_case(ref: oracle.RefHarness, name: str, mdl: M.TBModel, np_: int, kpts, omegas_ev,
             eta_ev: float, store_model: bool, full_first: bool = False,
             store_h: bool = True) -> None:
    irvec, hr = mdl.surface_cell()
    nw = mdl.num_wann
    n = nw * np_
    ref.set_model(nw, np_, irvec, hr)
    kpts = np.asarray(kpts, dtype=np.float64).reshape(-1, 2)
    omegas = np.asarray(omegas_ev, dtype=np.float64) * M.EV2HARTREE
    eta = eta_ev * M.EV2HARTREE
    nk, nom = kpts.shape[0], omegas.size
    h00 = np.zeros((nk, n, n), dtype=np.complex128)
    h01 = np.zeros((nk, n, n), dtype=np.complex128)
    dgll = np.zeros((nk, nom, n), dtype=np.complex128)
    dgrr = np.zeros_like(dgll)
    dgb = np.zeros_like(dgll)
    top = M.default_top_orbitals(nw)
    tr = np.zeros((3, nk, nom))
    full = None
    for ik in range(nk):
        a, b = ref.ham_qlayer2qlayer(kpts[ik])
        h00[ik], h01[ik] = a, b
        for iw in range(nom):
            gll, grr, gb = ref.surfgreen_1985(omegas[iw], eta, a, b)
            dgll[ik, iw] = np.diag(gll)
            dgrr[ik, iw] = np.diag(grr)
            dgb[ik, iw] = np.diag(gb)
            tr[:, ik, iw] = traces(gll, grr, gb, nw, top, top)
            if full_first and full is None:
                full = np.stack([gll, grr, gb])
    payload = dict(nw=nw, np=np_, kpts=kpts, omegas=omegas, eta=eta,
                   diag_gll=dgll, diag_grr=dgrr, diag_gb=dgb, dos_l=tr[0], dos_r=tr[1],
                   dos_bulk=tr[2], model_sha256=model_digest(irvec, hr))
    if store_h:
        payload.update(h00=h00, h01=h01)
    if store_model:
        payload.update(irvec=irvec, hr=hr)
    if full is not None:
        payload.update(full_g=full)
    path = os.path.join(OUT, f"{name}.npz")
    np.savez_compressed(path, **payload)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def mixed_ndim216_model() -> M.TBModel:
    """The Ndim = 216 member of the mixed-model set: 72 orbitals, R in [-2,2]^3
    (Np = 3)."""
    return M.synthetic_model(72, 2, 0.5, M.decay_l2(1.5), seed=43)


def main(only=None) -> None:
    oracle.build(ref=True)
    ref = oracle.RefHarness()
    if only:
        large(ref, only)
        return
    # Ndim = 16 (8 orbitals, Np = 2): k on the (-0.5,0)->(0.5,0) path, omega on
    # the surfstat grid of [-1, 1] eV with 401 points, eta = 1e-3 eV.
    kp = M.kpath_slab([((-0.5, 0.0), (0.5, 0.0))], 201)[[0, 37, 100, 163, 200]]
    om = M.omega_grid(-1.0, 1.0, 401)[[0, 57, 123, 200, 271, 333, 400]]
    run_case(ref, "ndim16", M.model_ndim16(), 2, kp, om, 1e-3, store_model=True, full_first=True)
    # Ndim = 72 (36 orbitals, Np = 2), 343 R vectors, eta = 1e-3 eV.
    kp72 = M.kpath_slab([((-0.5, 0.0), (0.5, 0.0))], 512)[[5, 300]]
    om72 = M.omega_grid(-1.0, 1.0, 512)[[40, 256, 470]]
    run_case(ref, "ndim72", M.model_ndim72(), 2, kp72, om72, 1e-3, store_model=False)
    # Np = 3 assembly (Ndim = 12), k off the high-symmetry lines.
    kp3 = np.array([[0.11, -0.23], [0.37, 0.41]])
    run_case(ref, "np3_ndim12", small_np3_model(), 3, kp3, [-0.4, 0.05, 0.6], 5e-3,
             store_model=True)
    large(ref, ("ndim216", "ndim288"))


def large(ref, names) -> None:
    """Orders beyond 80 (the kernels with the matrix in global memory).  No
    H00 / H01 (megabytes at these orders): the tests form them on the device,
    whose Fourier sum the smaller cases pin."""
    if "ndim216" in names:
        # the mixed-model set's largest member, k off the symmetry lines
        run_case(ref, "ndim216", mixed_ndim216_model(), 3, [[0.13, -0.21]], [-0.3, 0.05], 2e-3,
                 store_model=False, store_h=False)
    if "ndim288" in names:
        # the 144-orbital spinful k-plane model at omega = 0, eta = 5e-4 eV
        run_case(ref, "ndim288", M.model_ndim288(), 2, [[-0.25, 0.125], [0.3125, -0.4375]], [0.0],
                 5e-4, store_model=False, store_h=False)


if __name__ == "__main__":
    main(sys.argv[2].split(",") if len(sys.argv) > 2 and sys.argv[1] == "--only" else None)
