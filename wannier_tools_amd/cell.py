"""The surface cell of the SURFACE card and H(R) remapped onto it.

Host-side geometry that runs once per model, before anything reaches the
device:

* ``surface_umatrix``  -- the SURFACE card's integer matrix U (rows: R1', R2',
  R3' in units of a1, a2, a3), with the reference's handedness fix and its
  search for a third vector when the first two rows do not span a cell of
  the crystal's volume (src/readinput.f90:1575-1668, 4247-4463).
* ``surface_cell_positions`` -- the atoms in fractional coordinates of the new
  cell, wrapped into [0, 1) and shifted so that the top-surface atom sits at
  0.95 along R3' (src/readinput.f90:1749-1781).
* ``remap_hr``  -- get_hmnr_cell (src/readHmnR.f90:578-766): every hopping
  <i, 0|H|j, R> moves to the new cell's lattice vector
  R' = U^-T (R + tau_j - tau_i) - (tau'_j - tau'_i), rounded, dropped when a
  component exceeds 8.
* ``surface_frame`` -- Urot (x along R1', z along R1' x R2') and the surface
  reciprocal vectors Ka2, Kb2 (src/readinput.f90:1816-1855).

Lattices are rows a1, a2, a3 in Cartesian coordinates (Bohr, as the reference
keeps them); positions are fractional.
"""
from __future__ import annotations

import dataclasses

import numpy as np

EPS9 = float(np.float32(1e-9))
MAX_IR = 8          # src/readHmnR.f90:606
SEARCH_RMAX = 6     # src/readinput.f90:4261


def _triple(r1, r2, r3) -> float:
    """R1 . (R2 x R3) (get_volume, src/readinput.f90:4814-4826)."""
    return float(np.dot(np.cross(r2, r3), r1))


def _find_third_vector(lat: np.ndarray, u: np.ndarray) -> np.ndarray:
    """FindTheThirdLatticeVector (src/readinput.f90:4247-4463): replace rows 1
    and 2 by the shortest lattice vectors along them, then take as row 3 the
    shortest lattice vector giving the smallest non-zero volume."""
    u = u.copy()
    rng = range(-SEARCH_RMAX, SEARCH_RMAX + 1)
    grid = [(i1, i2, i3) for i1 in rng for i2 in rng for i3 in rng if (i1, i2, i3) != (0, 0, 0)]
    for row in (0, 1):
        best, best_len = None, 9999999.0
        for v in grid:
            cross = np.cross(u[row], np.array(v, dtype=float))
            if np.abs(cross).sum() < EPS9 and np.dot(v, u[row]) > 0:
                length = float(np.linalg.norm(np.array(v, dtype=float) @ lat))
                if length < best_len:
                    best_len, best = length, v
        if best is not None:
            u[row] = best
    r1, r2 = u[0] @ lat, u[1] @ lat
    vols = {v: abs(_triple(r1, r2, np.array(v, dtype=float) @ lat)) for v in grid}
    smallest = min((x for x in vols.values() if x >= EPS9), default=9999999.0)
    lengths = {v: float(np.linalg.norm(np.array(v, dtype=float) @ lat)) for v in grid}
    shortest = min((lengths[v] for v in grid if abs(vols[v] - smallest) < EPS9), default=9999999.0)
    for v in grid:
        if abs(vols[v] - smallest) < EPS9 and abs(lengths[v] - shortest) < EPS9:
            u[2] = v
            break
    if _triple(*(u @ lat)) < 0:
        u[2] = -u[2]
    if abs(_triple(*(u @ lat)) - _triple(*lat)) > EPS9:
        raise NotImplementedError(
            "SURFACE card: no cell of the crystal's volume spans the first two vectors; "
            "the Miller-index construction is not supported")
    return u


def surface_umatrix(lat: np.ndarray, rows) -> np.ndarray:
    """U (3 x 3) from the SURFACE card's two or three rows (the third defaults
    to 0 0 1), as src/readinput.f90:1618-1668 settles it."""
    lat = np.asarray(lat, dtype=float)
    rows = np.asarray(rows, dtype=float)
    u = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    u[: rows.shape[0]] = rows
    r = u @ lat
    vol2 = _triple(*r)
    if vol2 < 0:
        u[2] = -u[2]
    if abs(abs(vol2) - abs(_triple(*lat))) > 0.001:
        u = _find_third_vector(lat, u)
    return u


def _home(x: np.ndarray) -> np.ndarray:
    """transformtohomecell (src/readinput.f90:4495-4522): each component into [0, 1)."""
    x = np.array(x, dtype=float)
    for i in range(3):
        while not (0.0 <= x[i] < 1.0):
            x[i] += 1.0 if x[i] < 0.0 else -1.0
    return x


def surface_cell_positions(u: np.ndarray, positions: np.ndarray, top_atom: int = 0) -> np.ndarray:
    """Fractional positions in the new cell (src/readinput.f90:1749-1781).
    top_atom is 1-based; 0 picks the atom with the largest third coordinate
    (the first of equals)."""
    uinv_t = np.linalg.inv(np.asarray(u, dtype=float)).T
    new = np.array([_home(uinv_t @ p) for p in np.asarray(positions, dtype=float)])
    if top_atom == 0:
        top_atom = 1
        for ia in range(1, new.shape[0]):
            if new[ia, 2] > new[top_atom - 1, 2]:
                top_atom = ia + 1
    shift = np.array([0.0, 0.0, 0.95 - new[top_atom - 1, 2]])
    return np.array([_home(p + shift) for p in new])


def orbital_atoms(nprojs, soc: bool) -> np.ndarray:
    """0-based atom index of every orbital (spinorbital_to_atom_index): the
    projectors atom by atom, repeated for spin down with SOC."""
    idx = np.repeat(np.arange(len(nprojs)), nprojs)
    return np.concatenate([idx, idx]) if soc else idx


def remap_hr(irvec: np.ndarray, hr: np.ndarray, ndegen: np.ndarray, orb_atom: np.ndarray,
             pos_old: np.ndarray, pos_new: np.ndarray, u: np.ndarray):
    """get_hmnr_cell (src/readHmnR.f90:578-766): (irvec_surfacecell (nr', 3),
    hr_surfacecell (nr', nw, nw)) with 1/ndegen applied.  R' vectors are listed
    with the first component fastest; a later hopping landing on an occupied
    (R', i, j) replaces the earlier one, as the reference assigns them."""
    uinv_t = np.linalg.inv(np.asarray(u, dtype=float)).T
    irvec = np.asarray(irvec)
    nw = hr.shape[1]
    tau = np.asarray(pos_old, dtype=float)[orb_atom]        # (nw, 3)
    taun = np.asarray(pos_new, dtype=float)[orb_atom]
    dtau = tau[None, :, :] - tau[:, None, :]                # [i, j] = tau_j - tau_i
    dtaun = taun[None, :, :] - taun[:, None, :]
    rp = np.empty((irvec.shape[0], nw, nw, 3), dtype=np.int64)
    for ir in range(irvec.shape[0]):
        rmn = irvec[ir][None, None, :] + dtau
        rnew = rmn @ uinv_t.T - dtaun
        rp[ir] = np.rint(rnew).astype(np.int64)            # ANINT (halves away from zero differ
        half = np.abs(rnew - np.trunc(rnew)) == 0.5        # from rint only at exact .5)
        rp[ir][half] = (np.trunc(rnew) + np.sign(rnew))[half].astype(np.int64)
    keep = np.all(np.abs(rp) <= MAX_IR, axis=-1)
    used = np.zeros((2 * MAX_IR + 1,) * 3, dtype=bool)
    sel = rp[keep] + MAX_IR
    used[sel[:, 0], sel[:, 1], sel[:, 2]] = True
    # order: nir3 outer, nir2, nir1 inner (src/readHmnR.f90:660-671)
    order = np.argwhere(used.transpose(2, 1, 0))[:, ::-1] - MAX_IR
    index = {tuple(v): k for k, v in enumerate(order)}
    out = np.zeros((order.shape[0], nw, nw), dtype=np.complex128)
    for ir in range(irvec.shape[0]):
        for j in range(nw):
            for i in range(nw):
                if keep[ir, i, j]:
                    out[index[tuple(rp[ir, i, j])], i, j] = hr[ir, i, j] / ndegen[ir]
    return order.astype(np.int32), out


def surface_frame(r1, r2):
    """(Urot, Ka2, Kb2) of src/readinput.f90:1816-1855."""
    r1, r2 = np.asarray(r1, float), np.asarray(r2, float)
    urot = np.zeros((3, 3))
    urot[0] = r1 / np.linalg.norm(r1)
    urot[2] = np.cross(r1, r2)
    urot[2] /= np.linalg.norm(urot[2])
    urot[1] = np.cross(urot[2], urot[0])
    urot[1] /= np.linalg.norm(urot[1])
    ra2 = urot[:2] @ r1
    rb2 = urot[:2] @ r2
    vol = abs(ra2[0] * rb2[1] - rb2[0] * ra2[1])
    if vol < 1e-6:
        raise ValueError("surface cell of zero area")
    ka2 = 2 * np.pi / vol * np.array([rb2[1], -rb2[0]])
    kb2 = 2 * np.pi / vol * np.array([-ra2[1], ra2[0]])
    return urot, ka2, kb2


@dataclasses.dataclass
class SurfaceCell:
    """The remapped model and the frame of one SURFACE card."""
    umatrix: np.ndarray
    irvec: np.ndarray
    hr: np.ndarray
    urot: np.ndarray
    ka2: np.ndarray
    kb2: np.ndarray


def build_surface_cell(lat_bohr, positions, nprojs, soc: bool, irvec, hr, ndegen,
                       surface_rows=((1, 0, 0), (0, 1, 0)), top_atom: int = 0) -> SurfaceCell:
    """The whole chain for one model: U, the new cell's positions, the remapped
    H(R) (hr in the units the caller uses, ndegen applied) and the frame."""
    lat = np.asarray(lat_bohr, dtype=float)
    u = surface_umatrix(lat, surface_rows)
    pos_new = surface_cell_positions(u, positions, top_atom)
    atoms = orbital_atoms(nprojs, soc)
    irv, h = remap_hr(irvec, hr, ndegen, atoms, positions, pos_new, u)
    r = u @ lat
    urot, ka2, kb2 = surface_frame(r[0], r[1])
    return SurfaceCell(umatrix=u, irvec=irv, hr=h, urot=urot, ka2=ka2, kb2=kb2)
