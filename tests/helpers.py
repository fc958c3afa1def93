"""Shared checks for the parity tests."""
import hashlib

import numpy as np

from wannier_tools_amd import model as M


def model_digest(irvec, hr) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(irvec, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(hr, dtype=np.complex128).tobytes())
    return h.hexdigest()


def golden_model(g, name):
    """(nw, np, irvec, hr) of a golden case; regenerated models are checked
    against the digest stored with the fixture."""
    if "irvec" in g:
        irvec, hr = g["irvec"], g["hr"]
    else:
        mdl = {"ndim72": M.model_ndim72}[name]()
        irvec, hr = mdl.surface_cell()
    assert model_digest(irvec, hr) == str(g["model_sha256"]), "synthetic model generator changed"
    return int(g["nw"]), int(g["np"]), irvec, hr


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def chain_surface_g(z, e0, t):
    """Surface Green's function of a semi-infinite chain (onsite e0, hopping t):
    g = (z - e0 - sqrt((z-e0)^2 - 4t^2)) / (2 t^2) on the decaying branch."""
    x = z - e0
    s = np.sqrt(x * x - 4 * t * t + 0j)
    g1 = (x - s) / (2 * t * t)
    g2 = (x + s) / (2 * t * t)
    # physical branch: |t g| < 1 (decaying), Im g < 0 for Im z > 0
    return g1 if abs(t * g1) < 1 else g2


def chain_bulk_g(z, e0, t):
    """Bulk (infinite chain) diagonal Green's function 1/sqrt((z-e0)^2 - 4t^2),
    branch with Im < 0 for Im z > 0."""
    x = z - e0
    s = np.sqrt(x * x - 4 * t * t + 0j)
    g = 1.0 / s
    return g if g.imag <= 0 else -g
