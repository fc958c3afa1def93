"""CPU checks of the drop-in boundary: libwtsurf.so loads and exports every
entry point include/wtsurf.h declares, and the host mirror binds each one.
No device calls are made here."""
import os
import re

import pytest

from conftest import ROOT
from wannier_tools_amd import surface

HEADER = os.path.join(ROOT, "include", "wtsurf.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wtsurf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_surface():
    syms = declared_symbols()
    for s in ("wtsurf_ham_qlayer2qlayer", "wtsurf_surfgreen_1985", "wtsurf_spectral_grid",
              "wtsurf_set_model", "wtsurf_ctx_create"):
        assert s in syms


@pytest.mark.skipif(not os.path.exists(surface.LIB_PATH), reason="libwtsurf.so not built")
def test_library_exports_every_declared_symbol():
    lib = surface.load_library()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(declared_symbols()) == set(surface.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(surface.LIB_PATH), reason="libwtsurf.so not built")
def test_library_reports_largest_order():
    assert surface.max_ndim() >= 72


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        surface.load_library(str(tmp_path / "nope.so"))
