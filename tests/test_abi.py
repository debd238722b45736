"""The C ABI library loads and exports every symbol include/prc.h declares; host-only entry
points work and argument errors are reported without touching a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import prc_mlem_amd

HEADER = Path(__file__).resolve().parents[1] / "include" / "prc.h"


def declared_symbols():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(prc_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert set(declared_symbols()) == set(prc_mlem_amd.SIGNATURES)


def test_library_exports_every_declared_symbol(prc):
    L = C.CDLL(str(prc_mlem_amd.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_host_only_entry_points(prc, oracle_mod):
    assert prc.isotopes() == ["11C", "13N", "15O", "18F", "64Cu", "68Ga", "82Rb", "124I"]
    assert prc.lib().prc_version().startswith(b"prc-mi355x")
    for iso in prc.isotopes():
        # same FP32 host bisection as the oracle (prc.cu:83-91)
        assert prc.range_cutoff(iso) == oracle_mod.range_cutoff(iso)
    with pytest.raises(ValueError):
        prc.isotope_index("99Xx")


@pytest.mark.parametrize("kw", [dict(n_rays=3), dict(n_rays=0), dict(spacing=(0, 1, 1)), dict(air_cap_mm=0.0),
                                dict(support_rho=-1.0), dict(precision="bogus")])
def test_invalid_geometry_rejected_before_gpu(prc, kw):
    spacing = kw.pop("spacing", (1.0, 1.0, 1.0))
    rho = np.ones((2, 2, 2), np.float32)
    with pytest.raises((prc.PrcError, KeyError)) as e:
        prc.Projector(rho, spacing, **kw)
    if isinstance(e.value, prc.PrcError):
        assert e.value.code == -1
        assert "Invalid" in str(e.value)


def test_nonfinite_density_rejected_before_gpu(prc):
    rho = np.ones((2, 2, 2), np.float32)
    rho[1, 1, 1] = np.nan
    with pytest.raises(prc.PrcError) as e:
        prc.Projector(rho, (1.0, 1.0, 1.0))
    assert e.value.code == -1 and "Non-finite" in str(e.value)
