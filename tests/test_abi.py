"""The C-ABI library loads without a GPU and exports every symbol the header
declares; host-side setup works and reports errors through status codes."""
import os
import re

import pytest

import t9hip

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(REPO, "include", "t9hip.h")).read()
    return sorted(set(re.findall(r"\b(t9_\w+)\s*\(", text)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(t9hip.EXPORTS)


def test_library_exports_all_symbols():
    L = t9hip.lib()
    missing = [s for s in header_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_setup_error_is_a_status_code(tmp_path):
    with pytest.raises(t9hip.T9Error) as e:
        t9hip.System.from_files(str(tmp_path / "missing.xyz"))
    assert e.value.code == t9hip.T9_ERR_SETUP


def test_water_box_topology():
    s = t9hip.System.water_box(216, 6, 18.643, 7, key="ewald\newald-cutoff 7.0\n")
    assert s.n == 648
    ik, mdpu = s.scales()
    # three intramolecular pairs per water, all with m = d = p = 0, u = 1
    assert len(ik) == 648
    assert (mdpu[:, :3] == 0).all() and (mdpu[:, 3] == 1).all()
    frame = s.i32("frame", s.n)
    assert (frame[0::3] == 3).all() and (frame[1::3] == 2).all()  # O bisector, H z-then-x
    assert int(s.scalar("grid1")) == 24  # 1.2 points/Å rounded up to a smooth size


def test_ion_sites_keep_distance():
    import numpy as np
    s = t9hip.System.water_box(10000, 22, 66.9, 2024, 90, 90, key="ewald\newald-cutoff 7.0\npme-grid 80 80 80\n")
    x = s.f64("xyz", (s.n, 3))
    ions, O = x[30000:], x[0:30000:3]
    d = ions[:, None, :] - O[None]
    d -= 66.9 * np.round(d / 66.9)
    md = np.sqrt((d * d).sum(-1)).min(1)
    assert md.min() > 2.95
