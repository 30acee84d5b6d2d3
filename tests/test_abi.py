"""The C-ABI library without a GPU: it loads, exports every symbol that
include/rt6d.h declares, and its host-only entry points work."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden_case, golden_data, golden_index

from raytracer_3d_amd import cases, rt6d


def _declared():
    with open(os.path.join(ROOT, "include", "rt6d.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(rt6d_\w+)\s*\(", src, re.M)))


def test_header_and_binding_agree():
    assert sorted(rt6d.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    lib = rt6d.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_and_device_count_do_not_fail():
    lib = rt6d.load_library()
    assert b"gfx950" in lib.rt6d_version()
    assert rt6d.device_count() >= 0


@pytest.mark.parametrize("entry", golden_index(), ids=lambda e: e["stem"])
def test_build_grid_matches_reference(entry):
    case = golden_case(entry)
    g = golden_data(entry)
    for ax in range(6):
        fixed = case.fixed[ax - 3] if ax >= 3 else 0.0
        got = np.stack(rt6d.build_grid(ax, case.spacing[ax], fixed))
        assert np.array_equal(got, g[f"grid_{cases.AXES[ax]}"]), ax


def test_bad_arguments_are_errors_not_crashes():
    lib = rt6d.load_library()
    with pytest.raises(rt6d.RT6DError):
        rt6d.build_grid(7, np.ones(4))
    with pytest.raises(rt6d.RT6DError):
        rt6d.build_grid(0, np.ones(0))
    if rt6d.device_count() == 0:
        with pytest.raises(rt6d.RT6DError, match="no HIP device"):
            rt6d.Model((4, 4, 4, 4, 4, 4))
    assert lib.rt6d_query(None, 1, None) != 0
