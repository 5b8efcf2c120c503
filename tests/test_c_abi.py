"""The C-ABI library loads without a GPU and exports every entry point that
include/tdyn.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "tdyn.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tdyn_[a-z_]+)\s*\(", src)))


def test_header_declares_the_operator_surface():
    names = _declared()
    for n in ["tdyn_batch_create", "tdyn_realize_acceleration", "tdyn_multiply_by_m_inv",
              "tdyn_calc_residual_force_ignoring_constraints", "tdyn_calc_articulated_body_inertias",
              "tdyn_rollout_begin", "tdyn_rollout_advance", "tdyn_rollout"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from simbody_amd import matter
    if not os.path.exists(matter.LIB_PATH):
        pytest.skip("libtdyn.so not built")
    lib = ctypes.CDLL(matter.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(matter.EXPORTS) == _declared()


def test_invalid_trees_are_rejected_on_the_host():
    """Argument errors are caught before any device work (no GPU needed)."""
    import numpy as np
    from simbody_amd import matter, models as M
    if not os.path.exists(matter.LIB_PATH):
        pytest.skip("libtdyn.so not built")
    L = matter.lib()
    spec = M.spatial_chain(4, seed=1)
    bad_parent = M.TreeSpec(**{**spec.__dict__, "parent": np.array([0, 2, 1, 2], np.int32)})
    bad_mass = M.TreeSpec(**{**spec.__dict__, "mass": np.array([1.0, -1.0, 1.0, 1.0])})
    bad_bushing = M.TreeSpec(**{**spec.__dict__, "mobilizer": np.array([0, M.BUSHING, 0, 0], np.int32)})
    for s, code in [(bad_parent, 1), (bad_mass, 1), (bad_bushing, 1)]:
        with pytest.raises(matter.TdynError) as e:
            _create(L, s)
        assert e.value.code == code


def _create(L, s):
    import numpy as np
    from simbody_amd import matter
    arrs = [np.ascontiguousarray(s.parent, np.int32), np.ascontiguousarray(s.mobilizer, np.int32)] + \
           [np.ascontiguousarray(x, np.float64) for x in (s.mass, s.com_B, s.inertia_B, s.X_PF, s.X_BM)]
    d = matter.TreeDescC()
    d.nb = len(arrs[0])
    d.parent = arrs[0].ctypes.data_as(matter._i32); d.mobilizer = arrs[1].ctypes.data_as(matter._i32)
    d.mass, d.com_B, d.inertia_B, d.X_PF, d.X_BM = (a.ctypes.data_as(matter._d) for a in arrs[2:])
    h = ctypes.c_void_p()
    matter._check(L.tdyn_batch_create(ctypes.byref(d), 1, None, 4, ctypes.byref(h)))


def test_extended_model_descriptions_are_checked_on_the_host():
    """Mobilizer codes, Screw pitch, reversal of Free and BendStretch bodies:
    rejected by tdyn_batch_create before any device work (no GPU needed)."""
    import numpy as np
    from simbody_amd import matter, models as M
    if not os.path.exists(matter.LIB_PATH):
        pytest.skip("libtdyn.so not built")
    L = matter.lib()

    def create(spec):
        keep = []
        d = matter.fill_tree_desc(matter.TreeDescC(), spec, keep)
        h = ctypes.c_void_p()
        matter._check(L.tdyn_batch_create(ctypes.byref(d), 1, None, 2, ctypes.byref(h)))

    spec = M.spatial_chain(4, seed=2)
    cases = [
        ({"mobilizer": np.array([0, 13, 0, 0], np.int32)}, "unknown mobilizer"),
        ({"mobilizer": np.array([0, M.BUSHING, 0, 0], np.int32)}, "Bushing mobilizer allowed on Ground only"),
        ({"mobilizer": np.array([M.BUSHING, 0, 0, 0], np.int32), "reversed": np.array([1, 0, 0, 0], np.int32)},
         "reversed Bushing"),
        ({"mobilizer": np.array([0, M.SCREW, 0, 0], np.int32), "pitch": None}, "screw_pitch"),
        ({"mobilizer": np.array([M.FREE, 0, 0, 0], np.int32), "reversed": np.array([1, 0, 0, 0], np.int32)},
         "reversed Free"),
        ({"mobilizer": np.array([0, M.BENDSTRETCH, 0, 0], np.int32), "reversed": np.array([0, 1, 0, 0], np.int32)},
         "reversed BendStretch"),
    ]
    for change, msg in cases:
        bad = M.TreeSpec(**{**spec.__dict__, **change})
        with pytest.raises(matter.TdynError, match=msg) as e:
            create(bad)
        assert e.value.code == 1
