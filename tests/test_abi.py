"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(dem_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("dem_ctx_create", "dem_upload_state", "dem_set_materials", "dem_step", "dem_setup",
              "dem_compute_forces", "dem_download", "dem_get_pairs", "dem_last_error", "dem_ctx_destroy"):
        assert n in names


def test_library_exports_all_declared(eng):
    lib = ctypes.CDLL(eng.library_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_material_tables_host_helper_matches_oracle(eng, oracle_mod):
    """dem_material_tables is host code: compare with the oracle's tables."""
    import numpy as np

    nt = 2
    Y = np.array([0, 5e6, 2e7])
    nu = np.array([0, 0.3, 0.22])
    e = np.array([[0, 0, 0], [0, 0.9, 0.5], [0, 0.5, 0.3]])
    Ye, Ge, be = eng.material_tables(nt, Y, nu, e)
    o = oracle_mod.Oracle(dict(box_lo=[0, 0, 0], box_hi=[1, 1, 1], periodic=[0, 0, 0], dt=1e-6, skin=1e-3,
                               materials=dict(ntypes=nt, Y=Y, nu=nu, e=e, mu=np.zeros((3, 3)))),
                          dict(x=np.zeros((1, 3)), v=np.zeros((1, 3)), omega=np.zeros((1, 3)),
                               radius=np.ones(1) * 1e-3, rmass=np.ones(1), type=[1], tag=[1]))
    Yo = np.zeros((3, 3))
    Go = np.zeros((3, 3))
    bo = np.zeros((3, 3))
    dp = ctypes.POINTER(ctypes.c_double)
    o.L.or_get_material_tables(o.h, Yo.ctypes.data_as(dp), Go.ctypes.data_as(dp), bo.ctypes.data_as(dp))
    assert np.array_equal(Ye, Yo) and np.array_equal(Ge, Go) and np.array_equal(be, bo)


def test_no_device_gives_clean_error(eng):
    """Without a GPU the context creation fails loudly (no CPU fallback)."""
    try:
        import torch

        if torch.cuda.is_available():
            return
    except Exception:
        pass
    import pytest

    spec = dict(box_lo=[0, 0, 0], box_hi=[1, 1, 1], periodic=[0, 0, 0], dt=1e-6, skin=1e-3,
                materials=dict(ntypes=1, Y=[0, 5e6], nu=[0, 0.3], e=[[0.9, 0.9], [0.9, 0.9]],
                               mu=[[0.5, 0.5], [0.5, 0.5]]))
    import numpy as np

    atoms = dict(x=np.zeros((1, 3)), v=np.zeros((1, 3)), omega=np.zeros((1, 3)), radius=np.ones(1) * 1e-3,
                 rmass=np.ones(1), type=[1], tag=[1])
    with pytest.raises(eng.DemError):
        eng.Simulation(spec, atoms)
