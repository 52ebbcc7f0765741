"""Host-side pieces of batching and of the gradient front end (CPU only).

* the chain rule BatchObservables applies, d(Yeff, Geff, betaeff)/d(Y, e),
  against central differences of the engine's own table builder
  (dem_material_tables, global_properties.cpp:428-562), which runs on the
  host;
* batch_pours: instance sizes in [nmin, nmax], per-instance parameters in
  their ranges, no overlaps inside an instance, tags unique over the batch,
  and instance() returning exactly that instance's atoms and materials.
"""
import importlib

import numpy as np
import pytest

from conftest import load_engine


def test_table_jacobian_matches_table_builder():
    eng = load_engine()
    try:
        adj = importlib.import_module("liggghts_inl_amd.adjoint")
    except ImportError:
        pytest.skip("torch not importable")
    nu = 0.3
    for Y, e in ((1e6, 0.3), (5e6, 0.6), (1e7, 0.9)):
        dYeff, dGeff, dbeta = adj.table_jacobian(np.array([Y]), np.array([e]), nu)

        def tables(Yv, ev):
            Ye, Ge, be = eng.material_tables(1, np.array([0.0, Yv]), np.array([0.0, nu]),
                                             np.array([[0.0, 0.0], [0.0, ev]]))
            return Ye[1, 1], Ge[1, 1], be[1, 1]

        h = 1e-6 * Y
        p, m = tables(Y + h, e), tables(Y - h, e)
        assert abs((p[0] - m[0]) / (2 * h) - dYeff[0]) <= 1e-7 * dYeff[0]
        assert abs((p[1] - m[1]) / (2 * h) - dGeff[0]) <= 1e-7 * dGeff[0]
        h = 1e-6 * e
        p, m = tables(Y, e + h), tables(Y, e - h)
        assert abs((p[2] - m[2]) / (2 * h) - dbeta[0]) <= 1e-6 * abs(dbeta[0])


def test_batch_pours_structure():
    W = load_engine().workloads
    spec, atoms, inst, params = W.batch_pours(ninst=12, seed=3, nmin=300, nmax=900)
    counts = np.bincount(inst)
    assert len(counts) == 12 and counts.min() >= 300 and counts.max() <= 900
    assert np.all((params["Y"] >= 1e6) & (params["Y"] <= 1e7))
    assert np.all((params["mu"] >= 0.2) & (params["mu"] <= 0.7))
    assert np.all((params["e"] >= 0.3) & (params["e"] <= 0.9))
    assert len(np.unique(atoms["tag"])) == len(atoms["tag"])
    lo, hi = np.asarray(spec["box_lo"]), np.asarray(spec["box_hi"])
    from scipy.spatial import cKDTree
    for b in range(12):
        s1, a1 = W.instance(spec, atoms, inst, params, b)
        assert len(a1["x"]) == counts[b]
        assert np.all(a1["x"] - a1["radius"][:, None] >= lo) and np.all(a1["x"] + a1["radius"][:, None] <= hi)
        pr = cKDTree(a1["x"]).query_pairs(2 * a1["radius"].max(), output_type="ndarray")
        if len(pr):
            dist = np.linalg.norm(a1["x"][pr[:, 0]] - a1["x"][pr[:, 1]], axis=1)
            assert np.all(dist >= a1["radius"][pr[:, 0]] + a1["radius"][pr[:, 1]])
        assert s1["materials"]["Y"][1] == params["Y"][b] and s1["materials"]["mu"][1, 1] == params["mu"][b]
