"""The engine's host-side mesh topology (C-ABI dem_mesh_topology, no GPU)
against the oracle's restatement of SurfaceMesh::buildNeighbours and
handleCorner, on the reference's own STL floor, a generated hopper and small
hand-checkable meshes.  Bit-exact: these are integer flags."""
import os

import numpy as np
import pytest

from conftest import ROOT, load_engine
from oracle import Oracle
from test_oracle_mesh import base_spec, flat_grid, folded_pair, one_atom

GOLD = os.path.join(ROOT, "tests", "golden", "reference_gold")


def oracle_topology(tris, lo, hi):
    s = base_spec(box_lo=list(lo), box_hi=list(hi))
    o = Oracle(s, one_atom([0.5 * (a + b) for a, b in zip(lo, hi)]))
    o.add_mesh(tris, type=2)
    return o.mesh_topology()


def meshes():
    eng = load_engine()
    out = {}
    out["stl_floor"] = (eng.meshio.read_stl(os.path.join(GOLD, "rolling_resistance_bottom.stl")),
                        [-0.015, -0.005, -0.001], [0.035, 0.005, 0.01])
    out["flat_grid"] = (flat_grid(5, 0.01, flip=True), [-0.05] * 3, [0.05] * 3)
    out["folded"] = (folded_pair(), [-0.05] * 3, [0.05] * 3)
    spec, _ = eng.workloads.hopper(n=3000)
    out["hopper"] = (spec["meshes"][0]["nodes"], spec["box_lo"], spec["box_hi"])
    # a mesh crossing the box faces: corner activity of nodes outside stays on
    out["clipped_grid"] = (flat_grid(4, 0.02, x0=-0.01, y0=-0.01), [-0.005] * 3, [0.005] * 3)
    return out


@pytest.mark.parametrize("name", ["stl_floor", "flat_grid", "folded", "hopper", "clipped_grid"])
def test_topology_matches_oracle(name):
    eng = load_engine()
    tris, lo, hi = meshes()[name]
    got = eng.mesh_topology(tris, lo, hi)
    want = oracle_topology(tris, lo, hi)
    for k in ("edge_active", "corner_active", "obtuse", "nneigh"):
        assert np.array_equal(got[k], want[k]), k


def test_degenerate_triangles_skipped():
    eng = load_engine()
    tris = flat_grid(2, 0.01)
    bad = tris[0].copy()
    bad[1] = bad[0] + 1e-10    # two nodes equal within 1e-8
    t2 = np.concatenate([tris[:1], bad[None], tris[1:]])
    got = eng.mesh_topology(t2, [-0.05] * 3, [0.05] * 3)
    want = eng.mesh_topology(tris, [-0.05] * 3, [0.05] * 3)
    assert len(got["obtuse"]) == len(tris)
    for k in want:
        assert np.array_equal(got[k], want[k])
