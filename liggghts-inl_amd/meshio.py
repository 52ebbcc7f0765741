"""Triangle-mesh input for mesh walls (fix mesh/surface file ...).

ASCII and binary STL, read the way InputMeshTri::meshtrifile_stl does
(input_mesh_tri.cpp:305-590): facets in file order become triangles 0..n-1,
the three vertices in file order become nodes 0, 1, 2, the facet normal is
ignored (the mesh computes its own), coordinates are parsed as C doubles.
"""
import struct

import numpy as np


def read_stl(path):
    """Return the triangles of an STL file as a (ntri, 3, 3) float64 array."""
    with open(path, "rb") as f:
        data = f.read()
    head = data[:80].lstrip()
    if head[:5].lower() == b"solid" and b"facet" in data[:1024]:
        return _read_ascii(data.decode("ascii", errors="replace"))
    return _read_binary(data)


def _read_ascii(text):
    tris, cur = [], []
    in_facet = False
    for ln, line in enumerate(text.splitlines(), 1):
        w = line.split()
        if not w:
            continue
        k = w[0]
        if k == "facet":
            if in_facet:
                raise ValueError("STL line %d: new facet begins without closing prior facet" % ln)
            in_facet, cur = True, []
        elif k == "vertex":
            if not in_facet:
                raise ValueError("STL line %d: vertex outside a facet" % ln)
            if len(cur) == 3:
                raise ValueError("STL line %d: more than 3 vertices in a facet "
                                 "(only triangular meshes supported)" % ln)
            cur.append([float(w[1]), float(w[2]), float(w[3])])
        elif k == "endfacet":
            if not in_facet or len(cur) != 3:
                raise ValueError("STL line %d: facet without 3 vertices" % ln)
            tris.append(cur)
            in_facet = False
    if not tris:
        raise ValueError("STL: no facets")
    return np.asarray(tris, dtype=np.float64)


def _read_binary(data):
    if len(data) < 84:
        raise ValueError("STL: file too short")
    n = struct.unpack_from("<I", data, 80)[0]
    if len(data) < 84 + 50 * n:
        raise ValueError("STL: truncated binary file")
    rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", 9), ("a", "<u2")]),
                        count=n, offset=84)
    return rec["v"].astype(np.float64).reshape(n, 3, 3)


def write_stl(path, tris, name="mesh"):
    """ASCII STL with %.9e coordinates (used for generated geometries)."""
    tris = np.asarray(tris, dtype=np.float64)
    with open(path, "w") as f:
        f.write("solid %s\n" % name)
        for t in tris:
            nrm = np.cross(t[1] - t[0], t[2] - t[0])
            ln = np.linalg.norm(nrm)
            nrm = nrm / ln if ln > 0 else nrm
            f.write("  facet normal %.9e %.9e %.9e\n    outer loop\n" % tuple(nrm))
            for v in t:
                f.write("      vertex %.9e %.9e %.9e\n" % tuple(v))
            f.write("    endloop\n  endfacet\n")
        f.write("endsolid %s\n" % name)
