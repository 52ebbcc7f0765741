"""Writes the fixtures of the reference's microstructure_compression example
(examples/LIGGGHTS/INL/microstructure_compression) from its files:

* reference_data/microstructure_vox_1_occupancy.npz: the particle template
  vox_1.lmp (114,704 lines "x y z r type", every x, y, z an even integer in
  2..100 / 2..100 / 2..200, r = 1, type 2, lines in ascending (z, y, x)
  order) as a packed occupancy grid of the 2-unit lattice, shape (z, y, x) =
  (100, 50, 50): voxel (k, j, i) holds sphere (2i + 2, 2j + 2, 2k + 2), and
  the occupied voxels in C order are the file's lines in order;
* reference_data/microstructure_mesh_plate.stl: the plate mesh, unchanged;
* reference_gold/microstructure_compression_output.csv: gold/output.csv,
  unchanged (one row every 10,000 steps).

Run from the repository root, where /root/reference holds the reference.
"""
import os
import shutil

import numpy as np

SRC = "/root/reference/examples/LIGGGHTS/INL/microstructure_compression"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    a = np.loadtxt(os.path.join(SRC, "vox_1.lmp"))
    xyz = a[:, :3].astype(np.int64)
    assert (a[:, :3] == xyz).all() and (xyz % 2 == 0).all()
    assert (a[:, 3] == 1.0).all() and (a[:, 4] == 2.0).all()
    key = (xyz[:, 2] * 1000 + xyz[:, 1]) * 1000 + xyz[:, 0]
    assert (np.diff(key) > 0).all(), "lines not in (z, y, x) order"
    shape = (100, 50, 50)
    g = np.zeros(shape, bool)
    g[xyz[:, 2] // 2 - 1, xyz[:, 1] // 2 - 1, xyz[:, 0] // 2 - 1] = True
    assert g.sum() == len(a)
    np.savez_compressed(os.path.join(HERE, "reference_data", "microstructure_vox_1_occupancy.npz"),
                        occ=np.packbits(g.ravel()), shape=np.array(shape))
    shutil.copyfile(os.path.join(SRC, "mesh_plate.stl"),
                    os.path.join(HERE, "reference_data", "microstructure_mesh_plate.stl"))
    shutil.copyfile(os.path.join(SRC, "gold", "output.csv"),
                    os.path.join(HERE, "reference_gold", "microstructure_compression_output.csv"))


if __name__ == "__main__":
    main()
