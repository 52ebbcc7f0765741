"""read_data / write_data for atom_style sphere (granular): the text data
files a LIGGGHTS-INL script loads with `read_data` and writes with
`write_data`, read into and written from the atom dictionaries that
``engine.Simulation`` takes (x, v, omega, radius, rmass, type, tag, image).

Reading follows the reference on one process:
* header (ReadData::header, read_data.cpp:445-560): the first line is a
  title; anything from '#' on is dropped; "N atoms", "N atom types",
  "lo hi xlo xhi" (and y, z) are read; the header ends at the first section
  keyword;
* Atoms (Atom::data_atoms, atom.cpp:671-792): `id type diameter density x y
  z [ix iy iz]`, image flags when the line has more than 7 words; each
  position is remapped into the periodic box (Domain::remap, domain.cpp:
  1096-1175: whole periods added or removed, the image flag counted, then
  clamped to lo); an atom whose remapped position lies outside the box (the
  sub-domain of the only process, widened by 1e-6 of the period at periodic
  boundaries) is not created;
* per atom (AtomVecSphere::data_atom, atom_vec_sphere.cpp:1055-1100):
  radius = diameter / 2, rmass = 4/3 pi r^3 density (density itself when the
  radius is 0), velocities zero until a Velocities section sets them
  (AtomVecSphere::data_vel, :1129-1137: `id vx vy vz wx wy wz`).

Writing follows WriteData (write_data.cpp:245-420) with
AtomVecSphere::pack_data / write_data (atom_vec_sphere.cpp:1155-1222) and
write_vel (:1286-1293): density = rmass / (4/3 pi r^3), every real written
with "%-1.16e", so a file written here reads back to the same doubles.
"""
import math

import numpy as np

SECTIONS = ("Atoms", "Velocities", "Ellipsoids", "Lines", "Triangles", "Bodies", "Bonds", "Angles",
            "Dihedrals", "Impropers", "Masses", "Pair Coeffs", "PairIJ Coeffs", "Bond Coeffs",
            "Angle Coeffs", "Dihedral Coeffs", "Improper Coeffs", "BondBond Coeffs", "BondAngle Coeffs",
            "MiddleBondTorsion Coeffs", "EndBondTorsion Coeffs", "AngleTorsion Coeffs",
            "AngleAngleTorsion Coeffs", "BondBond13 Coeffs", "AngleAngle Coeffs")
EPSILON = 1.0e-6   # atom.cpp:81
NCOLS = 7          # id type diameter density x y z


class DataFileError(ValueError):
    pass


def _strip(line):
    k = line.find("#")
    return (line[:k] if k >= 0 else line).strip()


def _section_name(line):
    s = _strip(line)
    for name in SECTIONS:
        if s == name:
            return name
    return None


def _remap(x, img, lo, hi, periodic):
    """Domain::remap for one position (orthogonal box)."""
    x = list(x)
    img = list(img)
    for d in range(3):
        if not periodic[d]:
            continue
        prd = hi[d] - lo[d]
        while x[d] < lo[d]:
            x[d] += prd
            img[d] -= 1
        while x[d] >= hi[d]:
            x[d] -= prd
            img[d] += 1
        x[d] = max(x[d], lo[d])
    return x, img


def read_data(path, periodic=(0, 0, 0), dimension=3):
    """Returns dict(box_lo, box_hi, ntypes, natoms_header, atoms)."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    natoms = None
    ntypes = None
    lo = [None] * 3
    hi = [None] * 3
    k = 1   # line 0 is the title
    while k < len(lines):
        if _section_name(lines[k]):
            break
        s = _strip(lines[k])
        k += 1
        if not s:
            continue
        w = s.split()
        if "atom types" in s:
            ntypes = int(w[0])
        elif "atoms" in s:
            natoms = int(w[0])
        else:
            for d, key in enumerate(("xlo xhi", "ylo yhi", "zlo zhi")):
                if key in s:
                    lo[d], hi[d] = float(w[0]), float(w[1])
    if natoms is None or ntypes is None or None in lo or None in hi:
        raise DataFileError("data file header lacks atoms, atom types or box bounds")
    per = [bool(p) for p in periodic]
    eps = [(hi[d] - lo[d]) * EPSILON if per[d] else 0.0 for d in range(3)]
    sublo = [lo[d] - eps[d] for d in range(3)]
    subhi = [hi[d] + eps[d] for d in range(3)]
    rows = []
    vel = {}
    while k < len(lines):
        name = _section_name(lines[k])
        k += 1
        if name is None:
            if _strip(lines[k - 1]):
                raise DataFileError("Unknown identifier in data file: %s" % _strip(lines[k - 1]))
            continue
        # a blank line follows the keyword, then one line per atom
        while k < len(lines) and not _strip(lines[k]):
            k += 1
        if name == "Atoms":
            for _ in range(natoms):
                if k >= len(lines):
                    raise DataFileError("Unexpected end of data file")
                w = lines[k].split()
                k += 1
                if len(w) < NCOLS:
                    raise DataFileError("Incorrect atom format in data file")
                img = [0, 0, 0]
                if len(w) > NCOLS:
                    img = [int(w[-3]), int(w[-2]), int(w[-1])]
                x, img = _remap([float(w[4]), float(w[5]), float(w[6])], img, lo, hi, per)
                if not all(sublo[d] <= x[d] < subhi[d] for d in range(3)):
                    continue
                tag = int(w[0])
                if tag <= 0:
                    raise DataFileError("Invalid atom ID in Atoms section of data file")
                typ = int(w[1])
                if typ <= 0 or typ > ntypes:
                    raise DataFileError("Invalid atom type in Atoms section of data file")
                r = 0.5 * float(w[2])
                if r < 0.0:
                    raise DataFileError("Invalid radius in Atoms section of data file")
                dens = float(w[3])
                if dens <= 0.0:
                    raise DataFileError("Invalid density in Atoms section of data file")
                if r == 0.0:
                    m = dens
                elif dimension == 2:
                    m = math.pi * r * r * dens
                else:
                    m = 4.0 * math.pi / 3.0 * r * r * r * dens
                rows.append((tag, typ, r, m, x, img))
        elif name == "Velocities":
            if not rows and natoms:
                raise DataFileError("Must read Atoms before Velocities")
            for _ in range(natoms):
                if k >= len(lines):
                    raise DataFileError("Unexpected end of data file")
                w = lines[k].split()
                k += 1
                vel[int(w[0])] = [float(t) for t in w[1:7]]
        else:
            raise DataFileError("data file section %s is not read for atom_style sphere" % name)
    n = len(rows)
    atoms = dict(
        tag=np.array([r[0] for r in rows], dtype=np.int32),
        type=np.array([r[1] for r in rows], dtype=np.int32),
        radius=np.array([r[2] for r in rows], dtype=np.float64),
        rmass=np.array([r[3] for r in rows], dtype=np.float64),
        x=np.array([r[4] for r in rows], dtype=np.float64).reshape(n, 3),
        image=np.array([r[5] for r in rows], dtype=np.int32).reshape(n, 3),
        v=np.zeros((n, 3)),
        omega=np.zeros((n, 3)),
    )
    for i, t in enumerate(atoms["tag"]):
        if int(t) in vel:
            atoms["v"][i] = vel[int(t)][:3]
            atoms["omega"][i] = vel[int(t)][3:6]
    return dict(box_lo=lo, box_hi=hi, ntypes=ntypes, natoms_header=natoms, atoms=atoms)


def write_data(path, box_lo, box_hi, atoms, ntypes, timestep=0):
    """write_data of an atom dictionary (state() of a Simulation merged with
    its radius, rmass, type and tag); image flags default to zero."""
    n = len(atoms["radius"])
    img = np.asarray(atoms.get("image", np.zeros((n, 3), dtype=np.int32)))
    x = np.asarray(atoms["x"])
    v = np.asarray(atoms.get("v", np.zeros((n, 3))))
    w = np.asarray(atoms.get("omega", np.zeros((n, 3))))
    with open(path, "w") as fh:
        fh.write("LAMMPS data file via write_data, version liggghts-inl_amd, timestep = %d\n" % timestep)
        fh.write("\n%d atoms\n%d atom types\n\n" % (n, ntypes))
        for d, key in enumerate(("xlo xhi", "ylo yhi", "zlo zhi")):
            fh.write("%-1.16e %-1.16e %s\n" % (box_lo[d], box_hi[d], key))
        fh.write("\nAtoms\n\n")
        for i in range(n):
            r = float(atoms["radius"][i])
            m = float(atoms["rmass"][i])
            dens = m if r == 0.0 else m / (4.0 * math.pi / 3.0 * r * r * r)
            fh.write("%d %d %-1.16e %-1.16e %-1.16e %-1.16e %-1.16e %d %d %d\n"
                     % (int(atoms["tag"][i]), int(atoms["type"][i]), 2.0 * r, dens, x[i, 0], x[i, 1], x[i, 2],
                        int(img[i, 0]), int(img[i, 1]), int(img[i, 2])))
        fh.write("\nVelocities\n\n")
        for i in range(n):
            fh.write("%d %-1.16e %-1.16e %-1.16e %-1.16e %-1.16e %-1.16e\n"
                     % (int(atoms["tag"][i]), v[i, 0], v[i, 1], v[i, 2], w[i, 0], w[i, 1], w[i, 2]))
