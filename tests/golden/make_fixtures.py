"""Writes the golden fixtures under tests/golden/ from the reference's own
test data (reference tree: tests/data/).  Run in a container where the
reference tree is mounted:

    python tests/golden/make_fixtures.py [/root/reference]

Fixtures are data only (inputs and expected outputs of the reference's test
suite, tests/saxs_test.c):
  ref_chi.txt            52-pose score/c1/c2 table at z = 40, L = 15 (saxs_test.c:177-363)
  ref_saxs.txt           experimental curve used there
  ref_profile.txt        I(q) of the receptor at c1 = c2 = 1 (saxs_test.c:129-175)
  ref_fitted_profile.txt fitted profile of the native dimer (saxs_test.c:365-407)
  euler_z40.txt          the Euler-list lines of euler_coords.000.00 with z = 40
  ref_spf.npz            receptor SPF coefficients written by the reference (ref_spf)
  pdb_formfactor_mapping_clean.prm  the (residue, atom) -> form-factor type mapping
                         file the reference's tests load (saxs_test.c:74)
  atoms_<name>.npz       atom tables of the test PDB files: coordinates, residue and
                         atom names, vdW radius (atoms.0.0.6.prm...), vacuum and
                         dummy zero-angle form factors of the mapped type
                         (pdb_formfactor_mapping_clean.prm, form_factor_table.c:25-49)
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# zero-angle form factors per type name (form_factor_table.c:25-49; vacuum, dummy)
FF = {"H": (0.999953, 1.7201), "C": (5.9992, 5.49096), "N": (6.9946, 0.83166),
      "O": (7.9994, 3.04942), "S": (15.9998, 6.63324), "P": (14.9993, 1.91382),
      "CH": (6.99915, 7.21106), "CH2": (7.99911, 8.93116), "CH3": (8.99906, 10.6513),
      "NH": (7.99455, 2.55176), "NH2": (8.99451, 4.27186), "NH3": (9.99446, 5.99196),
      "OH": (8.99935, 4.76952), "SH": (16.9998, 8.35334)}


def read_pdb(path):
    xyz, res, name = [], [], []
    for line in open(path):
        if line.startswith("ATOM") or line.startswith("HETATM"):
            name.append(line[12:16].strip())
            res.append(line[17:21].strip())
            xyz.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
    return np.array(xyz), res, name


def read_prm(path):
    rad = {}
    for line in open(path):
        f = line.split()
        if len(f) >= 7 and f[0] == "atom":
            rad[(f[2], f[3])] = float(f[5])
    return rad


def read_mapping(path):
    m = {}
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        m[(f[0], f[1])] = f[2]
    return m


def atoms_table(pdb, rad, mapping):
    xyz, res, name = read_pdb(pdb)
    r = np.array([rad.get((a, b), np.nan) for a, b in zip(res, name)])
    types = [mapping.get((a, b), "?") for a, b in zip(res, name)]
    vac = np.array([FF[t][0] if t in FF else np.nan for t in types])
    dum = np.array([FF[t][1] if t in FF else np.nan for t in types])
    return dict(xyz=xyz, res=np.array(res), name=np.array(name), vdw=r,
                ff_type=np.array(types), vac=vac, dum=dum)


def read_spf(path):
    with open(path) as f:
        L, qnum, rm = f.readline().split()
        L, qnum, rm = int(L), int(qnum), float(rm)
        v = np.loadtxt(f).reshape(qnum, (L + 1) ** 2, 6)
    re = np.stack([v[:, :, 0], v[:, :, 2], v[:, :, 4]])
    im = np.stack([v[:, :, 1], v[:, :, 3], v[:, :, 5]])
    return L, qnum, rm, re, im


def main(ref="/root/reference"):
    data = os.path.join(ref, "tests", "data")
    for src, dst in [("ref_chi", "ref_chi.txt"), ("ref_saxs", "ref_saxs.txt"),
                     ("ref_profile", "ref_profile.txt"),
                     ("ref_fitted_profile", "ref_fitted_profile.txt"),
                     ("pdb_formfactor_mapping_clean.prm", "pdb_formfactor_mapping_clean.prm")]:
        shutil.copyfile(os.path.join(data, src), os.path.join(HERE, dst))
    with open(os.path.join(data, "euler_coords.000.00")) as f, \
            open(os.path.join(HERE, "euler_z40.txt"), "w") as g:
        for line in f:
            z = float(line.split()[1])
            if 39.999 < z < 40.001:
                g.write(line)
    L, qnum, rm, re, im = read_spf(os.path.join(data, "ref_spf"))
    np.savez_compressed(os.path.join(HERE, "ref_spf.npz"), L=L, qnum=qnum, rm=rm, re=re, im=im)
    rad = read_prm(os.path.join(data, "atoms.0.0.6.prm.ms.3cap+0.5ace.Hr0rec"))
    mapping = read_mapping(os.path.join(data, "pdb_formfactor_mapping_clean.prm"))
    for nm in ["4g9s_r_native", "4g9s_l_moved", "4g9s_l_native", "4g9s_native_dimer"]:
        t = atoms_table(os.path.join(data, nm + ".pdb"), rad, mapping)
        miss_r = int(np.isnan(t["vdw"]).sum())
        miss_f = int(np.isnan(t["vac"]).sum())
        print(f"{nm}: {len(t['xyz'])} atoms, unmapped radius {miss_r}, unmapped ff {miss_f}")
        np.savez_compressed(os.path.join(HERE, f"atoms_{nm}.npz"), **t)


if __name__ == "__main__":
    main(*sys.argv[1:])
