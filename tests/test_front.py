"""The atom-group side of the path and the correlate front end, on the CPU:
form-factor mapping (form_factor_table.c:138-339) against the fixture atom
tables of the reference's test PDBs, the C SASA against the Python one, the
Euler conversion (index.c:38-121) against closed-form rotations, and the
Euler-list snapping (tools/correlate.c:202-251) against the oracle on the
reference's own Euler list.  No GPU call is made here."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI_REF = 3.14159265358


class Vec3(C.Structure):
    _fields_ = [("X", C.c_double), ("Y", C.c_double), ("Z", C.c_double)]


class Mat3(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("m11", "m12", "m13", "m21", "m22", "m23", "m31", "m32", "m33")]


class AtomGroup(C.Structure):
    _fields_ = [("natoms", C.c_size_t), ("coords", C.POINTER(Vec3)),
                ("vdw_radius", C.POINTER(C.c_double)),
                ("residue_name", C.POINTER(C.c_char_p)), ("atom_name", C.POINTER(C.c_char_p))]


class Euler(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("z", "b1", "g1", "a2", "b2", "g2")]


@pytest.fixture(scope="module")
def L():
    import libfmftsaxs_amd as F
    lib = F.lib()
    lib.default_ff_table.restype = C.c_void_p
    lib.default_ff_table.argtypes = [C.c_char_p]
    lib.vacuum_ff.restype = C.c_double
    lib.vacuum_ff.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.dummy_ff.restype = C.c_double
    lib.dummy_ff.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.sxs_faccs.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
    lib.mol_atom_group_average_radius.restype = C.c_double
    lib.mol_atom_group_average_radius.argtypes = [C.c_void_p]
    lib.sxs_ft2euler.argtypes = [C.c_void_p] * 4
    lib.sxs_ft_file2euler_file.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p]
    lib.fsx_euler_list_read.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_double,
                                        C.POINTER(C.POINTER(C.c_int64)),
                                        C.POINTER(C.POINTER(C.c_int)),
                                        C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)]
    lib.fsx_scores_write.argtypes = [C.c_char_p] + [C.c_void_p] * 5 + [C.c_int]
    lib.fsx_free.argtypes = [C.c_void_p]
    return lib


def _group(xyz, vdw, res, name):
    n = len(vdw)
    keep = dict(coords=(Vec3 * n)(*[Vec3(*map(float, p)) for p in xyz]),
                vdw=np.ascontiguousarray(vdw, dtype=np.float64),
                res=(C.c_char_p * n)(*[r.encode() for r in res]),
                name=(C.c_char_p * n)(*[a.encode() for a in name]))
    ag = AtomGroup(n, keep["coords"], keep["vdw"].ctypes.data_as(C.POINTER(C.c_double)),
                   keep["res"], keep["name"])
    return ag, keep


def test_form_factor_mapping(L):
    t = L.default_ff_table(os.path.join(G, "pdb_formfactor_mapping_clean.prm").encode())
    assert t
    for nm in ["4g9s_r_native", "4g9s_l_native"]:
        d = np.load(os.path.join(G, f"atoms_{nm}.npz"))
        vdw = np.nan_to_num(d["vdw"])
        ag, keep = _group(d["xyz"], vdw, [str(x) for x in d["res"]], [str(x) for x in d["name"]])
        vac = np.array([L.vacuum_ff(t, C.byref(ag), i) for i in range(len(vdw))])
        dum = np.array([L.dummy_ff(t, C.byref(ag), i) for i in range(len(vdw))])
        mapped = ~np.isnan(d["vac"])
        assert mapped.mean() > 0.99
        assert np.array_equal(vac[mapped], d["vac"][mapped])
        assert np.array_equal(dum[mapped], d["dum"][mapped])
        assert np.all(vac[~mapped] == 0.0)
        assert L.mol_atom_group_average_radius(C.byref(ag)) == pytest.approx(vdw.mean(), 1e-15)


def test_sasa_matches_python_shrake_rupley(L):
    from libfmftsaxs_amd import synthetic as S
    mol = S.random_chain(300, 77)
    ag, keep = _group(mol.coords, mol.vdw, ["ALA"] * len(mol.vdw), ["CA"] * len(mol.vdw))
    frac = np.zeros(len(mol.vdw))
    L.sxs_faccs(frac.ctypes.data, C.byref(ag), 1.4)
    ref = S.shrake_rupley(mol, probe=1.4, npts=1500)
    assert np.abs(frac - ref).max() <= 3.0 / 1500
    assert np.mean(frac == ref) > 0.95
    assert 0.0 <= frac.min() and frac.max() <= 1.0


def _rot(a, b, g):
    """active zyz rotation R(alpha, beta, gamma), saxs_utils.c:65-79"""
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    return np.array([[cg * cb * ca - sg * sa, -sg * cb * ca - cg * sa, sb * ca],
                     [cg * cb * sa + sg * ca, -sg * cb * sa + cg * ca, sb * sa],
                     [-cg * sb, sg * sb, cb]])


def _pose(rng):
    z = float(rng.integers(20, 60))
    b1, b2 = rng.uniform(0.2, 2.9, 2)
    g1, a2, g2 = rng.uniform(0.1, 6.1, 3)
    vec = z * np.array([-np.sin(b1) * np.cos(g1), np.sin(b1) * np.sin(g1), np.cos(b1)])
    rm = _rot(0.0, b1, g1).T @ _rot(a2, b2, g2)
    return (z, b1, g1, a2, b2, g2), vec, rm


def test_ft2euler_recovers_angles(L):
    """sxs_ft2euler inverts the pose construction: translation z (cos/sin of
    beta1, gamma1) and ligand rotation R(0, b1, g1)^T R(a2, b2, g2)."""
    rng = np.random.default_rng(5)
    ref = Vec3(1.5, -2.0, 0.75)
    for _ in range(200):
        want, vec, rm = _pose(rng)
        tv = Vec3(vec[0] - ref.X, vec[1] - ref.Y, vec[2] - ref.Z)
        m = Mat3(*rm.reshape(-1))
        e = Euler()
        L.sxs_ft2euler(C.byref(e), C.byref(tv), C.byref(m), C.byref(ref))
        got = (e.z, e.b1, e.g1, e.a2, e.b2, e.g2)
        # angles past pi come back as 2 pi_ref - x: 2 (pi - pi_ref) ~ 2e-11 off
        assert np.allclose(got, want, rtol=0, atol=1e-9), (got, want)


def test_ft_file_to_euler_file(L, tmp_path):
    rng = np.random.default_rng(9)
    ref = Vec3(0.5, 0.25, -1.0)
    poses, rots = [], []
    for i in range(20):
        want, vec, rm = _pose(rng)
        poses.append((i, vec - [ref.X, ref.Y, ref.Z], want))
        rots.append(rm)
    rot_path, ft_path, eu_path = tmp_path / "rot.prm", tmp_path / "ft.txt", tmp_path / "eu.txt"
    with open(rot_path, "w") as f:
        for i, rm in enumerate(rots):
            f.write(f"{i} " + " ".join(f"{v:.17g}" for v in rm.reshape(-1)) + "\n")
    with open(ft_path, "w") as f:
        for i, tv, _ in poses[::-1]:
            f.write(f"{i} {tv[0]:.17g} {tv[1]:.17g} {tv[2]:.17g} 0 0 0 0 0 0\n")
    L.sxs_ft_file2euler_file(str(eu_path).encode(), str(ft_path).encode(),
                             str(rot_path).encode(), C.byref(ref))
    lines = open(eu_path).read().splitlines()
    assert len(lines) == 20
    for line, (i, _, want) in zip(lines, poses[::-1]):
        f = line.split("\t")
        assert int(f[0]) == i
        assert np.allclose([float(x) for x in f[1:]], want, atol=1.5e-3)


def test_euler_list_snapping_matches_oracle(L, tmp_path):
    """The reference's own Euler lines at z = 40 (euler_coords.000.00),
    snapped by the product and by the oracle (correlate.c:213-251)."""
    Lmax = 15
    zv = np.arange(1.0, 81.0)
    idx = C.POINTER(C.c_int64)()
    fid = C.POINTER(C.c_int)()
    order = C.POINTER(C.c_int)()
    n = C.c_int()
    rc = L.fsx_euler_list_read(os.path.join(G, "euler_z40.txt").encode(), zv.ctypes.data,
                               zv.size, Lmax, Lmax + 1, 2 * Lmax + 1, 2 * Lmax + 1, PI_REF,
                               C.byref(idx), C.byref(fid), C.byref(order), C.byref(n))
    assert rc == 0 and n.value == 52
    got = [idx[i] for i in range(n.value)]
    want, ids = [], []
    for line in open(os.path.join(G, "euler_z40.txt")):
        v = line.split()
        ids.append(int(v[0]))
        want += O.euler2index(*map(float, v[1:7]), zv, Lmax)
    assert got == want
    assert [fid[i] for i in range(n.value)] == ids
    assert [order[i] for i in range(n.value)] == list(range(52))
    # the score file of correlate.c:371-381
    sc = np.linspace(1.0, 2.0, n.value)
    out = tmp_path / "scores.txt"
    assert L.fsx_scores_write(str(out).encode(), C.cast(order, C.c_void_p),
                              C.cast(fid, C.c_void_p), sc.ctypes.data, sc.ctypes.data,
                              sc.ctypes.data, n.value) == 0
    first = open(out).readline()
    assert first == f"{0:<6d}\t{ids[0]}\t{sc[0]:.3f}\t{sc[0]:.3f}\t{sc[0]:.3f}\n"
    for p in (idx, fid, order):
        L.fsx_free(C.cast(p, C.c_void_p))
