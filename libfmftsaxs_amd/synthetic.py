"""Deterministic synthetic workloads for the FFT dimer SAXS scoring path.

Molecules are compact random-walk chains: residue centres step 3.8 A, the
walk is confined to a sphere sized so that the radius of gyration follows
Rg ~ 2.2 * N_res**0.38 A, and each residue carries 7 or 8 heavy atoms
(7.8 on average) scattered within 2.4 A of its centre.  Elements are drawn
C/N/O/S at 63/17/19/1 %, mapped to the scalar form factors and vdW radii
the reference uses for those atom types (form_factor_table.c:25-49,
prms/atoms.0.0.6.prm...: C 1.87, N 1.83, O 1.55, S 1.89 A).
"""
import numpy as np

# (vacuum, dummy) zero-angle form factors, form_factor_table.c:25-49
FF = {
    "C": (5.9992, 5.49096),
    "N": (6.9946, 0.83166),
    "O": (7.9994, 3.04942),
    "S": (15.9998, 6.63324),
}
VDW = {"C": 1.87, "N": 1.83, "O": 1.55, "S": 1.89}
H2O_FF = 3.50968  # zero_ff[s_OH2], form_factor_table.c:21
ELEMENTS = ["C", "N", "O", "S"]
ELEMENT_P = [0.63, 0.17, 0.19, 0.01]


class Molecule:
    def __init__(self, coords, elements):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.elements = list(elements)
        self.vac = np.array([FF[e][0] for e in elements])
        self.dum = np.array([FF[e][1] for e in elements])
        self.vdw = np.array([VDW[e] for e in elements])
        self.sa = None

    @property
    def natoms(self):
        return self.coords.shape[0]

    def centroid(self):
        return self.coords.mean(axis=0)

    def center_of_extrema(self):
        return 0.5 * (self.coords.min(axis=0) + self.coords.max(axis=0))

    def translated(self, shift):
        m = Molecule(self.coords + np.asarray(shift), self.elements)
        m.sa = self.sa
        return m


def random_chain(natoms, seed):
    rng = np.random.default_rng(seed)
    nres = max(1, int(round(natoms / 7.8)))
    rg = 2.2 * nres ** 0.38
    rmax = np.sqrt(5.0 / 3.0) * rg
    pos = np.zeros((nres, 3))
    for i in range(1, nres):
        for _ in range(100):
            v = rng.normal(size=3)
            v *= 3.8 / np.linalg.norm(v)
            cand = pos[i - 1] + v
            if np.linalg.norm(cand) <= rmax:
                break
        pos[i] = cand
    counts = np.full(nres, natoms // nres)
    counts[: natoms - counts.sum()] += 1
    coords = []
    for i in range(nres):
        for _ in range(counts[i]):
            d = rng.normal(size=3)
            d *= rng.uniform(0.8, 2.4) / np.linalg.norm(d)
            coords.append(pos[i] + d)
    coords = np.array(coords[:natoms])
    elements = list(rng.choice(ELEMENTS, size=natoms, p=ELEMENT_P))
    # keep every atom off the z axis (the angular map is singular there)
    xy = np.hypot(coords[:, 0], coords[:, 1])
    coords[xy < 1e-3, 0] += 1e-3
    return Molecule(coords, elements)


def shrake_rupley(mol, probe=1.4, npts=96, seed=7, occluders="all"):
    """Fraction of solvent-accessible surface per atom (Shrake-Rupley, probe
    1.4 A, Fibonacci sphere of npts points on r_vdw + probe), in the
    normalisation sxs_faccs returns (saxs_utils.c:30-48): contact area over
    4 pi r_vdw^2, i.e. the unburied fraction of the probe sphere.  The
    neighbours of an atom are the atoms within R_i + R_j of it (no other
    atom can bury a point of its sphere); atoms are processed in chunks with
    their neighbour lists padded, the test per (point, neighbour) being the
    squared distance against R_j^2."""
    del seed
    # Fibonacci sphere
    k = np.arange(npts) + 0.5
    phi = np.arccos(1 - 2 * k / npts)
    th = np.pi * (1 + 5 ** 0.5) * k
    sph = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    X = mol.coords
    R = mol.vdw + probe
    n = X.shape[0]
    occl = np.ones(n, dtype=bool) if occluders == "all" else mol.vdw > 0
    # neighbours: atoms closer than R_i + R_j (only they can bury a point of
    # atom i's sphere; a generous margin keeps every such atom)
    from scipy.spatial import cKDTree
    pairs = cKDTree(X).query_pairs(2 * R.max() + 1e-6, output_type="ndarray")
    pi, pj = pairs[:, 0], pairs[:, 1]
    d = np.sqrt(((X[pi] - X[pj]) ** 2).sum(1))
    keep = d < R[pi] + R[pj] + 1e-6
    src = np.concatenate([pi[keep], pj[keep]])
    dst = np.concatenate([pj[keep], pi[keep]])
    sel = occl[dst]
    src, dst = src[sel], dst[sel]
    order = np.argsort(src, kind="stable")
    src, dst = src[order], dst[order]
    starts = np.searchsorted(src, np.arange(n + 1))
    nbl = [dst[starts[i]:starts[i + 1]] for i in range(n)]
    free = np.full(n, npts, dtype=np.int64)
    chunk = 64
    for c0 in range(0, n, chunk):
        idx = np.arange(c0, min(n, c0 + chunk))
        mx = max(1, max(nbl[i].size for i in idx))
        nb = np.zeros((idx.size, mx), dtype=np.int64)
        valid = np.zeros((idx.size, mx), dtype=bool)
        for r, i in enumerate(idx):
            nb[r, :nbl[i].size] = nbl[i]
            valid[r, :nbl[i].size] = True
        pts = X[idx][:, None, :] + R[idx][:, None, None] * sph[None, :, :]   # (c, npts, 3)
        xn = X[nb]                                                            # (c, mx, 3)
        # squared distances (c, npts, mx), summed over x, y, z in that order
        dx = pts[:, :, None, 0] - xn[:, None, :, 0]
        d2 = dx * dx
        dx = pts[:, :, None, 1] - xn[:, None, :, 1]
        d2 += dx * dx
        dx = pts[:, :, None, 2] - xn[:, None, :, 2]
        d2 += dx * dx
        buried = ((d2 < (R[nb] ** 2)[:, None, :]) & valid[:, None, :]).any(2)
        free[idx] = npts - buried.sum(1)
    # contact surface (the accessible fraction scaled to the vdW sphere,
    # sxs_faccs calls accs with cont_acc = 1), divided by 4 pi r^2 as in
    # saxs_utils.c:40-47; radius-0 atoms (hydrogens) carry no SASA
    return np.where(mol.vdw == 0, 0.0, free / npts)


def make_dimer_inputs(n_rec, n_lig, seed_rec, seed_lig, water=True):
    """Receptor centred on its centre of extrema, ligand on its centroid
    (correlate.c:86-89, 100-103)."""
    rec = random_chain(n_rec, seed_rec)
    lig = random_chain(n_lig, seed_lig)
    rec = rec.translated(-rec.center_of_extrema())
    lig = lig.translated(-lig.centroid())
    if water:
        rec.sa = shrake_rupley(rec)
        lig.sa = shrake_rupley(lig)
    return rec, lig


def random_grid_poses(n, znum, nbeta, na, ng, seed, distinct=True):
    """Uniform random grid points (z, b1, b2, a2, g1, g2) as int64 mixed-radix
    indices [z][b1][b2][a2][g1][g2]."""
    rng = np.random.default_rng(seed)
    total = znum * nbeta * nbeta * na * ng * ng
    if distinct:
        idx = rng.choice(total, size=n, replace=False) if total < 2 ** 62 else None
    else:
        idx = rng.integers(0, total, size=n)
    return np.sort(np.asarray(idx, dtype=np.int64))
