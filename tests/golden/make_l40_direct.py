"""Writes tests/golden/l40_direct.npz, the fixture of the maximum-order test
(test_max_order_vs_direct): the oracle's direct per-pose cross terms
(rotation, translation and pairing of the pose itself, no grid correlation)
at the largest expansion order the library accepts, L = 40, for two poses of a
small synthetic dimer.  A direct evaluation takes about half a minute per pose
at this order, which is why it is stored rather than recomputed on the GPU
box.  Run from the repository root:

    python tests/golden/make_l40_direct.py

Inputs (rebuilt by the test from the same seeds): receptor 200 / ligand 100
atoms (synthetic.make_dimer_inputs seeds 51 / 52, water shell), 9 q on
[0, 0.5], z = 20..39 A, the native grid of L = 40 (41 betas, 81-point alpha
and gamma), exact pi.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from libfmftsaxs_amd import synthetic as S  # noqa: E402

L, NB, NG = 40, 41, 81


def inputs():
    """The dimer's coefficients, q and z grids (shared with the test)."""
    O.set_pi(O.PI_EXACT)
    rec, lig = S.make_dimer_inputs(200, 100, 51, 52, water=True)
    q = O.mkarray(0.0, 0.5, 9)
    zv = np.arange(20.0, 40.0)
    Ar, Ai, rmA = O.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
    Br, Bi, rmB = O.atoms2spf(lig.coords, lig.vac, lig.dum, lig.sa, lig.vdw, q, L)
    rm = (rmA * rec.natoms + rmB * lig.natoms) / (rec.natoms + lig.natoms)
    return q, zv, Ar, Ai, Br, Bi, rm


def main():
    q, zv, Ar, Ai, Br, Bi, rm = inputs()
    rng = np.random.default_rng(4040)
    idx = np.sort(rng.choice(zv.size * NB * NB * NG ** 3, 2, replace=False)).astype(np.int64)
    cross = []
    for i in idx:
        z, b1, b2, a2, g1, g2 = O.disassemble_index(int(i), NB, NG, NG)
        cross.append(O.direct_pose_cross_terms(
            Ar, Ai, Br, Bi, L, q, zv[z], b1 * np.pi / (NB - 1), 2 * np.pi * g1 / NG,
            2 * np.pi * a2 / NG, b2 * np.pi / (NB - 1), 2 * np.pi * g2 / NG))
        print(f"pose {int(i)} done", flush=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "l40_direct.npz")
    np.savez_compressed(out, index=idx, cross=np.array(cross), q=q, zvals=zv, rm=rm)
    print("wrote", out)


if __name__ == "__main__":
    main()
