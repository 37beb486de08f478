"""Seeded scoring problems built with the oracle (test infrastructure)."""
import numpy as np

import oracle as O
from libfmftsaxs_amd import synthetic as S


def build_problem(n_rec, n_lig, L, qmax, qnum, zvals, npose, seed=0, nbeta=None,
                  na=None, ng=None, noise=0.01, water=True):
    """Synthetic receptor/ligand SPFs (oracle atom expansion), a random set
    of distinct grid poses, and an 'experimental' curve taken from the first
    pose's I(q) at c1 = c2 = 1 plus Gaussian noise (sigma = 5 % I)."""
    qv = O.mkarray(0.0, qmax, qnum)
    rec, lig = S.make_dimer_inputs(n_rec, n_lig, 1000 + seed, 2000 + seed, water=water)
    Ar, Ai, rmA = O.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, qv, L)
    Br, Bi, rmB = O.atoms2spf(lig.coords, lig.vac, lig.dum, lig.sa, lig.vdw, qv, L)
    nb = nbeta or L + 1
    na = na or 2 * L + 1
    ng = ng or 2 * L + 1
    zvals = np.asarray(zvals, dtype=np.float64)
    idx = S.random_grid_poses(npose, zvals.size, nb, na, ng, 3000 + seed)
    dummy = O.Params(np.ones(6 * qnum), 1.6, 1.0)
    r0 = O.compute_saxs_scores(idx[:1], Ar, Ai, Br, Bi, dummy, qv, zvals, L, nbeta=nb,
                               na=na, ng=ng, want_cross=True, dense_threshold=1)
    ct = r0["cross"][0]
    I = ct[0] - ct[1] + ct[2] + ct[3] - ct[4] + ct[5]
    rng = np.random.default_rng(4000 + seed)
    Iexp = I * (1 + noise * rng.normal(size=qnum))
    err = 0.05 * Iexp
    a = O.scoring_helper(qv, Iexp, err, qv)
    rm = (rmA * rec.natoms + rmB * lig.natoms) / (rec.natoms + lig.natoms)
    P = O.Params(a, rm, Iexp[0])
    return dict(qvals=qv, zvals=zvals, L=L, nbeta=nb, na=na, ng=ng, A_re=Ar, A_im=Ai,
                B_re=Br, B_im=Bi, rm=rm, a=a, peak=Iexp[0], params=P, index=idx,
                exp_q=qv.copy(), exp_in=Iexp, exp_err=err, rec=rec, lig=lig)


def small_problem(L=6, qnum=21, npose=40, seed=0):
    pb = build_problem(300, 200, L, 0.3, qnum, np.arange(20.0, 31.0), npose, seed=seed)
    r = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                              pb["params"], pb["qvals"], pb["zvals"], L, want_cross=True)
    pb["cross"] = r["cross"]
    pb["result"] = r
    return pb
