"""Mixed-precision form (FSX_MIXED_FP32: rotated coefficients, translation
blocks, the cross-term contraction and its lane sums in FP32; Bessel,
Wigner-d, the per-pose accumulators and the fit in FP64) against the FP64
oracle.  Bars: max relative I(q) error <= 1e-5 and the identical top-100
score ranking (exact ties, scores within 1e-9, may come in either order); chi (the square root of the fitted objective) within 1e-4.
The chi bar follows from chi's conditioning: in the oracle itself, random
changes of 1e-6 relative in I(q) move chi by up to 4e-5 (L = 6 case below,
400 fits), because the L-BFGS-B stopping rule (pgtol 1e-5) then ends at a
different iterate.  The cases cover one and
several pairing slots per lane and gamma grids that four-row tiles do and do
not divide."""
import numpy as np
import pytest

import oracle as O
from tests.problem import build_problem

pytestmark = pytest.mark.gpu


def same_ranking(s_test, s_ref, k, tie=1e-9):
    """The top-k ranking of s_test equals that of s_ref, up to the order
    inside groups of exact ties (reference scores within tie relative of
    each other, e.g. poses that are images of one another): the sequences of
    tie groups of the two top-k lists are equal."""
    order = np.argsort(s_ref, kind="stable")
    srt = s_ref[order]
    group = np.empty(s_ref.size, dtype=np.int64)
    group[order] = np.concatenate([[0], np.cumsum(np.diff(srt) > tie * np.abs(srt[1:]))])
    return np.array_equal(group[np.argsort(s_test, kind="stable")[:k]], group[order[:k]])


def _I(c):
    return c[..., 0, :] - c[..., 1, :] + c[..., 2, :] + c[..., 3, :] - c[..., 4, :] + c[..., 5, :]


@pytest.mark.parametrize("L,Q,qmax,nat,nb,ng", [(6, 21, 0.3, (300, 200), 7, 13),
                                                (10, 31, 0.3, (800, 500), 11, 21),
                                                (15, 51, 0.5, (1500, 800), 19, 64)])
def test_mixed_vs_oracle(L, Q, qmax, nat, nb, ng):
    import libfmftsaxs_amd as F
    O.set_pi(O.PI_TRUNC)
    zv = np.arange(25.0, 61.0)
    pb = build_problem(nat[0], nat[1], L, qmax, Q, zv, 400, seed=7, nbeta=nb, na=ng, ng=ng)
    ro = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                               pb["params"], pb["qvals"], zv, L, nbeta=nb, na=ng, ng=ng,
                               want_cross=True, dense_threshold=1)
    gp = F.ScoringProblem(L, pb["qvals"], zv, pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                          pb["a"], pb["params"].s.mult, pb["peak"], nbeta=nb, nalpha=ng,
                          ngamma=ng)
    r = F.score(gp, pb["index"], want_cross=True, mixed=True)
    ok = ~np.isnan(ro["score"])
    assert np.array_equal(ok, ~np.isnan(r["score"]))
    Ig, Io = _I(r["cross"][ok]), _I(ro["cross"][ok])
    relI = np.abs(Ig - Io) / np.abs(Io)
    relc = np.abs(r["score"][ok] - ro["score"][ok]) / ro["score"][ok]
    print(f"L={L}: mixed max rel I(q) {relI.max():.2e}, max rel chi {relc.max():.2e}")
    assert relI.max() <= 1e-5
    assert relc.max() <= 1e-4
    assert same_ranking(r["score"][ok], ro["score"][ok], min(100, int(ok.sum())))
