"""Mixed-precision form (FSX_MIXED_FP32: rotated coefficients, translation
blocks and the cross-term contraction in FP32; Bessel, Wigner-d, lane
reductions, per-pose accumulators and the fit in FP64) against the FP64
oracle.  Bars: max relative I(q) error <= 1e-5, max relative chi error
<= 1e-5 and the identical top-100 score ranking."""
import numpy as np
import pytest

import oracle as O
from tests.problem import build_problem

pytestmark = pytest.mark.gpu


def _I(c):
    return c[..., 0, :] - c[..., 1, :] + c[..., 2, :] + c[..., 3, :] - c[..., 4, :] + c[..., 5, :]


@pytest.mark.parametrize("L,Q,qmax,nat,nb,ng", [(10, 31, 0.3, (800, 500), 11, 21),
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
    assert relc.max() <= 1e-5
    k = min(100, int(ok.sum()))
    assert np.array_equal(np.argsort(r["score"][ok], kind="stable")[:k],
                          np.argsort(ro["score"][ok], kind="stable")[:k])
