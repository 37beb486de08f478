"""The fit kernel's wavefront-wide evaluation against its lane-per-fit one.

Once the pose queue is empty and at most FSX_FIT_WIDE lanes of a wavefront
still hold a fit, the wavefront evaluates those fits together: q across the
lanes, each q's terms formed by the operations of the sequential loop and
every sum added in q order by one lane.  The fits therefore visit the same
points and end with the same bits whether the wide form is never used (0),
always used (64) or used in the tail only (the default).  Checked for one
round of q (Q <= 64), several rounds with the moments in LDS (Q = 150) and
from global memory (Q = 400), in both forms (the mixed form with its
compact-form fit, FSX_FIT_COMPACT); the evaluation counts must agree as well.
"""
import numpy as np
import pytest

from tests.problem import build_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import libfmftsaxs_amd as F
    if F.device_count() < 1:
        pytest.fail("no HIP device visible")
    return F


def _gp(F, pb):
    return F.ScoringProblem(pb["L"], pb["qvals"], pb["zvals"], pb["A_re"], pb["A_im"], pb["B_re"],
                            pb["B_im"], pb["a"], pb["params"].mult, pb["peak"],
                            nbeta=pb["nbeta"], nalpha=pb["na"], ngamma=pb["ng"])


@pytest.mark.parametrize("L,Q,npose,seed", [(8, 25, 3000, 61), (6, 150, 600, 62),
                                            (6, 400, 200, 63)])
@pytest.mark.parametrize("mixed", [False, True])
def test_wide_evaluation_bit_identical(F, monkeypatch, L, Q, npose, seed, mixed):
    pb = build_problem(300, 200, L, 0.5, Q, np.arange(18.0, 30.0), npose, seed=seed)
    gp = _gp(F, pb)
    res = {}
    for wide in ("0", "64", None):
        if wide is None:
            monkeypatch.delenv("FSX_FIT_WIDE", raising=False)
        else:
            monkeypatch.setenv("FSX_FIT_WIDE", wide)
        print(f"L={L} Q={Q} mixed={mixed} FSX_FIT_WIDE={wide}", flush=True)
        res[wide] = F.score(gp, pb["index"], mixed=mixed, fit_compact=mixed)
    base = res["0"]
    ok = ~np.isnan(base["score"])
    assert ok.sum() > 0.9 * npose
    for wide in ("64", None):
        r = res[wide]
        for k in ("score", "c1", "c2"):
            assert np.array_equal(r[k], base[k], equal_nan=True), (wide, k)
        assert r["stats"]["fit_evals"] == base["stats"]["fit_evals"]
        assert r["stats"]["n_capped"] == base["stats"]["n_capped"]
