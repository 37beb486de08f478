"""The mixed-precision form's fit on the GPU (k_fitg: L-BFGS-B with the
2 x 2 L-BFGS matrix held explicitly, one pose per lane group, the profile in
registers) against the same fit compiled for the host (fsx_fit_host_form
with FSX_MIXED_FP32), fed the device's own cross-term profiles, rescaled as
min_saxs.c:170-179 does and rounded to FP32 as the mixed form stores them for
its fit.  The two differ only in the order of the sums over
q (a lane's q block, then the butterfly across the group), so chi must agree
to 1e-9 relative and c1, c2 to 1e-5.  The q grids cover one register block
per lane with 8 lanes per fit (Q = 25), 64 lanes (Q = 150) and the form that
reads the profile per evaluation (Q = 400 > 64 lanes x 4 q); the boxes are
the reference's and the headline workload's."""
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


def _rescaled(ct, peak):
    """min_saxs.c:170-179: every term times peak / I(q_0) at c1 = 1, c2 = 1,
    rounded to FP32 as the mixed form stores the profiles its fit reads"""
    s = ct[0, 0] + ct[3, 0] + ct[5, 0] + ct[2, 0] - ct[1, 0] - ct[4, 0]
    return (ct * (peak / s)).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("L,Q,npose,seed", [(8, 25, 400, 71), (6, 150, 200, 72),
                                            (6, 400, 120, 73)])
@pytest.mark.parametrize("bounds", [((0.96, 1.04), (-2.0, 4.0)), ((0.95, 1.05), (0.0, 4.0))])
def test_device_group_fit_equals_host(F, L, Q, npose, seed, bounds):
    pb = build_problem(300, 200, L, 0.5, Q, np.arange(18.0, 30.0), npose, seed=seed)
    gp = F.ScoringProblem(pb["L"], pb["qvals"], pb["zvals"], pb["A_re"], pb["A_im"],
                          pb["B_re"], pb["B_im"], pb["a"], pb["params"].mult, pb["peak"],
                          c1=bounds[0], c2=bounds[1])
    print(f"L={L} Q={Q} bounds={bounds}: mixed form, lane-group fit", flush=True)
    r = F.score(gp, pb["index"], want_cross=True, mixed=True)
    ok = ~np.isnan(r["score"])
    assert ok.sum() > 0.9 * npose
    ds, d1, d2, ev = [], [], [], []
    for i in np.nonzero(ok)[0]:
        h = F.fit_host(_rescaled(r["cross"][i], pb["peak"]), pb["qvals"], pb["a"],
                       pb["params"].mult, c1=bounds[0], c2=bounds[1], mixed=True)
        ds.append(abs(r["score"][i] - h[0]) / h[0])
        d1.append(abs(r["c1"][i] - h[1]))
        d2.append(abs(r["c2"][i] - h[2]))
        ev.append(h[3])
    print(f"  device vs host: chi {max(ds):.1e}, c1 {max(d1):.1e}, c2 {max(d2):.1e}; "
          f"host evaluations {np.mean(ev):.2f} (max {max(ev):.0f}); device "
          f"{r['stats']['fit_evals'] / ok.sum():.2f} (max {r['stats']['fit_max_steps']})")
    assert max(ds) <= 1e-9
    assert max(d1) <= 1e-5 and max(d2) <= 1e-5
    assert r["stats"]["n_capped"] == 0
