"""The mixed-precision form of the atom expansion (fsx_atoms2spf_form with
FSX_MIXED_FP32): every j_l(q r) and f conj(Y_lm) split into three BF16
parts, the six leading part products on v_mfma_f32_16x16x32_bf16 with FP32
accumulation per 1024-atom chunk, the chunks summed in FP64.  Against the
oracle's FP64 expansion (pdb2spf.c:24-152 restated): every amplitude within
2e-6 of its (X, q) block's largest (the parts carry 24 bits; FP32 sums over
a chunk lose a few more).  And the precision form end to end: amplitudes
from this expansion scored by the mixed form, against the FP64 oracle with
FP64 amplitudes, at the bars of the mixed form (I(q) within 1e-5, chi within
1e-4, identical top-100 ranking)."""
import numpy as np
import pytest

import oracle as O
from tests.problem import build_problem
from tests.test_gpu_mixed import _I, same_ranking

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import libfmftsaxs_amd as F
    if F.device_count() < 1:
        pytest.fail("no HIP device visible")
    return F


@pytest.mark.parametrize("L,natoms,Q", [(6, 700, 26), (15, 700, 26), (25, 2500, 41)])
def test_mixed_atom_expansion_vs_oracle(F, L, natoms, Q):
    from libfmftsaxs_amd import synthetic as S
    rec = S.random_chain(natoms, 17)
    rec = rec.translated(-rec.center_of_extrema())
    rec.sa = S.shrake_rupley(rec)
    q = O.mkarray(0.0, 0.5, Q)
    print(f"L={L} atoms={natoms} Q={Q}: split-BF16 expansion", flush=True)
    rm_, im_, rmm = F.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L, mixed=True)
    ro, io, rmo = O.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
    assert rmm == rmo
    scale = np.maximum(np.abs(ro).max(axis=2, keepdims=True), np.abs(io).max(axis=2, keepdims=True))
    err = np.maximum(np.abs(rm_ - ro), np.abs(im_ - io)) / scale
    print(f"  max error relative to the block's largest amplitude {err.max():.2e}")
    assert err.max() <= 2e-6


def test_precision_form_end_to_end(F):
    """L = 15, 51 q on [0, 0.5], 19 betas, 64-point grids, 400 poses: both
    molecules expanded in split BF16 and scored in the mixed form, against the
    oracle's FP64 expansion and scoring."""
    L, nb, ng = 15, 19, 64
    zv = np.arange(25.0, 61.0)
    pb = build_problem(1500, 800, L, 0.5, 51, zv, 400, seed=7, nbeta=nb, na=ng, ng=ng)
    O.set_pi(O.PI_TRUNC)
    ro = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                               pb["params"], pb["qvals"], zv, L, nbeta=nb, na=ng, ng=ng,
                               want_cross=True, dense_threshold=1)
    rec, lig = pb["rec"], pb["lig"]
    print("precision form: split-BF16 expansion of both molecules", flush=True)
    Ar, Ai, _ = F.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, pb["qvals"], L,
                            mixed=True)
    Br, Bi, _ = F.atoms2spf(lig.coords, lig.vac, lig.dum, lig.sa, lig.vdw, pb["qvals"], L,
                            mixed=True)
    gp = F.ScoringProblem(L, pb["qvals"], zv, Ar, Ai, Br, Bi, pb["a"], pb["params"].s.mult,
                          pb["peak"], nbeta=nb, nalpha=ng, ngamma=ng)
    print("precision form: mixed scoring", flush=True)
    r = F.score(gp, pb["index"], want_cross=True, mixed=True)
    ok = ~np.isnan(ro["score"])
    assert np.array_equal(ok, ~np.isnan(r["score"]))
    relI = np.abs(_I(r["cross"][ok]) - _I(ro["cross"][ok])) / np.abs(_I(ro["cross"][ok]))
    relc = np.abs(r["score"][ok] - ro["score"][ok]) / ro["score"][ok]
    print(f"precision form vs FP64 oracle: max rel I(q) {relI.max():.2e}, chi {relc.max():.2e}")
    assert relI.max() <= 1e-5
    assert relc.max() <= 1e-4
    assert same_ranking(r["score"][ok], ro["score"][ok], min(100, int(ok.sum())))
