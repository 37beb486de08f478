"""GPU brute-force scorer (fsx_score_naive, tools/score_ft_naive.c:156-212)
against the oracle's real-space restatement of the same loop, and against
the FFT correlation on grid poses.  Bars:
  * cross-term profiles within 1e-10 of each (term, q) row's largest value
    (only the order of the atom sums and of the centroid differs);
  * score within 1e-7 relative, c1 and c2 within 1e-6;
  * naive vs FFT: I(q) within 1e-7 relative at L = 30, where truncation is
    below the accuracy of the reference's Bessel routine.
"""
import numpy as np
import pytest

import oracle as O
from tests.problem import build_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import libfmftsaxs_amd as F
    if F.device_count() < 1:
        pytest.fail("no HIP device visible")
    return F


def _rot(a, b, g):
    """active zyz rotation R(alpha, beta, gamma) in closed form (saxs_utils.c:65-79)"""
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    return np.array([[cg * cb * ca - sg * sa, -sg * cb * ca - cg * sa, sb * ca],
                     [cg * cb * sa + sg * ca, -sg * cb * sa + cg * ca, sb * sa],
                     [-cg * sb, sg * sb, cb]])


def _cat(rec, lig, f):
    return np.concatenate([getattr(rec, f), getattr(lig, f)])


def _oracle_naive_cross(rec, lig, q, L, pose, sa):
    z, b1, g1, a2, b2, g2 = pose
    x = np.vstack([rec.coords @ _rot(0.0, b1, g1).T,
                   lig.coords @ _rot(a2, b2, g2).T + np.array([0.0, 0.0, z])])
    x -= x.mean(axis=0)
    re, im, _ = O.atoms2spf(x, _cat(rec, lig, "vac"), _cat(rec, lig, "dum"), sa,
                            _cat(rec, lig, "vdw"), q, L)
    return O.spf2cross_terms(re, im, L, q.size)


def _mol(m):
    return (m.coords, m.vac, m.dum, m.vdw)


def _I(c):
    return c[..., 0, :] - c[..., 1, :] + c[..., 2, :] + c[..., 3, :] - c[..., 4, :] + c[..., 5, :]


def test_score_naive_vs_oracle(F):
    pb = build_problem(300, 200, 10, 0.3, 21, np.arange(20.0, 31.0), 4, seed=3)
    rec, lig, q, L, P = pb["rec"], pb["lig"], pb["qvals"], pb["L"], pb["params"]
    rng = np.random.default_rng(8)
    poses = np.column_stack([rng.uniform(18, 40, 12), rng.uniform(0, np.pi, 12),
                             rng.uniform(0, 2 * np.pi, 12),
                             rng.uniform(0, 2 * np.pi, 12), rng.uniform(0, np.pi, 12),
                             rng.uniform(0, 2 * np.pi, 12)])
    sa = _cat(rec, lig, "sa")
    print("naive: 12 poses, rec 300 / lig 200 atoms, L 10, Q 21", flush=True)
    r = F.score_naive(_mol(rec), _mol(lig), q, L, pb["a"], P.mult, poses, sa=sa,
                      want_cross=True)
    for i, p in enumerate(poses):
        co = _oracle_naive_cross(rec, lig, q, L, p, sa)
        scale = np.abs(co).max(axis=1, keepdims=True)
        assert np.all(np.abs(r["cross"][i] - co) <= 1e-10 * scale)
        sc, c1, c2, _, _ = O.fit_one(co, q, P, rescale=False)
        assert abs(r["score"][i] - sc) <= 1e-7 * sc
        assert abs(r["c1"][i] - c1) <= 1e-6 and abs(r["c2"][i] - c2) <= 1e-6


def test_score_naive_without_water(F):
    """sa = NULL drops the water shell: the W rows are zero."""
    pb = build_problem(120, 80, 6, 0.3, 11, np.arange(20.0, 24.0), 2, seed=4)
    rec, lig, q = pb["rec"], pb["lig"], pb["qvals"]
    poses = np.array([[22.0, 0.4, 1.0, 2.0, 1.1, 5.0]])
    r = F.score_naive(_mol(rec), _mol(lig), q, 6, pb["a"], pb["params"].mult, poses,
                      want_cross=True)
    assert np.all(r["cross"][0, [2, 4, 5]] == 0.0)
    co = _oracle_naive_cross(rec, lig, q, 6, poses[0], None)
    scale = np.abs(co).max(axis=1, keepdims=True) + 1e-300
    assert np.all(np.abs(r["cross"][0] - co) <= 1e-10 * scale)


def test_score_naive_matches_fft_on_grid_poses(F):
    """A grid pose of the correlation is the real-space pose with alpha2 and
    gamma2 mirrored to 2 pi - angle (correlate.c:222-223)."""
    from libfmftsaxs_amd import synthetic as S
    rec, lig = S.make_dimer_inputs(60, 40, 11, 12, water=True)
    L, N = 30, 61
    q = O.mkarray(0.0, 0.15, 6)
    Ar, Ai, _ = F.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
    Br, Bi, _ = F.atoms2spf(lig.coords, lig.vac, lig.dum, lig.sa, lig.vdw, q, L)
    z = np.array([25.0])
    grid = [(7, 11, 23, 19, 41), (3, 59, 5, 29, 9), (15, 0, 60, 1, 30)]
    idx = np.array([O.assemble_index(0, b1, b2, a2, g1, g2, L + 1, N, N)
                    for b1, g1, a2, b2, g2 in grid], dtype=np.int64)
    pb = F.ScoringProblem(L, q, z, Ar, Ai, Br, Bi, np.ones(6 * q.size), 0.01, 1.0)
    print("naive vs FFT: L 30, 3 grid poses", flush=True)
    rf = F.score(pb, idx, want_cross=True)
    poses = np.array([(25.0, b1 * np.pi / L, 2 * np.pi * g1 / N, 2 * np.pi - 2 * np.pi * a2 / N,
                       b2 * np.pi / L, 2 * np.pi - 2 * np.pi * g2 / N)
                      for b1, g1, a2, b2, g2 in grid])
    rn = F.score_naive(_mol(rec), _mol(lig), q, L, np.ones(6 * q.size), 0.01, poses,
                       sa=_cat(rec, lig, "sa"), want_cross=True)
    In, If = _I(rn["cross"]), _I(rf["cross"])
    rel = np.abs(In - If).max(axis=1) / np.abs(If).max(axis=1)
    print("naive vs FFT I(q) rel:", rel)
    assert rel.max() < 1e-7
