"""Edge shapes of the HIP scoring path against the oracle: the smallest
expansion orders (L = 1..4, down to three translation entries per row and
two pairing slots), a single pose, a single z plane, and alpha / gamma
grids of one point.  Bars as in the parity and mixed-form tests: the FP64
form within 1e-10 relative in I(q), scores within 1e-7, c1 / c2 within 1e-6
and the identical top ranking; the mixed form within 1e-5 in I(q), chi
within 1e-4 and the same ranking up to exact ties.  Q >= 3 throughout: with
two q points the fitted objective of some poses rounds below zero in the
oracle itself (chi = NaN), which says nothing about the path."""
import numpy as np
import pytest

import oracle as O
from tests.problem import build_problem
from tests.test_gpu_mixed import same_ranking
from tests.test_gpu_parity import _I, _gpu_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import libfmftsaxs_amd as F
    if F.device_count() < 1:
        pytest.fail("no HIP device visible")
    return F


def _check_fp64(r, ro):
    """The parity bars of the FP64 form, with the ranking compared up to
    exact ties (at small orders distinct grid poses can be images of one
    another and score the same)."""
    ok = ~np.isnan(ro["score"])
    assert np.array_equal(ok, ~np.isnan(r["score"]))
    cg, co = r["cross"][ok], ro["cross"][ok]
    Ig, Io = _I(cg), _I(co)
    assert (np.abs(Ig - Io) / np.abs(Io)).max() <= 1e-10
    assert (np.abs(cg - co) / np.abs(co).max(axis=2, keepdims=True)).max() <= 1e-10
    s_g, s_o = r["score"][ok], ro["score"][ok]
    assert np.all(np.abs(s_g - s_o) <= 1e-7 * np.abs(s_o))
    assert np.all(np.abs(r["c1"][ok] - ro["c1"][ok]) <= 1e-6)
    assert np.all(np.abs(r["c2"][ok] - ro["c2"][ok]) <= 1e-6)
    assert same_ranking(s_g, s_o, min(100, int(ok.sum())))


def _check_mixed(r, ro):
    ok = ~np.isnan(ro["score"])
    assert np.array_equal(ok, ~np.isnan(r["score"]))
    Ig, Io = _I(r["cross"][ok]), _I(ro["cross"][ok])
    assert (np.abs(Ig - Io) / np.abs(Io)).max() <= 1e-5
    assert (np.abs(r["score"][ok] - ro["score"][ok]) / ro["score"][ok]).max() <= 1e-4
    assert same_ranking(r["score"][ok], ro["score"][ok], min(100, int(ok.sum())))


@pytest.mark.parametrize("L,Q,zn,npose", [(1, 4, 3, 24), (2, 5, 3, 40), (3, 9, 2, 60),
                                          (4, 12, 4, 60)])
def test_small_orders_vs_oracle(F, L, Q, zn, npose):
    pb = build_problem(200, 150, L, 0.3, Q, np.arange(20.0, 20.0 + zn), npose, seed=70 + L)
    ro = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                               pb["params"], pb["qvals"], pb["zvals"], L, want_cross=True)
    assert not np.isnan(ro["score"]).any()
    gp = _gpu_problem(F, pb)
    print(f"L={L} Q={Q} zn={zn} n={npose}: FP64 form", flush=True)
    _check_fp64(F.score(gp, pb["index"], want_cross=True), ro)
    print(f"L={L} Q={Q} zn={zn} n={npose}: mixed form", flush=True)
    _check_mixed(F.score(gp, pb["index"], want_cross=True, mixed=True), ro)


def test_single_pose_and_single_plane(F):
    L = 6
    pb = build_problem(300, 200, L, 0.3, 21, np.array([24.0]), 30, seed=90)
    ro = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                               pb["params"], pb["qvals"], pb["zvals"], L, want_cross=True)
    gp = _gpu_problem(F, pb)
    print("one z plane, 30 poses", flush=True)
    _check_fp64(F.score(gp, pb["index"], want_cross=True), ro)
    for k in (0, 17):
        print(f"one pose (pose {k})", flush=True)
        one = pb["index"][k:k + 1]
        r1 = F.score(gp, one, want_cross=True)
        r1o = {key: ro[key][k:k + 1] for key in ("score", "c1", "c2", "cross")}
        _check_fp64(r1, r1o)
        assert np.array_equal(r1["score"], F.score(gp, pb["index"], want_cross=True)["score"][k:k + 1])


@pytest.mark.parametrize("na,ng,nb,npose", [(1, 1, 2, 8), (1, 5, 3, 20), (4, 1, 2, 20)])
def test_one_point_angle_grids_vs_oracle(F, na, ng, nb, npose):
    """alpha2 or gamma grids of a single point (every pose at angle 0), and
    two or three beta values."""
    L = 4
    O.set_pi(O.PI_TRUNC)
    pb = build_problem(200, 150, L, 0.3, 9, np.arange(20.0, 23.0), npose, seed=80, nbeta=nb,
                       na=na, ng=ng)
    ro = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                               pb["params"], pb["qvals"], pb["zvals"], L, nbeta=nb, na=na, ng=ng,
                               want_cross=True, dense_threshold=1)
    gp = _gpu_problem(F, pb)
    print(f"na={na} ng={ng} nbeta={nb}: FP64 form", flush=True)
    _check_fp64(F.score(gp, pb["index"], want_cross=True), ro)
    print(f"na={na} ng={ng} nbeta={nb}: mixed form", flush=True)
    _check_mixed(F.score(gp, pb["index"], want_cross=True, mixed=True), ro)


def test_batch_with_small_and_empty_complexes(F):
    """fsx_score_batch over an L = 1 complex, an empty pose list, an L = 3
    complex and a single pose: every non-empty complex gets exactly the
    results of its own fsx_score call, in both forms."""
    pbs = [build_problem(200, 150, 1, 0.3, 4, np.arange(20.0, 23.0), 24, seed=71),
           build_problem(200, 150, 3, 0.3, 9, np.arange(20.0, 22.0), 60, seed=73),
           build_problem(300, 200, 6, 0.3, 21, np.array([24.0]), 30, seed=90)]
    gps = [_gpu_problem(F, pb) for pb in pbs]
    probs = [gps[0], gps[1], gps[1], gps[2]]
    poses = [pbs[0]["index"], pbs[1]["index"][:0], pbs[1]["index"], pbs[2]["index"][3:4]]
    for mixed in (False, True):
        print(f"batch of 4 (one empty), mixed={mixed}", flush=True)
        rb = F.score_batch(probs, poses, want_cross=True, mixed=mixed)
        assert rb[1]["score"].size == 0
        for gp, p, r in zip(probs, poses, rb):
            if p.size == 0:
                continue
            rs = F.score(gp, p, want_cross=True, mixed=mixed)
            for k in ("score", "c1", "c2", "cross"):
                assert np.array_equal(r[k], rs[k], equal_nan=True), k
            assert not np.isnan(r["score"]).any()


def test_dense_buckets_vs_oracle(F):
    """Hundreds of poses in the buckets of one cross-term tile (one z, one
    beta2, gamma2 rows 0..3), so that a tile's poses run in several batches
    of the pose loop, followed by a z with a few poses: both forms against
    the oracle."""
    L = 6
    pb = build_problem(300, 200, L, 0.3, 21, np.arange(20.0, 23.0), 10, seed=95)
    nb, na, ng, zn = pb["nbeta"], pb["na"], pb["ng"], pb["zvals"].size
    rng = np.random.default_rng(96)

    def idx(z, b1, b2, a2, g1, g2):
        return ((((z * nb + b1) * nb + b2) * na + a2) * ng + g1) * ng + g2

    poses = []
    for g2, m in ((2, 300), (0, 150), (3, 40), (1, 5)):
        pick = rng.choice(nb * na * ng, size=m, replace=False)
        b1, r = np.divmod(pick, na * ng)
        a2, g1 = np.divmod(r, ng)
        poses.append(idx(1, b1, 4, a2, g1, g2))
    pick = rng.choice(nb * na * ng, size=7, replace=False)
    b1, r = np.divmod(pick, na * ng)
    a2, g1 = np.divmod(r, ng)
    poses.append(idx(2, b1, 4, a2, g1, 2))
    index = rng.permutation(np.concatenate(poses).astype(np.int64))
    assert index.max() < zn * nb * nb * na * ng * ng
    ro = O.compute_saxs_scores(index, pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                               pb["params"], pb["qvals"], pb["zvals"], L, want_cross=True)
    gp = _gpu_problem(F, pb)
    print(f"{index.size} poses, 495 in one tile: FP64 form", flush=True)
    _check_fp64(F.score(gp, index, want_cross=True), ro)
    print(f"{index.size} poses, 495 in one tile: mixed form", flush=True)
    _check_mixed(F.score(gp, index, want_cross=True, mixed=True), ro)
