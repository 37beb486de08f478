"""GPU parity of the HIP scoring path against the oracle (FP64 CPU
restatement of the reference).  Bars, stated per check:
  * Bessel table, fit of identical profiles: bit-exact except where the
    device and host libm (pow, exp) round differently (counted, <= 2 ulp);
  * cross-term profiles / I(q): within 1e-10 relative of the oracle
    (the parity bar of the FP64 form);
  * score, c1, c2: within 1e-7 relative / 1e-6 absolute and identical
    score ranking of the top poses.
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


def _I(c):
    return c[..., 0, :] - c[..., 1, :] + c[..., 2, :] + c[..., 3, :] - c[..., 4, :] + c[..., 5, :]


def _cmp_results(pb, r_gpu, r_orc, I_tol=1e-10):
    ok = ~np.isnan(r_orc["score"])
    assert np.array_equal(ok, ~np.isnan(r_gpu["score"]))
    cg, co = r_gpu["cross"][ok], r_orc["cross"][ok]
    Ig, Io = _I(cg), _I(co)
    relI = np.abs(Ig - Io) / np.abs(Io)
    assert relI.max() <= I_tol, relI.max()
    scale = np.abs(co).max(axis=2, keepdims=True)
    assert (np.abs(cg - co) / scale).max() <= I_tol
    s_g, s_o = r_gpu["score"][ok], r_orc["score"][ok]
    assert np.all(np.abs(s_g - s_o) <= 1e-7 * np.abs(s_o))
    assert np.all(np.abs(r_gpu["c1"][ok] - r_orc["c1"][ok]) <= 1e-6)
    assert np.all(np.abs(r_gpu["c2"][ok] - r_orc["c2"][ok]) <= 1e-6)
    k = min(100, ok.sum())
    assert np.array_equal(np.argsort(s_g, kind="stable")[:k], np.argsort(s_o, kind="stable")[:k])
    return relI.max()


def _gpu_problem(F, pb, **kw):
    return F.ScoringProblem(pb["L"], pb["qvals"], pb["zvals"], pb["A_re"], pb["A_im"], pb["B_re"],
                            pb["B_im"], pb["a"], pb["params"].mult, pb["peak"],
                            nbeta=pb["nbeta"], nalpha=pb["na"], ngamma=pb["ng"], **kw)


def test_sbessel_table(F):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(0, 45, 3000), [0.0, 1e-9, 40.0]])
    g = F.sbessel_batch(x, 40)
    o = np.array([[O.sbessel(p, xi) for p in range(41)] for xi in x])
    ulp = np.abs(g - o) / np.spacing(np.abs(o) + 1e-300)
    print("sbessel exact:", np.mean(g == o), "max ulp", ulp.max())
    assert ulp.max() <= 2
    assert np.mean(g == o) > 0.99


def test_fit_identical_profiles(F):
    pb = build_problem(300, 200, 8, 0.3, 25, np.arange(15.0, 35.0), 400, seed=21)
    r = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                              pb["params"], pb["qvals"], pb["zvals"], pb["L"], want_cross=True)
    cts, want = [], []
    for ct in r["cross"]:
        sc, c1, c2, nev, cs = O.fit_one(ct, pb["qvals"], pb["params"])
        cts.append(cs)
        want.append((sc, c1, c2, nev))
    want = np.array(want)
    got = F.fit_batch(np.array(cts), pb["qvals"], pb["a"], pb["params"].mult)
    exact = np.all(got == want, axis=1)
    print("bit-exact fits:", exact.mean())
    assert exact.mean() >= 0.95
    assert np.all(np.abs(got[:, 0] - want[:, 0]) <= 1e-9 * want[:, 0])
    assert np.all(np.abs(got[:, 1:3] - want[:, 1:3]) <= 1e-6)


@pytest.mark.parametrize("L,Q,qmax,nat", [(6, 21, 0.3, (300, 200)), (10, 31, 0.3, (600, 400)),
                                          (15, 50, 0.5, (800, 500))])
def test_native_grid_vs_oracle(F, L, Q, qmax, nat):
    zv = np.arange(18.0, 30.0)
    pb = build_problem(nat[0], nat[1], L, qmax, Q, zv, 150, seed=L)
    r_o = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], zv, L, want_cross=True)
    gp = _gpu_problem(F, pb)
    r_g = F.score(gp, pb["index"], want_cross=True)
    rel = _cmp_results(pb, r_g, r_o)
    print(f"L={L}: max rel I diff {rel:.2e}; stats {r_g['stats']}")


def test_many_q_vs_oracle(F):
    """400 q points: the moments, the q grid and the reciprocals of the
    q steps no longer fit the fit kernel's LDS beside the L-BFGS-B blocks
    (at most 160 q do), so the fits read them from global memory; the
    reference has no limit on the number of q points."""
    L = 6
    zv = np.arange(20.0, 26.0)
    pb = build_problem(200, 150, L, 0.5, 400, zv, 40, seed=14)
    r_o = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], zv, L, want_cross=True)
    r_g = F.score(_gpu_problem(F, pb), pb["index"], want_cross=True)
    rel = _cmp_results(pb, r_g, r_o)
    print(f"400 q: max rel I diff {rel:.2e}")


def test_extended_grid_vs_oracle(F):
    """nbeta = 19 (10 deg), 64-point alpha/gamma grids: beyond the reference's
    grid; the oracle evaluates the same generalised transform directly."""
    L = 8
    zv = np.arange(25.0, 31.0)
    pb = build_problem(400, 300, L, 0.5, 21, zv, 60, seed=5, nbeta=19, na=64, ng=64)
    O.set_pi(O.PI_TRUNC)
    r_o = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], zv, L, nbeta=19, na=64, ng=64,
                                want_cross=True)
    r_g = F.score(_gpu_problem(F, pb), pb["index"], want_cross=True)
    _cmp_results(pb, r_g, r_o)


def _break_symmetry(pb, L, seed):
    """The problem with A_{l,-m} (m < 0) scaled by 1 + 1e-3 noise, so that
    the receptor coefficients are no longer conjugate symmetric."""
    K2 = (L + 1) ** 2
    rng = np.random.default_rng(seed)
    Ar = np.array(pb["A_re"], dtype=np.float64).reshape(-1, K2).copy()
    Ai = np.array(pb["A_im"], dtype=np.float64).reshape(-1, K2).copy()
    neg = [l * l + l + m for l in range(1, L + 1) for m in range(-l, 0)]
    Ar[:, neg] *= 1.0 + 1e-3 * rng.standard_normal((Ar.shape[0], len(neg)))
    Ai[:, neg] *= 1.0 + 1e-3 * rng.standard_normal((Ai.shape[0], len(neg)))
    return dict(pb, A_re=Ar.reshape(np.shape(pb["A_re"])), A_im=Ai.reshape(np.shape(pb["A_im"])))


def _break_m0(pb, L, seed):
    """The problem with only the receptor's m = 0 coefficients moved off the
    symmetry: A_l0 must be real for even l and imaginary for odd l; a part
    of 1e-3 of each coefficient's size is added to the other component."""
    K2 = (L + 1) ** 2
    rng = np.random.default_rng(seed)
    Ar = np.array(pb["A_re"], dtype=np.float64).reshape(-1, K2).copy()
    Ai = np.array(pb["A_im"], dtype=np.float64).reshape(-1, K2).copy()
    for l in range(L + 1):
        e = l * l + l
        mag = np.hypot(Ar[:, e], Ai[:, e])
        if l % 2 == 0:
            Ai[:, e] += 1e-3 * mag * rng.standard_normal(Ar.shape[0])
        else:
            Ar[:, e] += 1e-3 * mag * rng.standard_normal(Ar.shape[0])
    return dict(pb, A_re=Ar.reshape(np.shape(pb["A_re"])), A_im=Ai.reshape(np.shape(pb["A_im"])))


def test_large_L_chunked_vs_oracle(F):
    """L = 25, the largest L of the workloads.  The translation blocks (6201
    entries) are too large for LDS and are read from global memory.  Physical
    coefficients use the half entry set (351 per row, 338 pairing slots: two
    m-chunk launches of at most 256 entries for FP64 tables, three of at most
    128 slots for FP32 tables, accumulated in place; both forms at their
    bars); coefficients without the conjugate symmetry use all 676 entries
    (three chunks)."""
    L = 25
    zv = np.arange(30.0, 34.0)
    pb = build_problem(500, 300, L, 0.6, 11, zv, 24, seed=25)
    r_o = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], zv, L, want_cross=True)
    gp = _gpu_problem(F, pb)
    r_g = F.score(gp, pb["index"], want_cross=True)
    st = r_g["stats"]
    assert st["cross_entries"] in (351, 676)
    assert st["cross_chunks"] == -(-st["cross_entries"] // 256)
    rel = _cmp_results(pb, r_g, r_o)
    r_m = F.score(gp, pb["index"], want_cross=True, mixed=True)
    assert r_m["stats"]["cross_chunks"] == (3 if st["cross_entries"] == 351 else 6)
    ok = ~np.isnan(r_o["score"])
    relI = np.abs(_I(r_m["cross"][ok]) - _I(r_o["cross"][ok])) / np.abs(_I(r_o["cross"][ok]))
    relc = np.abs(r_m["score"][ok] - r_o["score"][ok]) / r_o["score"][ok]
    print(f"L=25 ({st['cross_entries']} entries): FP64 max rel I {rel:.2e}; "
          f"mixed max rel I {relI.max():.2e}, chi {relc.max():.2e}")
    assert relI.max() <= 1e-5 and relc.max() <= 1e-4   # chi bar: see test_gpu_mixed.py
    assert np.array_equal(np.argsort(r_m["score"][ok], kind="stable"),
                          np.argsort(r_o["score"][ok], kind="stable"))
    pb2 = _break_symmetry(pb, L, 7)
    r_o2 = O.compute_saxs_scores(pb2["index"], pb2["A_re"], pb2["A_im"], pb2["B_re"],
                                 pb2["B_im"], pb2["params"], pb2["qvals"], zv, L, want_cross=True)
    r_g2 = F.score(_gpu_problem(F, pb2), pb2["index"], want_cross=True)
    assert r_g2["stats"]["cross_entries"] == 676 and r_g2["stats"]["cross_chunks"] == 3
    _cmp_results(pb2, r_g2, r_o2)


def test_entry_sets_symmetric_and_asymmetric(F):
    """Coefficients of real densities satisfy A_{l,-m} = (-1)^(l+m)
    conj(A_lm), so the kernels form only m >= 0 and count the m > 0 terms
    twice; coefficients that break the symmetry take the full m = -L..L
    entry set.  Both against the oracle, which always sums over all m."""
    L = 8
    zv = np.arange(20.0, 26.0)
    pb = build_problem(300, 200, L, 0.3, 21, zv, 60, seed=12)
    r_o = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], zv, L, want_cross=True)
    r_g = F.score(_gpu_problem(F, pb), pb["index"], want_cross=True)
    assert r_g["stats"]["cross_entries"] == (L + 1) * (L + 2) // 2
    _cmp_results(pb, r_g, r_o)
    K2 = (L + 1) ** 2
    pb2 = _break_symmetry(pb, L, 3)
    r_o2 = O.compute_saxs_scores(pb2["index"], pb2["A_re"], pb2["A_im"], pb2["B_re"],
                                 pb2["B_im"], pb2["params"], pb2["qvals"], zv, L, want_cross=True)
    r_g2 = F.score(_gpu_problem(F, pb2), pb2["index"], want_cross=True)
    assert r_g2["stats"]["cross_entries"] == K2
    _cmp_results(pb2, r_g2, r_o2)
    assert not np.allclose(r_g2["cross"], r_g["cross"], rtol=1e-9, atol=0)
    # symmetric for m >= 1 but not at m = 0: the full set as well
    pb3 = _break_m0(pb, L, 5)
    r_o3 = O.compute_saxs_scores(pb3["index"], pb3["A_re"], pb3["A_im"], pb3["B_re"],
                                 pb3["B_im"], pb3["params"], pb3["qvals"], zv, L, want_cross=True)
    r_g3 = F.score(_gpu_problem(F, pb3), pb3["index"], want_cross=True)
    assert r_g3["stats"]["cross_entries"] == K2
    _cmp_results(pb3, r_g3, r_o3)
    assert not np.allclose(r_g3["cross"], r_g["cross"], rtol=1e-9, atol=0)


def test_out_of_range_poses_untouched(F):
    L = 6
    zv = np.arange(20.0, 26.0)
    pb = build_problem(200, 150, L, 0.3, 17, zv, 30, seed=8)
    N = 2 * L + 1
    size5d = (L + 1) ** 2 * N ** 3
    # past the last z slot, negative, and z quotients at 2^31 + 5 and 2^32 + 1
    # (a 32-bit z digit would wrap them to a negative key and to z = 1)
    bad = [len(zv) * size5d + 5, -3, (2 ** 31 + 5) * size5d + 11, (2 ** 32 + 1) * size5d + 7]
    idx = np.concatenate([pb["index"], bad]).astype(np.int64)
    nb = len(bad)
    r = F.score(_gpu_problem(F, pb), idx)
    assert np.all(np.isnan(r["score"][-nb:])) and np.all(np.isnan(r["c1"][-nb:]))
    assert np.all(np.isnan(r["c2"][-nb:]))
    assert not np.any(np.isnan(r["score"][:-nb]))
    assert r["stats"]["n_scored"] == len(pb["index"])


def test_duplicate_poses_identical(F):
    L = 6
    zv = np.arange(20.0, 26.0)
    pb = build_problem(200, 150, L, 0.3, 17, zv, 20, seed=9)
    idx = np.concatenate([pb["index"], pb["index"][:7], pb["index"][:3]])
    r = F.score(_gpu_problem(F, pb), idx)
    assert np.array_equal(r["score"][20:27], r["score"][:7])
    assert np.array_equal(r["c2"][27:], r["c2"][:3])


def test_dropin_sxs_compute_saxs_scores(F):
    """The reference-signature entry point on reference structs."""
    import ctypes as C
    L = 6
    zv = np.arange(20.0, 26.0)
    pb = build_problem(200, 150, L, 0.3, 17, zv, 25, seed=10)
    lib = F.lib()
    A = F.spf_from_arrays(pb["A_re"], pb["A_im"], L)
    B = F.spf_from_arrays(pb["B_re"], pb["B_im"], L)
    exp_q = np.ascontiguousarray(pb["exp_q"])
    prof = lib.sxs_profile_create(exp_q.ctypes.data_as(C.POINTER(C.c_double)), exp_q.size, 0)
    C.memmove(prof.contents.in_, pb["exp_in"].ctypes.data, 8 * exp_q.size)
    C.memmove(prof.contents.err, pb["exp_err"].ctypes.data, 8 * exp_q.size)
    q = np.ascontiguousarray(pb["qvals"])
    params = lib.sxs_opt_params_create(prof, q.ctypes.data_as(C.POINTER(C.c_double)), q.size,
                                       pb["rm"])
    a = np.ctypeslib.as_array(params.contents.a, shape=(6 * q.size,))
    assert np.array_equal(a, pb["a"])
    n = len(pb["index"])
    idx = np.ascontiguousarray(pb["index"], dtype=np.int32)
    sc, c1, c2 = np.zeros(n), np.zeros(n), np.zeros(n)
    z = np.ascontiguousarray(zv)
    lib.sxs_compute_saxs_scores(sc.ctypes.data, c1.ctypes.data, c2.ctypes.data, idx.ctypes.data,
                                n, A, B, params, q.ctypes.data, q.size, z.ctypes.data, z.size, L, 1)
    r = F.score(_gpu_problem(F, pb), pb["index"])
    assert np.array_equal(sc, r["score"]) and np.array_equal(c1, r["c1"])
    r_o = O.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], zv, L)
    assert np.all(np.abs(sc - r_o["score"]) <= 1e-7 * r_o["score"])
    lib.sxs_opt_params_free(params)
    lib.sxs_spf_full_free(A)
    lib.sxs_spf_full_free(B)


@pytest.mark.parametrize("L,natoms,Q", [(6, 700, 26), (15, 700, 26), (25, 2500, 41)])
def test_atoms2spf_vs_oracle(F, L, natoms, Q):
    """GPU atom expansion (pdb2spf.c:24-152: angles, Legendre values and
    Bessel table on the device, sums over atoms on FP64 MFMA in chunks of
    1024 atoms) against the oracle restatement; only the order of the sums
    over atoms differs.  2500 atoms: three chunks, the last one partial."""
    from libfmftsaxs_amd import synthetic as S
    rec = S.random_chain(natoms, 17)
    rec = rec.translated(-rec.center_of_extrema())
    rec.sa = S.shrake_rupley(rec)
    q = O.mkarray(0.0, 0.5, Q)
    rg, ig, rmg = F.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
    ro, io, rmo = O.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
    assert rmg == rmo
    scale = np.maximum(np.abs(ro).max(axis=2, keepdims=True), np.abs(io).max(axis=2, keepdims=True))
    assert np.all(np.abs(rg - ro) <= 1e-12 * scale)
    assert np.all(np.abs(ig - io) <= 1e-12 * scale)


def test_large_complex_grid_vs_direct(F):
    """The large single-complex grid (L = 25, 81 q on [0, 0.6], 37 betas in
    5 degree steps, 128-point alpha/gamma grids, z in 0.5 A steps over
    20..70 A) with small molecules: six poses of both forms against the
    oracle's independent per-pose evaluation (rotation, translation and
    pairing per pose; exact pi on both sides)."""
    from libfmftsaxs_amd import synthetic as S
    O.set_pi(O.PI_EXACT)
    rec, lig = S.make_dimer_inputs(600, 300, 31, 32, water=True)
    L, nb, ng = 25, 37, 128
    q = O.mkarray(0.0, 0.6, 81)
    zv = np.arange(20.0, 70.0 + 0.25, 0.5)
    Ar, Ai, rmA = O.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
    Br, Bi, rmB = O.atoms2spf(lig.coords, lig.vac, lig.dum, lig.sa, lig.vdw, q, L)
    rng = np.random.default_rng(44)
    idx = np.sort(rng.choice(zv.size * nb * nb * ng ** 3, 6, replace=False)).astype(np.int64)
    cross_o = []
    for i in idx:
        z, b1, b2, a2, g1, g2 = O.disassemble_index(int(i), nb, ng, ng)
        cross_o.append(O.direct_pose_cross_terms(
            Ar, Ai, Br, Bi, L, q, zv[z], b1 * np.pi / (nb - 1), 2 * np.pi * g1 / ng,
            2 * np.pi * a2 / ng, b2 * np.pi / (nb - 1), 2 * np.pi * g2 / ng))
    cross_o = np.array(cross_o)
    I0 = _I(cross_o[0])
    Iexp = I0 * (1 + 0.01 * rng.normal(size=q.size))
    a = O.scoring_helper(q, Iexp, 0.05 * Iexp, q)
    rm = (rmA * rec.natoms + rmB * lig.natoms) / (rec.natoms + lig.natoms)
    P = O.Params(a, rm, Iexp[0])
    pb = F.ScoringProblem(L, q, zv, Ar, Ai, Br, Bi, a, P.mult, Iexp[0], nbeta=nb, nalpha=ng,
                          ngamma=ng, pi=np.pi)
    want = [O.fit_one(c, q, P) for c in cross_o]
    for mixed, tol in ((False, 1e-10), (True, 1e-5)):
        print(f"large grid, L 25, 81 q, mixed={mixed}", flush=True)
        r = F.score(pb, idx, want_cross=True, mixed=mixed)
        relI = np.abs(_I(r["cross"]) - _I(cross_o)) / np.abs(_I(cross_o))
        print(f"large grid mixed={mixed}: max rel I {relI.max():.2e}")
        assert relI.max() <= tol
        if not mixed:
            for k, (sc, c1, c2, _, _) in enumerate(want):
                assert abs(r["score"][k] - sc) <= 1e-7 * sc
                assert abs(r["c1"][k] - c1) <= 1e-6 and abs(r["c2"][k] - c2) <= 1e-6


def test_batch_equals_single_calls(F):
    """fsx_score_batch over complexes of different L, grids and pose counts
    gives, for each complex, exactly the results of its own fsx_score call,
    in both forms."""
    pbs = [build_problem(200, 150, 6, 0.3, 11, np.arange(20.0, 25.0), 40, seed=31),
           build_problem(300, 200, 10, 0.4, 21, np.arange(22.0, 30.0), 90, seed=32,
                         nbeta=11, na=24, ng=20),
           build_problem(300, 200, 15, 0.5, 25, np.arange(25.0, 31.0), 60, seed=33)]
    gps = [_gpu_problem(F, pb) for pb in pbs]
    for mixed in (False, True):
        print(f"batch of {len(gps)}, mixed={mixed}", flush=True)
        rb = F.score_batch(gps, [pb["index"] for pb in pbs], want_cross=True, mixed=mixed)
        for gp, pb, r in zip(gps, pbs, rb):
            rs = F.score(gp, pb["index"], want_cross=True, mixed=mixed)
            for k in ("score", "c1", "c2", "cross"):
                assert np.array_equal(r[k], rs[k], equal_nan=True), k
            assert r["stats"]["n_scored"] == rs["stats"]["n_scored"]


def test_max_order_vs_direct(F):
    """The largest expansion order accepted, L = 40, on its native grid (41
    betas, 81-point alpha/gamma): two poses of both forms against the
    oracle's direct per-pose evaluation stored in tests/golden/l40_direct.npz
    (tests/golden/make_l40_direct.py; exact pi).  The half set has 861
    entries: FP64 tables run in chunks of at most 256, FP32 tables in chunks
    of at most 128 pairing slots, and at this order both read the translation
    blocks from global memory instead of LDS."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_l40_direct as G
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "l40_direct.npz"))
    q, zv, Ar, Ai, Br, Bi, rm = G.inputs()
    assert np.array_equal(q, d["q"]) and np.array_equal(zv, d["zvals"])
    idx, cross_o = d["index"], d["cross"]
    rng = np.random.default_rng(41)
    Iexp = _I(cross_o[0]) * (1 + 0.01 * rng.normal(size=q.size))
    a = O.scoring_helper(q, Iexp, 0.05 * Iexp, q)
    P = O.Params(a, rm, Iexp[0])
    pb = F.ScoringProblem(G.L, q, zv, Ar, Ai, Br, Bi, a, P.mult, Iexp[0], nbeta=G.NB,
                          nalpha=G.NG, ngamma=G.NG, pi=np.pi)
    for mixed, tol in ((False, 1e-10), (True, 1e-5)):
        print(f"L 40, mixed={mixed}", flush=True)
        r = F.score(pb, idx, want_cross=True, mixed=mixed)
        st = r["stats"]
        assert st["n_scored"] == idx.size and st["cross_entries"] == 41 * 42 // 2
        assert st["cross_chunks"] >= (7 if mixed else 4)
        relI = np.abs(_I(r["cross"]) - _I(cross_o)) / np.abs(_I(cross_o))
        print(f"L 40 mixed={mixed}: max rel I {relI.max():.2e}, chunks {st['cross_chunks']}")
        assert relI.max() <= tol
        assert np.all(np.isfinite(r["score"]))
