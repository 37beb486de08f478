"""Pins the parts of the oracle that no reference build covers (the
correlation driver of fftsaxs.c, the chi^2 objective of min_saxs.c, the atom
expansion of pdb2spf.c) against mathematics that fixes them: closed forms,
an independent per-pose evaluation, brute-force chi^2, finite differences
and the Debye sum.  Each check is chosen so that a dropped term, a wrong sign
or index, or a transposed operand fails it."""
import numpy as np
import pytest

from tests.problem import build_problem


def test_wigner_d_closed_forms(orc):
    L = 12
    N = 2 * L + 1
    d0 = orc.wigner_d(L, 0.0).reshape(L + 1, N, N)
    for l in range(L + 1):
        sub = d0[l, L - l:L + l + 1, L - l:L + l + 1]
        assert np.allclose(sub, np.eye(2 * l + 1), atol=1e-13)
    dpi = orc.wigner_d(L, np.pi).reshape(L + 1, N, N)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            for m1 in range(-l, l + 1):
                want = (-1) ** (l - m1) if m == -m1 else 0.0
                assert abs(dpi[l, m + L, m1 + L] - want) < 1e-10
    db = orc.wigner_d(L, 0.83).reshape(L + 1, N, N)
    for l in range(L + 1):
        sub = db[l, L - l:L + l + 1, L - l:L + l + 1]
        assert np.allclose(sub @ sub.T, np.eye(2 * l + 1), atol=1e-12)
    # d^1 against the textbook matrix (rows m, columns m')
    b = 0.83
    c, s = np.cos(b), np.sin(b)
    d1 = np.array([[(1 + c) / 2, s / np.sqrt(2), (1 - c) / 2],
                   [-s / np.sqrt(2), c, s / np.sqrt(2)],
                   [(1 - c) / 2, -s / np.sqrt(2), (1 + c) / 2]])
    assert np.allclose(db[1, L - 1:L + 2, L - 1:L + 2], d1, atol=1e-15)


def test_sbessel_against_closed_forms(orc):
    for x in [0.05, 0.5, 1.3, 4.0, 9.0]:
        j0 = np.sin(x) / x
        j1 = np.sin(x) / x ** 2 - np.cos(x) / x
        j2 = (3 / x ** 2 - 1) * np.sin(x) / x - 3 * np.cos(x) / x ** 2
        # the reference's series stops at a 1e-5 relative term (sfbessel.c:46)
        assert abs(orc.sbessel(0, x) - j0) <= 2e-5 * max(abs(j0), 1e-3)
        assert abs(orc.sbessel(1, x) - j1) <= 2e-5 * max(abs(j1), 1e-3)
        assert abs(orc.sbessel(2, x) - j2) <= 2e-5 * max(abs(j2), 1e-3)


@pytest.fixture(scope="module")
def prob():
    import oracle as O
    O.set_pi(O.PI_TRUNC)
    return build_problem(250, 180, 6, 0.3, 17, np.array([0.0, 12.0, 21.0, 30.0]), 60, seed=3)


def test_zero_pose_equals_summed_spf(orc, prob):
    """z = 0 and all angles 0: the translation operator is the identity and
    no rotation applies, so the six cross terms are those of the summed
    expansion A + B (min_saxs.c:450-505)."""
    pb = prob
    L = pb["L"]
    idx = np.array([0], dtype=np.int64)    # z slot 0 holds z = 0
    r = orc.compute_saxs_scores(idx, pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], pb["zvals"], L, want_cross=True)
    want = orc.spf2cross_terms(pb["A_re"] + pb["B_re"], pb["A_im"] + pb["B_im"], L,
                               pb["qvals"].size)
    got = r["cross"][0]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_slab_dft_equals_direct_pose(orc, prob):
    """The reference-structured restatement (first sums, second sums, DFT to
    the Euler grid) against an independent per-pose rotation/translation
    evaluation, with exact pi and exact twiddles on both sides."""
    pb = prob
    orc.set_pi(orc.PI_EXACT)
    L = pb["L"]
    N = 2 * L + 1
    nb = L + 1
    idx = pb["index"][:25]
    r = orc.compute_saxs_scores(idx, pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                pb["params"], pb["qvals"], pb["zvals"], L, want_cross=True,
                                dense_threshold=1)
    for k, i in enumerate(idx):
        z, b1, b2, a2, g1, g2 = orc.disassemble_index(i, nb, N, N)
        d = orc.direct_pose_cross_terms(pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"], L,
                                        pb["qvals"], pb["zvals"][z], b1 * np.pi / L,
                                        2 * np.pi * g1 / N, 2 * np.pi * a2 / N,
                                        b2 * np.pi / L, 2 * np.pi * g2 / N)
        scale = np.abs(d).max(axis=1, keepdims=True)
        assert np.all(np.abs(r["cross"][k] - d) <= 1e-12 * scale)


def test_truncated_twiddles_close_to_exact(orc, prob):
    """fftsaxs.c:899: slabs with < 30 poses use truncated-pi twiddles; the
    effect on I(q) is ~1e-12 relative and both paths are exercised."""
    pb = prob
    L = pb["L"]
    kw = dict(want_cross=True)
    rt = orc.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                 pb["params"], pb["qvals"], pb["zvals"], L,
                                 dense_threshold=10 ** 9, **kw)
    re = orc.compute_saxs_scores(pb["index"], pb["A_re"], pb["A_im"], pb["B_re"], pb["B_im"],
                                 pb["params"], pb["qvals"], pb["zvals"], L,
                                 dense_threshold=1, **kw)
    assert np.all(rt["twiddle_exact"] == 0) and np.all(re["twiddle_exact"] == 1)
    I = lambda c: c[:, 0] - c[:, 1] + c[:, 2] + c[:, 3] - c[:, 4] + c[:, 5]
    It, Ie = I(rt["cross"]), I(re["cross"])
    rel = np.abs(It - Ie) / np.abs(Ie)
    assert rel.max() < 1e-10 and rel.max() > 0


def _brute_chi2(orc, ct, q, P, c1, c2, k, eq, ein, eerr):
    inn = orc.compile_intensity(ct, q, P.s.rm, c1, c2)
    model = np.interp(eq, q, inn)
    return np.mean((ein - k * model) ** 2 / eerr ** 2)


def test_chi2_expansion_identity_and_gradient(orc, prob):
    pb = prob
    P = pb["params"]
    q = pb["qvals"]
    r = orc.compute_saxs_scores(pb["index"][:6], pb["A_re"], pb["A_im"], pb["B_re"],
                                pb["B_im"], P, q, pb["zvals"], pb["L"], want_cross=True)
    for ct in r["cross"]:
        for c1, c2 in [(1.0, 0.0), (0.97, 1.3), (1.03, -1.1)]:
            k = orc.best_scale(ct, q, P, c1, c2)
            f, g = orc.gradient(ct, q, P, k, [c1, c2])
            fb = _brute_chi2(orc, ct, q, P, c1, c2, k, pb["exp_q"], pb["exp_in"], pb["exp_err"])
            assert abs(f - fb) <= 1e-9 * abs(fb)
            # k minimises chi^2 over the scale
            for kk in (k * 0.999, k * 1.001):
                assert orc.gradient(ct, q, P, kk, [c1, c2])[0] > f
            # gradient by central differences at fixed k
            h = 1e-6
            fp = orc.gradient(ct, q, P, k, [c1 + h, c2])[0]
            fm = orc.gradient(ct, q, P, k, [c1 - h, c2])[0]
            assert abs((fp - fm) / (2 * h) - g[0]) <= 1e-5 * (abs(g[0]) + abs(f))
            fp = orc.gradient(ct, q, P, k, [c1, c2 + h])[0]
            fm = orc.gradient(ct, q, P, k, [c1, c2 - h])[0]
            assert abs((fp - fm) / (2 * h) - g[1]) <= 1e-5 * (abs(g[1]) + abs(f))


def test_fit_satisfies_bound_kkt(orc, prob):
    """At the returned (c1, c2) the projected gradient is small: interior
    coordinates have |g| small, coordinates at a bound point outward."""
    pb = prob
    P = pb["params"]
    q = pb["qvals"]
    r = orc.compute_saxs_scores(pb["index"][:20], pb["A_re"], pb["A_im"], pb["B_re"],
                                pb["B_im"], P, q, pb["zvals"], pb["L"], want_cross=True)
    for i, ct in enumerate(r["cross"]):
        sc, c1, c2, nev, cts = orc.fit_one(ct, q, P)
        assert sc == r["score"][i] and c1 == r["c1"][i] and c2 == r["c2"][i]
        k = orc.best_scale(cts, q, P, c1, c2)
        f, g = orc.gradient(cts, q, P, k, [c1, c2])
        assert abs(np.sqrt(f) - sc) <= 1e-12 * sc
        for x, gx, lo, hi in [(c1, g[0], 0.96, 1.04), (c2, g[1], -2.0, 4.0)]:
            if x <= lo:
                assert gx >= -1e-3 * (1 + f)
            elif x >= hi:
                assert gx <= 1e-3 * (1 + f)
            else:
                assert abs(gx) <= 2e-2 * (1 + f)


def test_atom_expansion_debye_limit(orc):
    """sum_lm |A^V_lm(q)|^2 -> Debye sum_ij f_i f_j sinc(q r_ij) for a small
    cluster as L grows (pdb2spf.c:24-152; 4 pi i^l factor and Y_lm norm)."""
    rng = np.random.default_rng(2)
    X = rng.normal(scale=3.0, size=(12, 3))
    X[:, 0] += 0.01
    f = rng.uniform(2, 8, 12)
    q = orc.mkarray(0.0, 0.4, 9)
    re, im, rm = orc.atoms2spf(X, f, np.zeros(12), None, np.full(12, 1.7), q, 18)
    I = (re[0] ** 2 + im[0] ** 2).sum(axis=1)
    d = np.linalg.norm(X[:, None] - X[None], axis=-1)
    for iq, qq in enumerate(q):
        x = qq * d
        s = np.where(x > 0, np.sin(x) / np.where(x > 0, x, 1), 1.0)
        deb = (f[:, None] * f[None] * s).sum() * (4 * np.pi) ** 2 / (4 * np.pi)
        assert abs(I[iq] - deb) <= 3e-4 * deb
    assert abs(rm - 1.7) < 1e-15


def test_translation_along_axis_matches_moved_atoms(orc):
    """A ligand translated by +z along the axis (pose z, all angles 0): the
    cross terms from the translation operator agree with the expansion of
    the physically moved dimer at low q, where the L = 15 expansion has
    converged (limited by the reference's 1e-5 Bessel series)."""
    from libfmftsaxs_amd import synthetic as S
    rec = S.random_chain(60, 11)
    lig = S.random_chain(40, 12)
    rec = rec.translated(-rec.center_of_extrema())
    lig = lig.translated(-lig.centroid())
    L = 15
    q = orc.mkarray(0.0, 0.1, 6)
    z = 14.0
    Ar, Ai, _ = orc.atoms2spf(rec.coords, rec.vac, rec.dum, None, rec.vdw, q, L)
    Br, Bi, _ = orc.atoms2spf(lig.coords, lig.vac, lig.dum, None, lig.vdw, q, L)
    d = orc.direct_pose_cross_terms(Ar, Ai, Br, Bi, L, q, z, 0, 0, 0, 0, 0)
    X = np.vstack([rec.coords, lig.coords + [0, 0, z]])
    vac = np.concatenate([rec.vac, lig.vac])
    dum = np.concatenate([rec.dum, lig.dum])
    Dr, Di, _ = orc.atoms2spf(X, vac, dum, None, np.ones(len(vac)), q, L)
    want = orc.spf2cross_terms(Dr, Di, L, q.size)
    for k in (0, 1, 3):   # VV, VD, DD (W is zero here)
        assert np.all(np.abs(d[k] - want[k]) <= 2e-4 * np.abs(want[k]) + 1e-9)


def test_index_round_trip_and_carry(orc):
    L = 15
    N = 2 * L + 1
    for t in [(0, 0, 0, 0, 0, 0), (3, 7, 15, 30, 12, 1), (79, 15, 15, 30, 30, 30)]:
        i = orc.assemble_index(*t, L + 1, N, N)
        assert orc.disassemble_index(i, L + 1, N, N) == t
    # a2 = 2 pi - 0.0005 rounds to digit N, which carries into b2 (tools/correlate.c:222-238)
    zv = np.arange(1.0, 81.0)
    pi = orc.PI_TRUNC
    (idx,) = orc.euler2index(40.0, 0.0, 0.0, 0.0005, 0.0, 0.0, zv, L)
    z, b1, b2, a2, g1, g2 = orc.disassemble_index(idx, L + 1, N, N)
    assert (z, b1, a2) == (39, 0, 0) and b2 == 1
    # a pose whose z is not within 0.001 of a grid z is dropped (tools/correlate.c:219-220)
    assert orc.euler2index(40.002, 1.0, 1.0, 1.0, 1.0, 1.0, zv, L) == []
    assert pi == 3.14159265358


def _rot(a, b, g):
    """active zyz rotation R(alpha, beta, gamma) in closed form (saxs_utils.c:65-79)"""
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    return np.array([[cg * cb * ca - sg * sa, -sg * cb * ca - cg * sa, sb * ca],
                     [cg * cb * sa + sg * ca, -sg * cb * sa + cg * ca, sb * sa],
                     [-cg * sb, sg * sb, cb]])


def _real_space_cross(orc, rec, lig, q, L, z, b1, g1, a2, b2, g2):
    """tools/score_ft_naive.c:156-190: receptor R(0, b1, g1) x, ligand
    R(a2, b2, g2) x + (0, 0, z), the complex centred at its centroid and
    expanded directly, then its cross terms (min_saxs.c:450-505)."""
    xr = rec.coords @ _rot(0.0, b1, g1).T
    xl = lig.coords @ _rot(a2, b2, g2).T + np.array([0.0, 0.0, z])
    x = np.vstack([xr, xl])
    x -= x.mean(axis=0)
    cat = lambda f: np.concatenate([getattr(rec, f), getattr(lig, f)])
    re, im, _ = orc.atoms2spf(x, cat("vac"), cat("dum"), cat("sa"), cat("vdw"), q, L)
    return orc.spf2cross_terms(re, im, L, q.size)


def test_fft_pose_equals_real_space_complex(orc):
    """The correlation's pose geometry against an independent real-space
    construction of the same complex.  A grid pose (b1, g1, a2, b2, g2) is
    the real-space pose with alpha2 and gamma2 mirrored to 2 pi - angle (the
    mirroring of the Euler-list reader, correlate.c:222-223); the cross terms
    agree to the truncation error at L, which falls with L down to the
    accuracy of the reference's spherical Bessel routine (~6e-7 relative
    against scipy, so ~3e-8 in I(q) here), and the unmirrored pose does not
    agree."""
    from libfmftsaxs_amd import synthetic as S
    orc.set_pi(orc.PI_EXACT)
    rec, lig = S.make_dimer_inputs(60, 40, 11, 12, water=True)
    q = orc.mkarray(0.0, 0.15, 6)
    z = 25.0
    I = lambda c: c[0] - c[1] + c[2] + c[3] - c[4] + c[5]
    err = {}
    for L in (6, 30):
        N = 2 * L + 1
        Ar, Ai, _ = orc.atoms2spf(rec.coords, rec.vac, rec.dum, rec.sa, rec.vdw, q, L)
        Br, Bi, _ = orc.atoms2spf(lig.coords, lig.vac, lig.dum, lig.sa, lig.vdw, q, L)
        P = orc.Params(np.ones(6 * q.size), 1.6, 1.0)
        for (ib1, ig1, ia2, ib2, ig2) in [(L // 4, 11, 23, 2 * L // 3, 41 % N), (3, N - 2, 5, L - 1, 9)]:
            idx = np.array([orc.assemble_index(0, ib1, ib2, ia2, ig1, ig2, L + 1, N, N)],
                           dtype=np.int64)
            cf = orc.compute_saxs_scores(idx, Ar, Ai, Br, Bi, P, q, np.array([z]), L,
                                         want_cross=True)["cross"][0]
            b1, b2 = ib1 * np.pi / L, ib2 * np.pi / L
            g1, a2, g2 = 2 * np.pi * ig1 / N, 2 * np.pi * ia2 / N, 2 * np.pi * ig2 / N
            cm = _real_space_cross(orc, rec, lig, q, L, z, b1, g1, 2 * np.pi - a2, b2,
                                   2 * np.pi - g2)
            rel = np.abs(I(cm) - I(cf)).max() / np.abs(I(cf)).max()
            err[L] = max(err.get(L, 0.0), rel)
            if L == 30:
                scale = np.abs(cf).max(axis=1, keepdims=True)
                assert np.all(np.abs(cm - cf) <= 1e-6 * scale)
                cu = _real_space_cross(orc, rec, lig, q, L, z, b1, g1, a2, b2, g2)
                assert np.abs(I(cu) - I(cf)).max() / np.abs(I(cf)).max() > 1e-3
    print("I(q) rel. error by L:", err)
    assert err[30] < 1e-7 and err[6] > 100 * err[30], err
