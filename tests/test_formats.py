"""On-disk formats of the path, through the product library (CPU only):

  * SPF coefficient file, sxs_spf_full_write / sxs_spf_full_read /
    sxs_spf_full_fread (src/pdb2spf.c:267-349): header "% i % i % .4f",
    then per q and (l, m = -l..l) one line of six "% .4e" values (V, D, W,
    real and imaginary parts);
  * profile text, sxs_profile_write / sxs_profile_read (src/profile.c:69-124):
    "%.4f %.4f %.4f" lines of q, I and sigma, sigma = I * rerr where the
    stored error is not positive.

The expected text is produced here with Python's own printf-style
formatting, which follows C's for these conversions; the read-back values
must be the nearest doubles of the printed decimals (what Python's float()
and C's strtod both return), and the reference's own profile files read by
the product must equal numpy.loadtxt of the same files bit for bit."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def F():
    import libfmftsaxs_amd as F
    return F


@pytest.fixture(scope="module")
def libc():
    c = C.CDLL(None)
    c.fopen.restype = C.c_void_p
    c.fopen.argtypes = [C.c_char_p, C.c_char_p]
    c.fclose.argtypes = [C.c_void_p]
    return c


def _spf_text(re, im, L, rm):
    Q = re.shape[1]
    lines = ["% i % i % .4f" % (L, Q, rm)]
    for q in range(Q):
        for l in range(L + 1):
            for m in range(-l, l + 1):
                i = l * (l + 1) + m
                lines.append("% .4e % .4e % .4e % .4e % .4e % .4e" % (
                    re[0, q, i], im[0, q, i], re[1, q, i], im[1, q, i], re[2, q, i], im[2, q, i]))
    return "\n".join(lines) + "\n"


def _rand_spf(L, Q, seed):
    rng = np.random.default_rng(seed)
    K2 = (L + 1) ** 2
    mag = 10.0 ** rng.uniform(-12, 6, size=(2, 3, Q, K2))
    sign = rng.choice([-1.0, 1.0], size=(2, 3, Q, K2))
    v = mag * sign
    v[:, :, 0, 0] = 0.0       # exact zeros print as " 0.0000e+00"
    return v[0], v[1]


def test_spf_write_exact_text_and_read_back(F, libc, tmp_path):
    L, Q, rm = 6, 5, 1.873456789
    re, im = _rand_spf(L, Q, 1)
    p = F.spf_from_arrays(re, im, L, rm)
    path = str(tmp_path / "rec.spf")
    F.lib().sxs_spf_full_write(path.encode(), p)
    F.lib().sxs_spf_full_free(p)
    text = open(path).read()
    assert text == _spf_text(re, im, L, rm)
    # read back: the nearest doubles of the printed values
    want_re = np.vectorize(lambda x: float("%.4e" % x))(re)
    want_im = np.vectorize(lambda x: float("%.4e" % x))(im)
    for reader in ("read", "fread"):
        if reader == "read":
            s = F.lib().sxs_spf_full_read(path.encode())
        else:
            F.lib().sxs_spf_full_fread.argtypes = [C.c_void_p]
            F.lib().sxs_spf_full_fread.restype = C.POINTER(F.SpfFull)
            fp = libc.fopen(path.encode(), b"r")
            s = F.lib().sxs_spf_full_fread(fp)
            libc.fclose(fp)
        assert s.contents.L == L and s.contents.qnum == Q
        assert s.contents.rm == float("%.4f" % rm)
        r2, i2, _ = F.spf_to_arrays(s)
        assert np.array_equal(r2, want_re) and np.array_equal(i2, want_im)
        F.lib().sxs_spf_full_free(s)


def test_spf_reference_amplitudes_round_trip(F, tmp_path):
    """The reference's receptor amplitudes (ref_spf, printed by the
    reference's own writer at "% .4e") are reproduced to the last printed
    digit by a write and read through the product."""
    spf = np.load(os.path.join(G, "ref_spf.npz"))
    L = int(spf["L"])
    p = F.spf_from_arrays(spf["re"], spf["im"], L, float(spf["rm"]))
    path = str(tmp_path / "ref.spf")
    F.lib().sxs_spf_full_write(path.encode(), p)
    F.lib().sxs_spf_full_free(p)
    s = F.lib().sxs_spf_full_read(path.encode())
    r2, i2, rm2 = F.spf_to_arrays(s)
    F.lib().sxs_spf_full_free(s)
    assert np.array_equal(r2, spf["re"]) and np.array_equal(i2, spf["im"])
    assert rm2 == float(spf["rm"])
    lines = open(path).read().splitlines()
    assert len(lines) == 1 + int(spf["qnum"]) * (L + 1) ** 2
    assert lines[0] == "% i % i % .4f" % (L, int(spf["qnum"]), float(spf["rm"]))


def test_profile_write_exact_text_and_read_back(F, tmp_path):
    lib = F.lib()
    q = F.mkarray(0.0, 0.5, 11)
    p = lib.sxs_profile_create(q.ctypes.data_as(C.POINTER(C.c_double)), q.size, 0)
    inn = np.linspace(1e6, 2.5, q.size)
    err = 0.03 * inn
    err[[2, 7]] = 0.0          # not positive: written as I * rerr
    err[5] = -1.0
    C.memmove(p.contents.in_, inn.ctypes.data, 8 * q.size)
    C.memmove(p.contents.err, err.ctypes.data, 8 * q.size)
    rerr = p.contents.rerr
    assert rerr == 0.05
    path = str(tmp_path / "prof.txt")
    lib.sxs_profile_write(path.encode(), p)
    lib.sxs_profile_free(p)
    e_out = np.where(err > 0.0, err, inn * rerr)
    want = "".join("%.4f %.4f %.4f\n" % (q[i], inn[i], e_out[i]) for i in range(q.size))
    assert open(path).read() == want
    r = lib.sxs_profile_read(path.encode())
    n = r.contents.qnum
    assert n == q.size
    got = [np.ctypeslib.as_array(getattr(r.contents, k), shape=(n,)).copy()
           for k in ("qvals", "in_", "err")]
    for g, w in zip(got, (q, inn, e_out)):
        assert np.array_equal(g, np.array([float("%.4f" % x) for x in w]))
    # a profile read from a file owns its q grid (the caller frees it,
    # correlate.c:131); destroy leaves it alone
    qv = C.cast(r.contents.qvals, C.c_void_p)
    lib.sxs_profile_free(r)
    C.CDLL(None).free.argtypes = [C.c_void_p]
    C.CDLL(None).free(qv)


@pytest.mark.parametrize("name", ["ref_saxs.txt", "ref_profile.txt", "ref_fitted_profile.txt"])
def test_reference_profiles_read(F, name):
    """The reference's own profile files (ref_saxs: 1,714 points in
    exponent notation; ref_profile, ref_fitted_profile: "%.4f" lines)."""
    path = os.path.join(G, name)
    r = F.lib().sxs_profile_read(path.encode())
    n = r.contents.qnum
    got = np.stack([np.ctypeslib.as_array(getattr(r.contents, k), shape=(n,)).copy()
                    for k in ("qvals", "in_", "err")], axis=1)
    F.lib().sxs_profile_free(r)
    want = np.loadtxt(path)
    assert got.shape == want.shape and np.array_equal(got, want)


def test_malformed_files_exit_like_the_reference(tmp_path):
    """Bad formats are fatal in the reference (ERROR_MSG: message, then
    exit(EXIT_FAILURE), common.h:26-36); the product behaves the same."""
    bad_prof = tmp_path / "bad.txt"
    bad_prof.write_text("0.1 2.0 0.1\n0.2 x 0.1\n")
    bad_spf = tmp_path / "bad.spf"
    bad_spf.write_text(" 2  1  1.5000\n 1.0 2.0 3.0\n")
    for fn, path in (("sxs_profile_read", bad_prof), ("sxs_spf_full_read", bad_spf)):
        code = ("import sys; sys.path.insert(0, %r); import libfmftsaxs_amd as F; "
                "F.lib().%s(%r.encode()); print('returned')" % (ROOT, fn, str(path)))
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 1, (fn, r.returncode, r.stdout, r.stderr)
        assert "returned" not in r.stdout
        assert "Wrong input file format" in r.stderr
