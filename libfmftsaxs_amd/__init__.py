"""Python binding (ctypes) of libfmftsaxs_amd, the MI355X implementation of
libfmftsaxs's FFT dimer SAXS scoring path.

The library is a C ABI (include/*.h); this module mirrors it for Python
callers and tests.  Importing it loads the in-tree libfmftsaxs_amd.so and
fails loudly if it has not been built (python -m libfmftsaxs_amd.build).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfmftsaxs_amd.so")
PI_REFERENCE = 3.14159265358

FSX_DEVICE_POSES = 1
FSX_DEVICE_OUTPUT = 2
FSX_MIXED_FP32 = 8
FSX_FIT_COMPACT = 16


class FsxError(RuntimeError):
    pass


class Problem(C.Structure):
    _fields_ = [("L", C.c_int), ("nbeta", C.c_int), ("nalpha", C.c_int), ("ngamma", C.c_int),
                ("qnum", C.c_int), ("qvals", C.POINTER(C.c_double)),
                ("znum", C.c_int), ("zvals", C.POINTER(C.c_double)),
                ("A_re", C.POINTER(C.c_double)), ("A_im", C.POINTER(C.c_double)),
                ("B_re", C.POINTER(C.c_double)), ("B_im", C.POINTER(C.c_double)),
                ("a", C.POINTER(C.c_double)), ("mult", C.c_double), ("peak", C.c_double),
                ("c1_lo", C.c_double), ("c1_hi", C.c_double), ("c2_lo", C.c_double),
                ("c2_hi", C.c_double), ("pi", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("n_scored", C.c_int), ("n_capped", C.c_int), ("fit_evals", C.c_long),
                ("ms_tables", C.c_double), ("ms_bessel_t", C.c_double),
                ("ms_rotate", C.c_double), ("ms_group", C.c_double),
                ("ms_cross", C.c_double), ("ms_fit", C.c_double), ("ms_total", C.c_double),
                ("cross_bytes", C.c_double), ("cross_flops", C.c_double),
                ("cross_items", C.c_long), ("cross_chunks", C.c_int), ("cross_entries", C.c_int),
                ("ms_fit_prep", C.c_double), ("fit_steps", C.c_long),
                ("fit_wide_steps", C.c_long), ("fit_max_steps", C.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpfSing(C.Structure):
    _fields_ = [("L", C.c_int), ("re", C.POINTER(C.c_double)), ("im", C.POINTER(C.c_double))]


class SpfFull(C.Structure):
    _fields_ = [("L", C.c_int), ("qnum", C.c_int), ("rm", C.c_double),
                ("V", C.POINTER(C.POINTER(SpfSing))), ("D", C.POINTER(C.POINTER(SpfSing))),
                ("W", C.POINTER(C.POINTER(SpfSing)))]


class Profile(C.Structure):
    _fields_ = [("qnum", C.c_int), ("rerr", C.c_double)] + [
        (n, C.POINTER(C.c_double)) for n in ("in_", "err", "qvals", "VV", "VD", "VW", "DD",
                                               "DW", "WW")] + [
        (n, C.c_double) for n in ("score", "scale", "c1", "c2", "b1", "g1", "a2", "b2", "g2")]


class OptParams(C.Structure):
    _fields_ = [("rm", C.c_double), ("mult", C.c_double), ("peak", C.c_double),
                ("a", C.POINTER(C.c_double)), ("wa", C.c_double * 43251),
                ("iwa", C.c_long * 3072), ("dsave", C.c_double * 29),
                ("isave", C.c_long * 44), ("lsave", C.c_long * 4)]


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FsxError("libfmftsaxs_amd.so is not built: run python -m libfmftsaxs_amd.build")
        L = C.CDLL(LIB_PATH)
        L.fsx_score.argtypes = [C.POINTER(Problem), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_uint, C.POINTER(Stats)]
        L.fsx_score.restype = C.c_int
        _vpp = C.POINTER(C.c_void_p)
        L.fsx_score_batch.argtypes = [C.c_int, C.POINTER(Problem), _vpp, C.POINTER(C.c_int), _vpp,
                                      _vpp, _vpp, _vpp, C.c_uint, C.POINTER(Stats)]
        L.fsx_score_batch.restype = C.c_int
        L.fsx_fit_host.argtypes = [_dp, _dp, C.c_int, _dp] + [C.c_double] * 5 + [_dp]
        L.fsx_fit_host.restype = C.c_int
        L.fsx_fit_host_form.argtypes = [_dp, _dp, C.c_int, _dp] + [C.c_double] * 5 + [C.c_uint,
                                                                                       _dp]
        L.fsx_fit_host_form.restype = C.c_int
        L.fsx_fit_batch.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp] + [C.c_double] * 5 + [_dp]
        L.fsx_fit_batch.restype = C.c_int
        L.fsx_sbessel_batch.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.fsx_sbessel_batch.restype = C.c_int
        L.fsx_atoms2spf.argtypes = [C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_double, _dp, _dp,
                                    C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_double)]
        L.fsx_atoms2spf.restype = C.c_int
        L.fsx_atoms2spf_form.argtypes = [C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_double, _dp,
                                         _dp, C.c_int, C.c_int, C.c_uint, _dp, _dp,
                                         C.POINTER(C.c_double)]
        L.fsx_atoms2spf_form.restype = C.c_int
        L.fsx_score_naive.argtypes = ([C.c_int, C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_double,
                                       _dp, _dp, C.c_int, C.c_int, _dp] + [C.c_double] * 5 +
                                      [_dp, C.c_int, _dp, C.c_void_p])
        L.fsx_score_naive.restype = C.c_int
        L.fsx_device_count.restype = C.c_int
        L.fsx_error_string.argtypes = [C.c_int]
        L.fsx_error_string.restype = C.c_char_p
        L.sxs_sbessel.argtypes = [C.c_int, C.c_double]
        L.sxs_sbessel.restype = C.c_double
        L.sxs_spf_full_create.argtypes = [C.c_int, C.c_int]
        L.sxs_spf_full_create.restype = C.POINTER(SpfFull)
        L.sxs_spf_full_free.argtypes = [C.POINTER(SpfFull)]
        L.sxs_spf_full_read.argtypes = [C.c_char_p]
        L.sxs_spf_full_read.restype = C.POINTER(SpfFull)
        L.sxs_spf_full_write.argtypes = [C.c_char_p, C.POINTER(SpfFull)]
        L.sxs_profile_create.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int]
        L.sxs_profile_create.restype = C.POINTER(Profile)
        L.sxs_profile_free.argtypes = [C.POINTER(Profile)]
        L.sxs_profile_read.argtypes = [C.c_char_p]
        L.sxs_profile_read.restype = C.POINTER(Profile)
        L.sxs_profile_write.argtypes = [C.c_char_p, C.POINTER(Profile)]
        L.sxs_profile_from_spf.argtypes = [C.POINTER(Profile), C.POINTER(SpfFull), C.c_double,
                                           C.c_double]
        L.sxs_opt_params_create.argtypes = [C.POINTER(Profile), C.POINTER(C.c_double), C.c_int,
                                            C.c_double]
        L.sxs_opt_params_create.restype = C.POINTER(OptParams)
        L.sxs_opt_params_free.argtypes = [C.POINTER(OptParams)]
        L.sxs_lbfgs_fitting.argtypes = [C.POINTER(Profile), C.POINTER(OptParams)]
        L.sxs_spf2fitted_profile.argtypes = [C.POINTER(Profile), C.POINTER(SpfFull),
                                             C.POINTER(OptParams)]
        L.sxs_best_scale.argtypes = [C.POINTER(Profile), C.POINTER(OptParams), C.c_double,
                                     C.c_double]
        L.sxs_best_scale.restype = C.c_double
        L.sxs_compute_saxs_scores.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_int, C.POINTER(SpfFull), C.POINTER(SpfFull),
                                              C.POINTER(OptParams), C.c_void_p, C.c_int,
                                              C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.sxs_assemble_index.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.sxs_assemble_index.restype = C.c_size_t
        L.sxs_disassemble_index.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int]
        L.sxs_mkarray.argtypes = [C.c_double, C.c_double, C.c_int]
        L.sxs_mkarray.restype = C.POINTER(C.c_double)
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _chk(rc):
    if rc != 0:
        raise FsxError(lib().fsx_error_string(rc).decode())


def device_count():
    return lib().fsx_device_count()


def release():
    """fsx_release: drop the device buffers and host tables; the next call is
    a cold call (tables built and uploaded again)."""
    lib().fsx_release()


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class ScoringProblem:
    """Flat-array scoring problem (fsx.h fsx_problem).  Arrays are kept alive
    by this object."""

    def __init__(self, L, qvals, zvals, A_re, A_im, B_re, B_im, a, mult, peak, nbeta=None,
                 nalpha=None, ngamma=None, c1=(0.96, 1.04), c2=(-2.0, 4.0), pi=PI_REFERENCE):
        self.keep = [f64(qvals), f64(zvals), f64(A_re).reshape(-1), f64(A_im).reshape(-1),
                     f64(B_re).reshape(-1), f64(B_im).reshape(-1), f64(a)]
        q, z, ar, ai, br, bi, aa = self.keep
        K2 = (L + 1) ** 2
        assert ar.size == 3 * q.size * K2 and br.size == ar.size and aa.size == 6 * q.size
        self.s = Problem(L, nbeta or L + 1, nalpha or 2 * L + 1, ngamma or 2 * L + 1, q.size,
                         _ptr(q), z.size, _ptr(z), _ptr(ar), _ptr(ai), _ptr(br), _ptr(bi),
                         _ptr(aa), mult, peak, c1[0], c1[1], c2[0], c2[1], pi)

    @property
    def qnum(self):
        return self.s.qnum


def score(problem, poses, want_cross=False, out=None, mixed=False, fit_compact=False):
    """Score poses (int64 mixed-radix indices).  Returns dict with score, c1,
    c2 (NaN where the z slot is out of range), optional cross [n,6,Q], stats.
    mixed=True selects the mixed-precision form (FP32 contraction, lane-group
    fit); fit_compact=True gives that form the compact-form L-BFGS-B fit."""
    poses = np.ascontiguousarray(poses, dtype=np.int64)
    n = poses.size
    sc = np.full(n, np.nan)
    c1 = np.full(n, np.nan)
    c2 = np.full(n, np.nan)
    cross = np.full((n, 6, problem.qnum), np.nan) if want_cross else None
    st = Stats()
    _chk(lib().fsx_score(C.byref(problem.s), poses.ctypes.data, n, sc.ctypes.data,
                         c1.ctypes.data, c2.ctypes.data,
                         None if cross is None else cross.ctypes.data,
                         (FSX_MIXED_FP32 if mixed else 0) | (FSX_FIT_COMPACT if fit_compact else 0),
                         C.byref(st)))
    return {"score": sc, "c1": c1, "c2": c2, "cross": cross, "stats": st.as_dict()}


def _batch_call(problems, poses, ns, sc, c1, c2, cross, flags):
    k = len(problems)
    arr = (Problem * k)(*[p.s for p in problems])
    vp = lambda xs: (C.c_void_p * k)(*xs)
    st = (Stats * k)()
    _chk(lib().fsx_score_batch(k, arr, vp(poses), (C.c_int * k)(*ns), vp(sc), vp(c1), vp(c2),
                               None if cross is None else vp(cross), flags, st))
    return [s_.as_dict() for s_ in st]


def score_batch(problems, poses_list, want_cross=False, mixed=False):
    """Batched scoring (fsx_score_batch): one call for several complexes, each
    with its own problem and int64 pose indices.  Returns one result dict per
    complex, as score()."""
    poses_list = [np.ascontiguousarray(p, dtype=np.int64) for p in poses_list]
    out = []
    for pb, p in zip(problems, poses_list):
        r = {"score": np.full(p.size, np.nan), "c1": np.full(p.size, np.nan),
             "c2": np.full(p.size, np.nan)}
        if want_cross:
            r["cross"] = np.full((p.size, 6, pb.qnum), np.nan)
        out.append(r)
    stats = _batch_call(problems, [p.ctypes.data for p in poses_list],
                        [p.size for p in poses_list], [r["score"].ctypes.data for r in out],
                        [r["c1"].ctypes.data for r in out], [r["c2"].ctypes.data for r in out],
                        [r["cross"].ctypes.data for r in out] if want_cross else None,
                        FSX_MIXED_FP32 if mixed else 0)
    for r, st in zip(out, stats):
        r["stats"] = st
    return out


def score_batch_device(problems, d_poses, ns, d_score, d_c1, d_c2, mixed=False):
    """Device-resident batched form: per-complex lists of device addresses."""
    flags = FSX_DEVICE_POSES | FSX_DEVICE_OUTPUT | (FSX_MIXED_FP32 if mixed else 0)
    return _batch_call(problems, d_poses, ns, d_score, d_c1, d_c2, None, flags)


def score_device(problem, d_poses, n, d_score, d_c1, d_c2, d_cross=0, mixed=False,
                 fit_compact=False):
    """Device-resident form: all pointers are device addresses (ints)."""
    st = Stats()
    flags = (FSX_DEVICE_POSES | FSX_DEVICE_OUTPUT | (FSX_MIXED_FP32 if mixed else 0)
             | (FSX_FIT_COMPACT if fit_compact else 0))
    _chk(lib().fsx_score(C.byref(problem.s), d_poses, n, d_score, d_c1, d_c2, d_cross or None,
                         flags, C.byref(st)))
    return st.as_dict()


def fit_host(ct, qvals, a, mult, c1=(0.96, 1.04), c2=(-2.0, 4.0), mixed=False,
             compact=False):
    """Host fit of one rescaled profile ct [6,Q] -> (score, c1, c2, evals).  mixed: the
    mixed-precision form's fit (one-pass evaluation, L-BFGS-B with the explicit 2 x 2
    matrix; with compact, the compact-form L-BFGS-B) instead of the reference's."""
    ct = f64(ct)
    out = np.zeros(4)
    flags = (FSX_MIXED_FP32 | (FSX_FIT_COMPACT if compact else 0)) if mixed else 0
    _chk(lib().fsx_fit_host_form(ct.reshape(-1), f64(qvals), ct.shape[-1], f64(a), mult,
                                 c1[0], c1[1], c2[0], c2[1], flags, out))
    return out


def fit_batch(ct, qvals, a, mult, c1=(0.96, 1.04), c2=(-2.0, 4.0)):
    """GPU fits of rescaled profiles ct [n,6,Q] -> [n,4] (score, c1, c2, evals)."""
    ct = f64(ct)
    n = ct.shape[0]
    out = np.zeros((n, 4))
    _chk(lib().fsx_fit_batch(ct.reshape(-1), n, f64(qvals), ct.shape[-1], f64(a), mult,
                             c1[0], c1[1], c2[0], c2[1], out.reshape(-1)))
    return out


def sbessel_batch(x, pmax):
    x = f64(x)
    out = np.zeros((x.size, pmax + 1))
    _chk(lib().fsx_sbessel_batch(x, x.size, pmax, out.reshape(-1)))
    return out


def atoms2spf(coords, vac, dum, sa, vdw, qvals, L, h2o_ff=3.50968, mixed=False):
    """GPU atom expansion (pdb2spf.c:24-152).  Returns (re, im, rm) with
    re/im shaped [3, Q, (L+1)^2].  mixed: the sums over atoms in split BF16
    on MFMA with FP32 accumulation (fsx_atoms2spf_form, FSX_MIXED_FP32)."""
    coords = f64(coords).reshape(-1, 3)
    qvals = f64(qvals)
    re = np.zeros((3, qvals.size, (L + 1) ** 2))
    im = np.zeros_like(re)
    rm = C.c_double(0.0)
    sa_arr = None if sa is None else f64(sa)
    _chk(lib().fsx_atoms2spf_form(coords.shape[0], coords.ravel(), f64(vac), f64(dum),
                                  None if sa_arr is None else sa_arr.ctypes.data, h2o_ff,
                                  f64(vdw), qvals, qvals.size, L,
                                  FSX_MIXED_FP32 if mixed else 0, re.reshape(-1), im.reshape(-1),
                                  C.byref(rm)))
    return re, im, rm.value


def score_naive(rec, lig, qvals, L, a, mult, poses, c1=(0.96, 1.04), c2=(-2.0, 4.0),
                sa=None, h2o_ff=3.50968, want_cross=False):
    """Brute-force per-pose scores (tools/score_ft_naive.c:156-212): rec/lig
    are (coords [n,3], vac, dum, vdw) tuples in their starting frames, sa the
    per-atom SASA fraction of the starting complex (receptor atoms first) or
    None, poses [n,6] = (z, b1, g1, a2, b2, g2).  Returns a dict with score,
    c1, c2 and, if asked, the cross-term profiles [n,6,Q]."""
    xyz = f64(np.vstack([f64(rec[0]).reshape(-1, 3), f64(lig[0]).reshape(-1, 3)]))
    nrec, nlig = f64(rec[0]).reshape(-1, 3).shape[0], f64(lig[0]).reshape(-1, 3).shape[0]
    vac = f64(np.concatenate([rec[1], lig[1]]))
    dum = f64(np.concatenate([rec[2], lig[2]]))
    vdw = f64(np.concatenate([rec[3], lig[3]]))
    poses = f64(poses).reshape(-1, 6)
    qvals = f64(qvals)
    n = poses.shape[0]
    out = np.zeros((n, 3))
    cross = np.zeros((n, 6, qvals.size)) if want_cross else None
    sa_arr = None if sa is None else f64(sa)
    _chk(lib().fsx_score_naive(nrec, nlig, xyz.ravel(), vac, dum,
                               None if sa_arr is None else sa_arr.ctypes.data, h2o_ff, vdw,
                               qvals, qvals.size, L, f64(a), mult, c1[0], c1[1], c2[0], c2[1],
                               poses.ravel(), n, out.reshape(-1),
                               None if cross is None else cross.ctypes.data))
    r = {"score": out[:, 0].copy(), "c1": out[:, 1].copy(), "c2": out[:, 2].copy()}
    if want_cross:
        r["cross"] = cross
    return r


def sbessel(l, x):
    return lib().sxs_sbessel(l, x)


def mkarray(begin, end, n):
    """saxs_utils.c:50-63 q grid (cumulative addition) from the library."""
    L = lib()
    p = L.sxs_mkarray(begin, end, n)
    out = np.ctypeslib.as_array(p, shape=(n,)).copy()
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    libc.free(C.cast(p, C.c_void_p))
    return out


def fit_params(exp_q, exp_in, exp_err, qvals, rm):
    """sxs_opt_params_create (min_saxs.c:108-124) for an experimental curve
    given as arrays: returns (a [6*Q], mult, peak)."""
    L = lib()
    eq, ei, ee, q = f64(exp_q), f64(exp_in), f64(exp_err), f64(qvals)
    prof = L.sxs_profile_create(_ptr(eq), eq.size, 0)
    C.memmove(prof.contents.in_, ei.ctypes.data, 8 * eq.size)
    C.memmove(prof.contents.err, ee.ctypes.data, 8 * eq.size)
    par = L.sxs_opt_params_create(prof, _ptr(q), q.size, rm)
    a = np.ctypeslib.as_array(par.contents.a, shape=(6 * q.size,)).copy()
    mult, peak = par.contents.mult, par.contents.peak
    L.sxs_opt_params_free(par)
    L.sxs_profile_free(prof)
    return a, mult, peak


# ---- reference-struct helpers (drop-in API) -------------------------------------
def spf_from_arrays(re, im, L, rm=0.0):
    """Build a struct sxs_spf_full (owned by the library) from [3,Q,(L+1)^2]."""
    re = f64(re)
    im = f64(im)
    Q = re.shape[1]
    p = lib().sxs_spf_full_create(L, Q)
    p.contents.rm = rm
    K2 = (L + 1) ** 2
    for q in range(Q):
        for X, arr in enumerate((p.contents.V, p.contents.D, p.contents.W)):
            s = arr[q].contents
            C.memmove(s.re, re[X, q].ctypes.data, 8 * K2)
            C.memmove(s.im, im[X, q].ctypes.data, 8 * K2)
    return p


def spf_to_arrays(p):
    s = p.contents
    L, Q = s.L, s.qnum
    K2 = (L + 1) ** 2
    re = np.zeros((3, Q, K2))
    im = np.zeros((3, Q, K2))
    for q in range(Q):
        for X, arr in enumerate((s.V, s.D, s.W)):
            re[X, q] = np.ctypeslib.as_array(arr[q].contents.re, shape=(K2,))
            im[X, q] = np.ctypeslib.as_array(arr[q].contents.im, shape=(K2,))
    return re, im, s.rm
