"""ctypes binding of the oracle (FP64 CPU restatement of libfmftsaxs's FFT
dimer SAXS scoring path).

TEST INFRASTRUCTURE: only tests/, the smoke entry point and bench.py's
cpu_baseline leg may import this module, and only as the checker (or as the
timed CPU baseline).  The shipped library never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

PI_TRUNC = 3.14159265358
PI_EXACT = 3.14159265358979323846

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build(ref=True):
    """Compile liboracle.so (and oracle/_ref when /root/reference exists)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


class FitParams(C.Structure):
    _fields_ = [("a", C.POINTER(C.c_double)), ("rm", C.c_double),
                ("mult", C.c_double), ("peak", C.c_double),
                ("c1_lo", C.c_double), ("c1_hi", C.c_double),
                ("c2_lo", C.c_double), ("c2_hi", C.c_double),
                ("factr", C.c_double), ("pgtol", C.c_double), ("m", C.c_int)]


class Grid(C.Structure):
    _fields_ = [("L", C.c_int), ("nbeta", C.c_int), ("na", C.c_int),
                ("ng", C.c_int), ("dense_threshold", C.c_int),
                ("nthreads", C.c_int)]


FG_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double),
                    C.POINTER(C.c_double), C.POINTER(C.c_double))


class LbResult(C.Structure):
    _fields_ = [("task", C.c_int), ("iter", C.c_int), ("nfgv", C.c_int),
                ("nevals", C.c_int), ("f", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_set_pi.argtypes = [C.c_double]
        L.orc_get_pi.restype = C.c_double
        L.orc_sbessel.argtypes = [C.c_int, C.c_double]
        L.orc_sbessel.restype = C.c_double
        L.orc_sph_norm.argtypes = [C.c_int, _dp]
        L.orc_legendre.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_wigner_d.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_w3j.argtypes = [C.c_int] * 6
        L.orc_w3j.restype = C.c_double
        L.orc_d_symb.argtypes = [C.c_int, _dp]
        L.orc_atoms2spf.argtypes = [C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_double,
                                    _dp, _dp, C.c_int, C.c_int, _dp, _dp]
        L.orc_atoms2spf.restype = C.c_double
        L.orc_profile_from_spf.argtypes = [_dp, _dp, C.c_int, _dp, C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.orc_spf2cross_terms.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_scoring_helper.argtypes = [_dp, _dp, _dp, C.c_int, _dp, C.c_int, _dp]
        L.orc_params_init.argtypes = [C.POINTER(FitParams), _dp, C.c_double, C.c_double]
        L.orc_best_scale.argtypes = [_dp, _dp, C.c_int, C.POINTER(FitParams),
                                     C.c_double, C.c_double]
        L.orc_best_scale.restype = C.c_double
        L.orc_gradient.argtypes = [_dp, _dp, C.c_int, C.POINTER(FitParams), C.c_double,
                                   _dp, _dp, _dp]
        L.orc_compile_intensity.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double,
                                            C.c_double, _dp]
        L.orc_fit_one.argtypes = [_dp, _dp, C.c_int, C.POINTER(FitParams), C.c_int, _dp]
        L.orc_fit_one.restype = C.c_int
        L.orc_lbfgsb_minimize.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _ip, FG_FN,
                                          C.c_void_p, C.c_double, C.c_double, C.c_int,
                                          C.POINTER(LbResult)]
        L.orc_lbfgsb_minimize.restype = C.c_int
        L.orc_assemble_index.argtypes = [C.c_int] * 9
        L.orc_assemble_index.restype = C.c_int64
        L.orc_disassemble_index.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, _ip]
        L.orc_euler2index.argtypes = [C.c_double] * 6 + [_dp, C.c_int, C.c_int, _ip, C.c_int]
        L.orc_euler2index.restype = C.c_int
        L.orc_compute_saxs_scores.argtypes = [_dp, _dp, _dp, _lp, C.c_int, _dp, _dp, _dp,
                                              _dp, C.POINTER(FitParams), _dp, C.c_int, _dp,
                                              C.c_int, C.POINTER(Grid), C.c_int,
                                              C.c_void_p, C.c_void_p]
        L.orc_direct_pose_cross_terms.argtypes = [_dp, _dp, _dp, _dp, C.c_int, _dp, C.c_int] \
            + [C.c_double] * 6 + [_dp]
        _lib = L
    return _lib


def set_pi(pi):
    lib().orc_set_pi(pi)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- L1 math ---------------------------------------------------------------
def sbessel(l, x):
    return lib().orc_sbessel(int(l), float(x))


def sph_norm(N):
    out = np.zeros(N * (2 * N - 1))
    lib().orc_sph_norm(N, out)
    return out


def legendre(N, x):
    out = np.zeros(N * (2 * N - 1))
    lib().orc_legendre(N, float(x), out)
    return out


def wigner_d(L, beta):
    out = np.zeros((L + 1) * (2 * L + 1) ** 2)
    lib().orc_wigner_d(L, float(beta), out)
    return out


def w3j(j1, j2, j3, m1, m2, m3):
    return lib().orc_w3j(j1, j2, j3, m1, m2, m3)


def d_symb(L):
    out = np.zeros((L + 1) ** 3 * (2 * L + 1))
    lib().orc_d_symb(L, out)
    return out


# ---- atoms / profiles -------------------------------------------------------
def atoms2spf(coords, vac, dum, sa, vdw, qvals, L, h2o_ff=3.50968):
    """Returns (re, im, rm) with re/im shaped [3, Q, (L+1)^2]."""
    coords = f64(coords).reshape(-1, 3)
    n = coords.shape[0]
    qvals = f64(qvals)
    Q = qvals.size
    re = np.zeros((3, Q, (L + 1) ** 2))
    im = np.zeros_like(re)
    sa_arr = None if sa is None else f64(sa)
    rm = lib().orc_atoms2spf(n, coords.ravel(), f64(vac), f64(dum),
                             None if sa_arr is None else sa_arr.ctypes.data_as(C.c_void_p),
                             h2o_ff, f64(vdw), qvals, Q, L, re.reshape(-1), im.reshape(-1))
    return re, im, rm


def profile_from_spf(re, im, L, qvals, rm, c1, c2, rerr=0.05):
    qvals = f64(qvals)
    Q = qvals.size
    inn = np.zeros(Q)
    err = np.zeros(Q)
    lib().orc_profile_from_spf(f64(re).reshape(-1), f64(im).reshape(-1), L, qvals, Q,
                               rm, c1, c2, rerr, inn, err)
    return inn, err


def spf2cross_terms(re, im, L, qnum):
    out = np.zeros(6 * qnum)
    lib().orc_spf2cross_terms(f64(re).reshape(-1), f64(im).reshape(-1), L, qnum, out)
    return out.reshape(6, qnum)


# ---- fit -------------------------------------------------------------------
def scoring_helper(exp_q, exp_in, exp_err, qvals):
    qvals = f64(qvals)
    a = np.zeros(6 * qvals.size)
    lib().orc_scoring_helper(f64(exp_q), f64(exp_in), f64(exp_err), len(exp_q), qvals,
                             qvals.size, a)
    return a


class Params:
    """Holds an orc_fit_params struct and keeps its moment array alive."""

    def __init__(self, a, rm, peak, c1=(0.96, 1.04), c2=(-2.0, 4.0)):
        self.a = f64(a).copy()
        self.s = FitParams()
        lib().orc_params_init(C.byref(self.s), self.a, rm, peak)
        self.s.c1_lo, self.s.c1_hi = c1
        self.s.c2_lo, self.s.c2_hi = c2

    @property
    def ref(self):
        return C.byref(self.s)

    @property
    def mult(self):
        return self.s.mult


def best_scale(ct, qvals, params, c1, c2):
    ct = f64(ct)
    return lib().orc_best_scale(ct.reshape(-1), f64(qvals), ct.shape[-1], params.ref, c1, c2)


def gradient(ct, qvals, params, k, x):
    ct = f64(ct)
    g = np.zeros(2)
    f = np.zeros(1)
    lib().orc_gradient(ct.reshape(-1), f64(qvals), ct.shape[-1], params.ref, k, f64(x), g, f)
    return f[0], g


def compile_intensity(ct, qvals, rm, c1, c2):
    ct = f64(ct)
    out = np.zeros(ct.shape[-1])
    lib().orc_compile_intensity(ct.reshape(-1), f64(qvals), ct.shape[-1], rm, c1, c2, out)
    return out


def fit_one(ct, qvals, params, rescale=True):
    """Returns (score, c1, c2, nevals, rescaled_ct)."""
    ct = f64(ct).copy()
    out = np.zeros(3)
    n = lib().orc_fit_one(ct.reshape(-1), f64(qvals), ct.shape[-1], params.ref,
                          1 if rescale else 0, out)
    return out[0], out[1], out[2], n, ct


def lbfgsb_minimize(fg, x0, lo, hi, m=3, factr=1e7, pgtol=1e-5, max_evals=100000):
    """fg(x) -> (f, g); returns (x, LbResult, trace_of_x_evaluated)."""
    x = f64(x0).copy()
    n = x.size
    trace = []

    def cb(_ctx, xp, fp, gp):
        xv = np.array([xp[i] for i in range(n)])
        trace.append(xv)
        f, g = fg(xv)
        fp[0] = f
        for i in range(n):
            gp[i] = g[i]

    cfun = FG_FN(cb)
    res = LbResult()
    nbd = np.full(n, 2, dtype=np.int32)
    lib().orc_lbfgsb_minimize(n, m, x, f64(lo), f64(hi), nbd, cfun, None, factr, pgtol,
                              max_evals, C.byref(res))
    return x, res, trace


# ---- index ------------------------------------------------------------------
def assemble_index(z, b1, b2, a2, g1, g2, nbeta, na, ng):
    return lib().orc_assemble_index(z, b1, b2, a2, g1, g2, nbeta, na, ng)


def disassemble_index(index, nbeta, na, ng):
    out = np.zeros(6, dtype=np.int32)
    lib().orc_disassemble_index(int(index), nbeta, na, ng, out)
    return tuple(int(v) for v in out)


def euler2index(z, b1, g1, a2, b2, g2, zvals, L):
    out = np.zeros(4, dtype=np.int32)
    n = lib().orc_euler2index(z, b1, g1, a2, b2, g2, f64(zvals), len(zvals), L, out, 4)
    return [int(v) for v in out[:min(n, 4)]]


# ---- correlation driver ---------------------------------------------------------
def compute_saxs_scores(index_list, A_re, A_im, B_re, B_im, params, qvals, zvals, L,
                        nbeta=None, na=None, ng=None, skip=1, nthreads=0,
                        want_cross=False, dense_threshold=30):
    """Oracle sxs_compute_saxs_scores.  Returns dict with score/c1/c2 (NaN where
    the reference leaves the output untouched), cross [n,6,Q] and twiddle flags."""
    idx = np.ascontiguousarray(index_list, dtype=np.int64)
    n = idx.size
    qvals = f64(qvals)
    zvals = f64(zvals)
    g = Grid(L, nbeta or L + 1, na or 2 * L + 1, ng or 2 * L + 1, dense_threshold, nthreads)
    sc = np.full(n, np.nan)
    c1 = np.full(n, np.nan)
    c2 = np.full(n, np.nan)
    cross = np.zeros((n, 6, qvals.size)) if want_cross else None
    tw = np.full(n, -1, dtype=np.int32)
    lib().orc_compute_saxs_scores(sc, c1, c2, idx, n, f64(A_re).reshape(-1),
                                  f64(A_im).reshape(-1), f64(B_re).reshape(-1),
                                  f64(B_im).reshape(-1), params.ref, qvals, qvals.size,
                                  zvals, zvals.size, C.byref(g), skip,
                                  None if cross is None else cross.ctypes.data_as(C.c_void_p),
                                  tw.ctypes.data_as(C.c_void_p))
    return {"score": sc, "c1": c1, "c2": c2, "cross": cross, "twiddle_exact": tw}


def direct_pose_cross_terms(A_re, A_im, B_re, B_im, L, qvals, z, b1, g1, a2, b2, g2):
    qvals = f64(qvals)
    out = np.zeros(6 * qvals.size)
    lib().orc_direct_pose_cross_terms(f64(A_re).reshape(-1), f64(A_im).reshape(-1),
                                      f64(B_re).reshape(-1), f64(B_im).reshape(-1), L,
                                      qvals, qvals.size, z, b1, g1, a2, b2, g2, out)
    return out.reshape(6, qvals.size)


def mkarray(begin, end, n):
    """saxs_utils.c:50-63: q grid by cumulative addition (q[i] = q[i-1] + step)."""
    step = (end - begin) / (n - 1)
    q = np.zeros(n)
    q[0] = begin
    for i in range(1, n):
        q[i] = q[i - 1] + step
    return q


# ---- reference-built checker objects (oracle/_ref) ----------------------------
class Ref:
    """Loads oracle/_ref/*.so, compiled from the reference's own sources."""

    def __init__(self):
        self.math = C.CDLL(os.path.join(REF_DIR, "libref_math.so"))
        self.lb = C.CDLL(os.path.join(REF_DIR, "libref_lbfgsb.so"))
        m = self.math
        m.sxs_sbessel.argtypes = [C.c_int, C.c_double]
        m.sxs_sbessel.restype = C.c_double
        m.generate_spherical_norm.argtypes = [C.c_int]
        m.generate_spherical_norm.restype = C.POINTER(C.c_double)
        m.fill_assoc_Legendre_array.argtypes = [C.c_int, C.c_double, _dp]

        class DArray(C.Structure):
            _fields_ = [("L", C.c_int), ("data", C.POINTER(C.c_double))]
        self.DArray = DArray
        m.generate_d_array.argtypes = [C.c_int, C.c_double]
        m.generate_d_array.restype = C.POINTER(DArray)
        m.deallocate_d_array.argtypes = [C.POINTER(DArray)]
        self.libc = C.CDLL(None)
        self.libc.free.argtypes = [C.c_void_p]
        lb = self.lb
        LP = C.POINTER(C.c_long)
        DP = C.POINTER(C.c_double)
        lb.setulb.argtypes = [LP, LP, DP, DP, DP, LP, DP, DP, DP, DP, DP, LP, LP, LP, LP,
                              LP, LP, DP]
        lb.setulb.restype = C.c_int

    @staticmethod
    def available():
        return (os.path.exists(os.path.join(REF_DIR, "libref_math.so"))
                and os.path.exists(os.path.join(REF_DIR, "libref_lbfgsb.so")))

    def sbessel(self, l, x):
        return self.math.sxs_sbessel(int(l), float(x))

    def sph_norm(self, N):
        p = self.math.generate_spherical_norm(N)
        out = np.array([p[i] for i in range(N * (2 * N - 1))])
        self.libc.free(C.cast(p, C.c_void_p))
        return out

    def legendre(self, N, x):
        out = np.zeros(N * (2 * N - 1))
        self.math.fill_assoc_Legendre_array(N, float(x), out)
        return out

    def wigner_d(self, L, beta):
        p = self.math.generate_d_array(L, float(beta))
        n = (L + 1) * (2 * L + 1) ** 2
        out = np.ctypeslib.as_array(p.contents.data, shape=(n,)).copy()
        self.math.deallocate_d_array(p)
        return out

    def d_symb(self, L):
        """fftsaxs.c:138-174 evaluated with the reference's GMP 3j routine."""
        from ._gmp3j import ref_d_symb
        return ref_d_symb(self, L)

    def lbfgsb_minimize(self, fg, x0, lo, hi, m=3, factr=1e7, pgtol=1e-5, max_evals=100000):
        """Drive the reference setulb by reverse communication exactly like
        min_saxs.c:196-259.  Returns (x, f, n_evals, task, trace)."""
        n = len(x0)
        ln = C.c_long
        N_ = ln(n)
        M_ = ln(m)
        x = (C.c_double * n)(*x0)
        lo_ = (C.c_double * n)(*lo)
        hi_ = (C.c_double * n)(*hi)
        nbd = (ln * n)(*([2] * n))
        f = C.c_double(0.0)
        g = (C.c_double * n)()
        factr_ = C.c_double(factr)
        pgtol_ = C.c_double(pgtol)
        wa = (C.c_double * (2 * m * n + 5 * n + 11 * m * m + 8 * m + 43251))()
        iwa = (ln * (3 * n + 3072))()
        task = ln(1)
        iprint = ln(-1)
        csave = ln(0)
        lsave = (ln * 4)()
        isave = (ln * 44)()
        dsave = (C.c_double * 29)()
        trace = []
        nev = 0
        while True:
            self.lb.setulb(C.byref(N_), C.byref(M_), x, lo_, hi_, nbd, C.byref(f), g,
                           C.byref(factr_), C.byref(pgtol_), wa, iwa, C.byref(task),
                           C.byref(iprint), C.byref(csave), lsave, isave, dsave)
            t = task.value
            if 10 <= t <= 15:
                xv = np.array(x[:])
                trace.append(xv)
                fv, gv = fg(xv)
                f.value = fv
                for i in range(n):
                    g[i] = gv[i]
                nev += 1
                if nev >= max_evals:
                    break
                continue
            if t == 2:
                continue
            break
        return np.array(x[:]), f.value, nev, task.value, trace
