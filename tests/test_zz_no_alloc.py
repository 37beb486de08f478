"""fsx_atoms2spf_noalloc, the allocation-free form of the atom expansion,
through the C ABI on device arrays, against the allocating entry point
fsx_atoms2spf on the same atoms (bit-identical: the same kernels over the
same 1024-atom chunks, whatever the pass structure the fixed working memory
imposes) and against the oracle (1e-12 of each (X, q) block's largest
coefficient, the atom-expansion bar of tests/test_gpu_parity.py).  Two calls
in a row, the second smaller; then a problem too large for the fixed
working memory, which must be refused with FSX_ERR_SCRATCH and leave the
outputs untouched."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FSX_ERR_SCRATCH = -5


@pytest.fixture(scope="module")
def env():
    import libfmftsaxs_amd as F
    if F.device_count() < 1:
        pytest.fail("no HIP device visible")
    lib = F.lib()
    lib.fsx_atoms2spf_noalloc.argtypes = [C.c_int] + [C.c_void_p] * 4 + [C.c_double, C.c_void_p,
                                                                          C.c_int, C.c_int,
                                                                          C.c_void_p, C.c_void_p]
    lib.fsx_atoms2spf_noalloc.restype = C.c_int
    # the caller's device arrays, through the HIP runtime the library itself
    # links (a second runtime in the process, such as torch's, would not see
    # the device once this one holds it)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return F, lib, hip


H2D, D2H = 1, 2


class _Dev:
    """Device copies of host arrays, freed on exit."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *a):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _mol(natoms, seed, water=True):
    from libfmftsaxs_amd import synthetic as S
    m = S.random_chain(natoms, seed)
    m = m.translated(-m.center_of_extrema())
    if water:
        m.sa = S.shrake_rupley(m)
    return m


def _noalloc(env, m, q, L, fill=np.nan):
    F, lib, hip = env
    K2 = (L + 1) ** 2
    shape = (3, q.size, K2)
    with _Dev(hip) as d:
        xyz, vac, dum, qv = d.put(m.coords), d.put(m.vac), d.put(m.dum), d.put(q)
        sa = d.put(m.sa) if m.sa is not None else None
        re, im = d.put(np.full(shape, fill)), d.put(np.full(shape, fill))
        rc = lib.fsx_atoms2spf_noalloc(m.natoms, xyz, vac, dum, sa, 3.50968, qv, q.size, L,
                                       re, im)
        return rc, d.get(re, shape), d.get(im, shape)


def test_noalloc_matches_allocating_entry(env):
    F = env[0]
    cases = [(9000, 15, 51, 0.5, True), (700, 6, 26, 0.3, False)]   # second call smaller
    for natoms, L, Q, qmax, water in cases:
        print(f"noalloc atom expansion: {natoms} atoms, L={L}, {Q} q, water={water}", flush=True)
        m = _mol(natoms, 50 + L, water)
        q = O.mkarray(0.0, qmax, Q)
        rc, re, im = _noalloc(env, m, q, L)
        assert rc == 0, F.lib().fsx_error_string(rc)
        rg, ig, _ = F.atoms2spf(m.coords, m.vac, m.dum, m.sa, m.vdw, q, L)
        assert np.array_equal(re, rg) and np.array_equal(im, ig)
        ro, io, _ = O.atoms2spf(m.coords, m.vac, m.dum, m.sa, m.vdw, q, L)
        scale = np.maximum(np.abs(ro).max(axis=2, keepdims=True),
                           np.abs(io).max(axis=2, keepdims=True))
        scale = np.where(scale > 0, scale, 1.0)
        assert np.all(np.abs(re - ro) <= 1e-12 * scale)
        assert np.all(np.abs(im - io) <= 1e-12 * scale)
        if not water:
            assert np.all(re[2] == 0.0) and np.all(im[2] == 0.0)


def test_noalloc_refuses_oversized_problem(env):
    """5000 atoms at L = 40 with 400 q need about 231 MB of working memory."""
    m = _mol(5000, 61, water=False)
    q = O.mkarray(0.0, 0.5, 400)
    rc, re, im = _noalloc(env, m, q, 40, fill=7.0)
    assert rc == FSX_ERR_SCRATCH
    assert np.all(re == 7.0) and np.all(im == 7.0)
