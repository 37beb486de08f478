"""Reference d-symbol table through the reference's own GMP 3j routine
(borrowed.c:73-222 as compiled into oracle/_ref/libref_math.so), evaluated
with the reference loop of fftsaxs.c:138-174.  Test infrastructure only."""
import ctypes as C
import math

import numpy as np

_GMP_PATHS = ["/opt/conda/lib/libgmp.so.10", "libgmp.so.10"]


def _gmp():
    for p in _GMP_PATHS:
        try:
            return C.CDLL(p)
        except OSError:
            continue
    raise OSError("libgmp not found")


def ref_w3j_fn(ref, nmax):
    gmp = _gmp()
    gmp.__gmpf_init.argtypes = [C.c_void_p]
    gmp.__gmpf_clear.argtypes = [C.c_void_p]
    gmp.__gmpf_get_d.argtypes = [C.c_void_p]
    gmp.__gmpf_get_d.restype = C.c_double
    m = ref.math
    m.generate_desc_fact_array_arb.argtypes = [C.c_int]
    m.generate_desc_fact_array_arb.restype = C.c_void_p
    m.wigner_3j_symbol_arb.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.c_void_p]
    fact = m.generate_desc_fact_array_arb(nmax)
    res = C.create_string_buffer(64)
    gmp.__gmpf_init(res)

    def w3j(j1, j2, j3, m1, m2, m3):
        m.wigner_3j_symbol_arb(res, j1, j2, j3, m1, m2, m3, fact)
        return gmp.__gmpf_get_d(res)
    return w3j


def ref_d_symb(ref, L):
    w3j = ref_w3j_fn(ref, 4 * L + 1)
    N = 2 * L + 1
    out = np.zeros((L + 1) ** 3 * N)
    for l in range(L + 1):
        k1 = float(2 * l + 1)
        for l1 in range(L + 1):
            k2 = math.sqrt((2 * l1 + 1) * k1)
            for p in range(abs(l - l1), l + l1 + 1):
                k3 = (2 * p + 1) * k2 * w3j(l, p, l1, 0, 0, 0)
                mm = min(l, l1)
                for mv in range(-mm, mm + 1):
                    k4 = w3j(l, p, l1, -mv, 0, mv)
                    idx = ((l * (l + 1) + mv) * (L + 1) + l1) * N + p
                    out[idx] = k3 * k4
    return out
