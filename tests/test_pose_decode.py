"""The pose decode of the bucketing kernel (csrc/fsx_pose.h), compiled for the
host with gcc and checked against the mixed-radix decode of the reference
(src/index.c:9-29) written out in Python, including indices whose z digit
does not fit a 32-bit int: those are outside the grid and must be reported
as not scored, never wrapped into a valid z slot."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include "fsx_pose.h"
int pose_key(long long idx, int znum, int nbeta, int na, int ng, int *row, int *alpha)
{
    return fsx_pose_key(idx, znum, nbeta, na, ng, row, alpha);
}
"""


@pytest.fixture(scope="module")
def keyfn(tmp_path_factory):
    d = tmp_path_factory.mktemp("pose")
    src, so = d / "pose.c", d / "libpose.so"
    src.write_text(SRC)
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-shared", "-fPIC",
                    "-I" + os.path.join(ROOT, "libfmftsaxs_amd", "csrc"), str(src), "-o",
                    str(so)], check=True)
    lib = C.CDLL(str(so))
    lib.pose_key.argtypes = [C.c_longlong] + [C.c_int] * 4 + [C.POINTER(C.c_int)] * 2
    lib.pose_key.restype = C.c_int

    def f(idx, znum, nb, na, ng):
        r, a = C.c_int(-7), C.c_int(-7)
        k = lib.pose_key(int(idx), znum, nb, na, ng, C.byref(r), C.byref(a))
        return k, r.value, a.value
    return f


def _expect(idx, znum, nb, na, ng):
    if idx < 0:
        return -1, 0, 0
    g2 = idx % ng; idx //= ng
    g1 = idx % ng; idx //= ng
    a2 = idx % na; idx //= na
    b2 = idx % nb; idx //= nb
    b1 = idx % nb
    z = idx // nb
    if z >= znum:
        return -1, 0, 0
    return (z * nb + b2) * ng + g2, b1 * ng + g1, a2


@pytest.mark.parametrize("grid", [(6, 7, 13, 13), (36, 19, 64, 64), (101, 37, 128, 128)])
def test_pose_key_matches_mixed_radix(keyfn, grid):
    znum, nb, na, ng = grid
    size5d = nb * nb * na * ng * ng
    rng = np.random.default_rng(znum)
    idx = list(rng.integers(0, znum * size5d, 2000))
    # grid edges, out-of-range z slots, negative indices, and z quotients at
    # and beyond 2^31 / 2^32 (which a 32-bit z digit would wrap into range)
    idx += [0, size5d - 1, znum * size5d - 1, znum * size5d, znum * size5d + 5, -1, -3,
            (2 ** 31 - 1) * size5d, (2 ** 31) * size5d + 5, (2 ** 32 + 1) * size5d + 7,
            (2 ** 32) * size5d, 2 ** 62, 2 ** 63 - 1]
    for i in idx:
        assert keyfn(i, znum, nb, na, ng) == _expect(int(i), znum, nb, na, ng), i
