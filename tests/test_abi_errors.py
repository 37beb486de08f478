"""Argument checking of the extended C ABI (include/fsx.h) on the host.  Every
call here is either rejected or finished before the library touches a
device, so the tests run without a GPU.  Where the reference exits on bad
input (common.h:26-36), fsx_score returns FSX_ERR_ARG; an empty pose list is
a successful call that writes nothing."""
import ctypes as C

import numpy as np

import libfmftsaxs_amd as F

FSX_OK, FSX_ERR_ARG = 0, -1


def _problem(L=3, Q=4, nz=2):
    K2 = (L + 1) ** 2
    q = np.linspace(0.0, 0.3, Q)
    z = np.arange(20.0, 20.0 + nz)
    A = np.zeros((3, Q, K2))
    return F.ScoringProblem(L, q, z, A, A, A, A, np.ones(6 * Q), 1.0, 1.0)


def _score_rc(pb, n=1, outputs=True):
    poses = np.zeros(max(n, 1), dtype=np.int64)
    out = [np.full(max(n, 1), 7.0) for _ in range(3)]
    ptrs = [o.ctypes.data if outputs else None for o in out]
    st = F.Stats()
    rc = F.lib().fsx_score(C.byref(pb.s), poses.ctypes.data, n, *ptrs, None, 0, C.byref(st))
    return rc, out


def test_empty_pose_list_writes_nothing():
    pb = _problem()
    rc, out = _score_rc(pb, n=0)
    assert rc == FSX_OK
    assert all(np.all(o == 7.0) for o in out)
    r = F.score(pb, np.zeros(0, dtype=np.int64))
    assert r["score"].size == 0 and r["stats"]["n_scored"] == 0


def test_expansion_order_out_of_range_rejected():
    for L in (0, 41):
        pb = _problem()
        pb.s.L = L
        rc, out = _score_rc(pb)
        assert rc == FSX_ERR_ARG
        assert all(np.all(o == 7.0) for o in out)


def test_grid_and_q_sizes_rejected():
    for field, bad in (("nbeta", 1), ("nalpha", 0), ("ngamma", 0), ("qnum", 1), ("znum", 0)):
        pb = _problem()
        setattr(pb.s, field, bad)
        assert _score_rc(pb)[0] == FSX_ERR_ARG, field


def test_missing_arrays_rejected():
    for field in ("qvals", "zvals", "A_re", "A_im", "B_re", "B_im", "a"):
        pb = _problem()
        setattr(pb.s, field, None)
        assert _score_rc(pb)[0] == FSX_ERR_ARG, field
    assert F.lib().fsx_score(None, None, 0, None, None, None, None, 0, None) == FSX_ERR_ARG


def test_negative_count_and_missing_outputs_rejected():
    pb = _problem()
    assert _score_rc(pb, n=-1)[0] == FSX_ERR_ARG
    assert _score_rc(pb, n=1, outputs=False)[0] == FSX_ERR_ARG


def test_batch_arguments():
    lib = F.lib()
    assert lib.fsx_score_batch(-1, None, None, None, None, None, None, None, 0, None) \
        == FSX_ERR_ARG
    assert lib.fsx_score_batch(0, None, None, None, None, None, None, None, 0, None) == FSX_OK
    # a rejected problem stops the batch with its code
    pb = _problem()
    pb.s.L = 41
    arr = (F.Problem * 1)(pb.s)
    poses = np.zeros(1, dtype=np.int64)
    out = [np.zeros(1) for _ in range(3)]
    vp = lambda a: (C.c_void_p * 1)(a.ctypes.data)
    n = (C.c_int * 1)(1)
    assert lib.fsx_score_batch(1, arr, vp(poses), n, vp(out[0]), vp(out[1]), vp(out[2]),
                               None, 0, None) == FSX_ERR_ARG


def test_error_strings():
    lib = F.lib()
    seen = set()
    for code in (0, -1, -2, -3, -4):
        s = lib.fsx_error_string(code)
        assert s and s.decode().strip()
        seen.add(s)
    assert len(seen) == 5
    assert lib.fsx_error_string(-99)
