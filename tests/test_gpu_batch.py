"""The batched entry point's graph path (fsx_score_batch): the first call
captures every complex's stages into a HIP graph, later calls with the same
launch parameters replay it.  Every replay is checked against the eager
launches of the same batch (FSX_BATCH_GRAPH=0) and against single calls,
with the program's own eager work (single fsx_score calls, which use other
device buffers) and synchronisations between the replays; a replay with
new pose contents in the same buffers must give the new poses' results.
The tile-shape tuner's table is written to FSX_TUNE_FILE.  The persistent
launches (one cross-term launch per m-chunk and kernel shape, one fit launch,
over all complexes) are checked against each complex's own pipeline.
Bit-identical throughout: the graph replays the same kernels on the same
data, and the tuned tile rows change no arithmetic."""
import os

import numpy as np
import pytest

from tests.problem import build_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import libfmftsaxs_amd as F
    if F.device_count() < 1:
        pytest.fail("no HIP device visible")
    return F


def _gp(F, pb):
    return F.ScoringProblem(pb["L"], pb["qvals"], pb["zvals"], pb["A_re"], pb["A_im"], pb["B_re"],
                            pb["B_im"], pb["a"], pb["params"].mult, pb["peak"],
                            nbeta=pb["nbeta"], nalpha=pb["na"], ngamma=pb["ng"])


def _same(r, q):
    for k in ("score", "c1", "c2", "cross"):
        if not np.array_equal(r[k], q[k], equal_nan=True):
            return False
    return True


def test_graph_replays_match_eager(F, tmp_path, monkeypatch):
    tune = tmp_path / "tune.txt"
    monkeypatch.setenv("FSX_TUNE_FILE", str(tune))
    F.release()   # a fresh tuner table (and graph)
    pbs = [build_problem(300, 200, 10, 0.4, 21, np.arange(22.0, 28.0), 120, seed=51),
           build_problem(400, 250, 15, 0.5, 25, np.arange(25.0, 31.0), 90, seed=52, nbeta=19,
                         na=64, ng=64),
           build_problem(400, 300, 20, 0.5, 17, np.arange(30.0, 34.0), 60, seed=53),
           build_problem(300, 200, 15, 0.5, 25, np.arange(25.0, 29.0), 70, seed=54)]
    other = build_problem(200, 150, 6, 0.3, 11, np.arange(20.0, 24.0), 30, seed=55)
    gps = [_gp(F, pb) for pb in pbs]
    poses = [pb["index"] for pb in pbs]
    for mixed in (False, True):
        print(f"batch graph, mixed={mixed}: eager", flush=True)
        monkeypatch.setenv("FSX_BATCH_GRAPH", "0")
        eager = F.score_batch(gps, poses, want_cross=True, mixed=mixed)
        singles = [F.score(gp, p, want_cross=True, mixed=mixed) for gp, p in zip(gps, poses)]
        for r, q in zip(eager, singles):
            assert _same(r, q)
        monkeypatch.setenv("FSX_BATCH_GRAPH", "1")
        for rep in range(4):
            print(f"batch graph, mixed={mixed}: launch {rep}", flush=True)
            rb = F.score_batch(gps, poses, want_cross=True, mixed=mixed)
            for r, q in zip(rb, eager):
                assert _same(r, q), rep
                assert r["stats"]["n_scored"] == q["stats"]["n_scored"]
            # the program's own eager work between the replays
            F.score(_gp(F, other), other["index"], want_cross=True, mixed=mixed)
        # new pose contents in the same buffers: a replay with new data
        rng = np.random.default_rng(56)
        poses2 = [np.sort(rng.choice(p.max() + 1, p.size, replace=False)).astype(np.int64)
                  for p in poses]
        rb2 = F.score_batch(gps, poses2, want_cross=True, mixed=mixed)
        for gp, p, r in zip(gps, poses2, rb2):
            assert _same(r, F.score(gp, p, want_cross=True, mixed=mixed))
        assert not np.array_equal(rb2[0]["score"], eager[0]["score"], equal_nan=True)
    lines = [l.split() for l in open(tune)]
    print("tuner table:", lines)
    # one entry per (order bucket, form) met: L 10, 15, 20 in both forms
    assert sorted((int(l[0]), int(l[1])) for l in lines) == [(0, 0), (0, 1), (1, 0), (1, 1),
                                                              (2, 0), (2, 1)]
    assert all(int(l[2]) in (1, 2, 4) for l in lines)


def test_persistent_launches_match_per_complex(F, monkeypatch):
    """The persistent form (FSX_BATCH_PERSIST=1: one cross-term launch per
    m-chunk index and kernel shape over all complexes, one fit launch over
    all complexes' pose queues) against the per-complex form (the default,
    FSX_BATCH_PERSIST=0: each complex's whole pipeline on one of four
    streams), eager and replayed, in both forms.  The complexes cover one
    and two m-chunks (L = 20 in the mixed form), two tile shapes and
    q grids of 11 to 70 points (lane groups of 2 to 8 lanes per fit): the
    same kernels on the same data, so bit-identical."""
    pbs = [build_problem(300, 200, 10, 0.4, 21, np.arange(22.0, 28.0), 150, seed=61),
           build_problem(400, 300, 20, 0.5, 17, np.arange(30.0, 34.0), 80, seed=62),
           build_problem(400, 250, 15, 0.5, 70, np.arange(25.0, 31.0), 100, seed=63, nbeta=19,
                         na=64, ng=64),
           build_problem(200, 150, 6, 0.3, 11, np.arange(20.0, 24.0), 40, seed=64)]
    gps = [_gp(F, pb) for pb in pbs]
    poses = [pb["index"] for pb in pbs]
    for mixed in (False, True):
        res = {}
        for persist in ("0", "1"):
            for graph in ("0", "1"):
                monkeypatch.setenv("FSX_BATCH_PERSIST", persist)
                monkeypatch.setenv("FSX_BATCH_GRAPH", graph)
                print(f"mixed={mixed} persist={persist} graph={graph}", flush=True)
                res[(persist, graph)] = F.score_batch(gps, poses, want_cross=True, mixed=mixed)
        base = res[("0", "0")]
        for k, rb in res.items():
            for r, q in zip(rb, base):
                assert _same(r, q), (mixed, k)
                assert r["stats"]["n_scored"] == q["stats"]["n_scored"]
                assert r["stats"]["fit_evals"] == q["stats"]["fit_evals"]
        for gp, p, r in zip(gps, poses, base):
            assert _same(r, F.score(gp, p, want_cross=True, mixed=mixed))
