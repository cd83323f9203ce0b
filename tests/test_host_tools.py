"""Host-side tools: the bench contract's multi-GPU refusal, the training-set
cell generator, and argument checks of the host entry points."""
import os
import subprocess
import sys

import numpy as np
import pytest

import largetb_amd as L
from conftest import ROOT
from largetb_amd.learned import batch_structures


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_training_cells_follow_the_batched_distribution():
    cells = batch_structures(4096, seed=2026)
    assert cells == batch_structures(4096, seed=2026)  # reproducible from the seed
    L_, M_, N_, sig, seeds = (np.array(x) for x in zip(*cells))
    assert set(L_) == {2, 3, 4, 5} and set(M_) == {2, 3, 4, 5} and set(N_) == {2, 3, 4}
    assert sig.min() >= 0.02 and sig.max() <= 0.15 and abs(sig.mean() - 0.085) < 0.005
    assert len(set(seeds.tolist())) == len(cells)
    atoms = 9 * L_ * M_ * N_
    assert atoms.min() == 72 and atoms.max() == 900


def test_host_entry_points_reject_bad_arguments(tmp_path):
    with pytest.raises(L.LtbError):
        L.kpm_dos(np.ones(1), 1.0, 0.0, 10, 4.0)          # fewer than 2 moments
    with pytest.raises(L.LtbError):
        L.kpm_dos(np.ones(8), -1.0, 0.0, 10, 4.0)         # a <= 0
    with pytest.raises(L.LtbError):
        L.random_vectors(1, 0, 5)
    with pytest.raises(L.LtbError):
        L.write_dos(tmp_path / "no" / "dir" / "dos.txt", [0.0], [1.0])


def test_bench_traffic_comes_from_the_committed_counter_summary():
    """roofline.traffic: the scaled FETCH_SIZE + WRITE_SIZE figure of the profiled
    tiling, and None for any other workload or tiling."""
    sys.path.insert(0, ROOT)
    import bench
    t = dict(chunk_cols=4096, cols_per_lane=4, schedule=2, blocks_per_cu=2, nontemporal=1, order=2)
    v = bench.measured_traffic([12, 12, 12], 512, t, "fp32")
    assert v is not None and 3.0e9 < v < 5.0e9        # algorithmic 3.062 GB per launch
    assert bench.measured_traffic([8, 8, 8], 512, t, "fp32") is None
    assert bench.measured_traffic([12, 12, 12], 512, dict(t, cols_per_lane=2), "fp32") is None
    assert bench.measured_traffic([12, 12, 12], 512, t, "fp64") is None



def test_host_vectors_are_checked_before_the_c_side_reads_them():
    """The C entry points read natoms doubles from every per-atom array, so the
    front end refuses a vector of another length before the call."""
    v = L._vec([1, 2, 3], 3, "p")
    assert v.dtype == np.float64 and v.flags["C_CONTIGUOUS"]
    with pytest.raises(ValueError, match="2 values given, 3 expected"):
        L._vec(np.zeros(2), 3, "p")
    with pytest.raises(ValueError):
        L._vec(np.zeros((2, 2)), 3, "p")
