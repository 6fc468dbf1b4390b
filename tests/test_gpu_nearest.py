"""Nearest-node search (prrtc_nearest_async) on the GPU against a float64
brute force: the returned node is a nearest one up to float32 rounding,
exact ties go to the smaller index, and every launch shape (1, 2, 4 waves
per query) returns the same answers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(gpu, trees, n, queries, waves):
    import torch
    dev = torch.device("cuda:0")
    n_trees, dof, cap = trees.shape
    qpt = queries.shape[1]
    d_t = torch.from_numpy(np.ascontiguousarray(trees)).to(dev)
    d_q = torch.from_numpy(np.ascontiguousarray(queries)).to(dev)
    d_i = torch.zeros(n_trees * qpt, dtype=torch.int32, device=dev)
    d_d = torch.zeros(n_trees * qpt, dtype=torch.float32, device=dev)
    gpu.nearest_async(dof, d_t.data_ptr(), cap, n, n_trees, d_q.data_ptr(), qpt, waves, d_i.data_ptr(),
                      d_d.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return d_i.cpu().numpy().reshape(n_trees, qpt), d_d.cpu().numpy().reshape(n_trees, qpt)


@pytest.mark.parametrize("dof,n", [(7, 1500), (8, 1024), (14, 700), (3, 5000)])
def test_nearest_matches_brute_force(dof, n, gpu):
    rng = np.random.default_rng(dof * 1000 + n)
    cap = ((n + 1023) // 1024) * 1024
    n_trees, qpt = 3, 17
    trees = np.full((n_trees, dof, cap), np.nan, np.float32)     # storage past n must never be used
    trees[:, :, :n] = rng.uniform(-2, 2, (n_trees, dof, n)).astype(np.float32)
    queries = rng.uniform(-2, 2, (n_trees, qpt, dof)).astype(np.float32)
    # exact ties: node 900 % n duplicates node 5, queried at node 5
    dup = 900 % n
    trees[0, :, dup] = trees[0, :, 5]
    queries[0, 0] = trees[0, :, 5]
    answers = [run(gpu, trees, n, queries, w) for w in (1, 2, 4)]
    idx, d2 = answers[0]
    for i2, dd2 in answers[1:]:
        assert np.array_equal(i2, idx) and np.array_equal(dd2.view(np.int32), d2.view(np.int32))
    assert idx[0, 0] == min(5, dup) and d2[0, 0] == 0.0
    for t in range(n_trees):
        nodes = trees[t, :, :n].T.astype(np.float64)
        for j in range(qpt):
            dist = ((nodes - queries[t, j].astype(np.float64)) ** 2).sum(1)
            i = idx[t, j]
            assert 0 <= i < n
            assert dist[i] <= dist.min() * (1 + 1e-5) + 1e-6
            assert abs(d2[t, j] - dist[i]) <= 1e-5 * max(1.0, dist[i])



def test_nearest_overflowing_distances(gpu):
    """Distances that overflow float32 (nodes at 3e19, squares past FLT_MAX)
    or are NaN (a NaN query) compare below no bound: the answer is node 0
    with +inf, an index inside the tree, never one past it."""
    dof, n, cap = 7, 1500, 2048
    trees = np.full((2, dof, cap), np.nan, np.float32)
    trees[:, :, :n] = 3e19
    queries = np.zeros((2, 3, dof), np.float32)
    queries[1, 1, 2] = np.nan
    for w in (1, 2, 4):
        idx, d2 = run(gpu, trees, n, queries, w)
        assert np.array_equal(idx, np.zeros_like(idx)), (w, idx)
        assert np.all(np.isinf(d2)), (w, d2)
