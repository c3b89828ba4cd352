"""Bucketing of shots by channel count (qpc_gpu.bucket_shots): the groups
partition the shots, are contiguous in channel count, and reach the least
padded channel count over every split into at most K contiguous groups
(brute force on small cases)."""
import itertools

import numpy as np

import qpc_gpu


def _padded(ldims, groups):
    return sum(len(g) * max(ldims[k] for k in g) for g in groups)


def _brute(ldims, K):
    ls = sorted(ldims)
    n = len(ls)
    best = None
    for k in range(1, min(K, n) + 1):
        for cuts in itertools.combinations(range(1, n), k - 1):
            b = [0, *cuts, n]
            v = sum((b[i + 1] - b[i]) * ls[b[i + 1] - 1] for i in range(k))
            best = v if best is None else min(best, v)
    return best


def test_bucket_shots_optimal_and_partition():
    rng = np.random.default_rng(3)
    for _ in range(60):
        n = int(rng.integers(1, 9))
        ldims = [int(x) for x in rng.integers(3, 61, n)]
        for K in (1, 2, 3, 4):
            groups = qpc_gpu.bucket_shots(ldims, K)
            assert 1 <= len(groups) <= K
            assert sorted(k for g in groups for k in g) == list(range(n))
            # contiguous in channel count: every shot of a group is at most the smallest of the next
            for a, b in zip(groups, groups[1:]):
                assert max(ldims[k] for k in a) <= min(ldims[k] for k in b)
            assert _padded(ldims, groups) == _brute(ldims, K)


def test_bucket_shots_config4_split():
    """The 64 config-4 shots in 4 buckets: 2546 padded channels for 2324 real
    ones, against 3840 when every shot is padded to the largest."""
    import bench
    shots = bench.batch_shots(bench.CFG_BATCH)
    ldims = [s["L"] for s in shots]
    assert sum(ldims) == 2324 and len(ldims) * max(ldims) == 3840
    assert _padded(ldims, qpc_gpu.bucket_shots(ldims, 4)) == 2546
