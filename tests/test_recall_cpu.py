"""CPU checks of the recall helpers (turboquant/recall.py) against
definitions that fix them: recall of a ranking against itself is 1, against
a disjoint one 0; the exact top-k equals a brute force by sorting; the fp64
brute force is chunk-independent."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _recall():
    from turboquant import recall
    return recall


def test_recall_at_k_definition():
    r = _recall()
    a = np.arange(40).reshape(4, 10)
    assert r.recall_at_k(a, a, 10) == 1.0
    assert r.recall_at_k(a, a + 1000, 10) == 0.0
    b = a.copy()
    b[:, 5:] += 1000  # half of each row replaced
    assert r.recall_at_k(b, a, 10) == 0.5
    assert r.recall_at_k(b, a, 5) == 1.0
    with pytest.raises(ValueError):
        r.recall_at_k(a[:3], a, 10)


def test_topk_and_exact_against_sorting():
    r = _recall()
    rng = np.random.default_rng(0)
    base = rng.standard_normal((3000, 24))
    q = rng.standard_normal((7, 24))
    d = ((q[:, None, :] - base[None, :, :]) ** 2).sum(-1)
    want = np.argsort(d, axis=1, kind="stable")[:, :10]
    np.testing.assert_array_equal(r.topk_from_distances(d, 10), want)
    got = r.exact_topk(torch.from_numpy(q), torch.from_numpy(base), 10).numpy()
    np.testing.assert_array_equal(got, want)
    got_small_chunks = r.exact_topk(torch.from_numpy(q), torch.from_numpy(base), 10, chunk=257).numpy()
    np.testing.assert_array_equal(got_small_chunks, want)
