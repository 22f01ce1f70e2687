"""Admission rule of the serving loop, with stand-in requests (CPU only)."""
from collections import deque

from particleholography.serve import select_batch


class _R:
    def __init__(self, rid, fp):
        self.rid, self._fp = rid, fp

    def footprint(self):
        return self._fp


def _ids(b):
    return [r.rid for r in b]


def test_fifo_up_to_max_batch():
    w = deque(_R(i, 10) for i in range(10))
    assert _ids(select_batch(w, 0, 1000, 4, False)) == [0, 1, 2, 3]
    assert _ids(select_batch(w, 0, 1000, 4, True)) == [4, 5, 6, 7]
    assert len(w) == 2


def test_budget_stops_admission_in_order():
    w = deque([_R(0, 40), _R(1, 50), _R(2, 5)])
    # 40 + 50 > 80: the second request waits, and so does the third (no overtaking)
    assert _ids(select_batch(w, 0, 80, 8, False)) == [0]
    assert _ids(select_batch(w, 40, 80, 8, True)) == []
    assert _ids(select_batch(w, 0, 80, 8, False)) == [1, 2]


def test_oversized_request_runs_alone_when_idle():
    w = deque([_R(0, 500), _R(1, 1)])
    assert _ids(select_batch(w, 0, 100, 8, True)) == []        # others in flight: wait
    assert _ids(select_batch(w, 0, 100, 8, False)) == [0]      # idle: admitted alone
    assert _ids(select_batch(w, 0, 100, 8, False)) == [1]
