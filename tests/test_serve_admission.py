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


class _W(_R):
    def __init__(self, rid, fp, work):
        super().__init__(rid, fp)
        self._work = work

    def work_bytes(self):
        return self._work


def test_work_cap_closes_groups_but_admits_one():
    w = deque([_W(0, 1, 30), _W(1, 1, 30), _W(2, 1, 50), _W(3, 1, 5)])
    assert _ids(select_batch(w, 0, 1000, 8, False, max_work=64)) == [0, 1]   # 30 + 30 <= 64 < 110
    assert _ids(select_batch(w, 0, 1000, 8, False, max_work=40)) == [2]      # first one always fits
    assert _ids(select_batch(w, 0, 1000, 8, False, max_work=40)) == [3]
