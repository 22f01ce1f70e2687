"""Admission rule of the serving loop, with stand-in requests (CPU only)."""
from collections import deque

from particleholography.serve import group_work_cap, select_batch


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


class _S(_W):
    def __init__(self, rid, fp, work, size):
        super().__init__(rid, fp, work)
        self.size = size


def test_same_size_lookahead_groups_one_size_and_keeps_the_others_in_order():
    w = deque(_S(i, 1, 10, s) for i, s in enumerate(["a", "b", "a", "c", "a", "b", "a"]))
    key = lambda r: r.size  # noqa: E731
    # the oldest request and the same-size ones within the window of 5 (ids 0, 2, 4; 6 is beyond)
    assert _ids(select_batch(w, 0, 1000, 8, False, lookahead=5, key=key)) == [0, 2, 4]
    assert _ids(w) == [1, 3, 5, 6]
    assert _ids(select_batch(w, 0, 1000, 8, False, lookahead=5, key=key)) == [1, 5]
    assert _ids(w) == [3, 6]
    # limits still apply: max_batch, the work cap (the first request always fits), the budget
    w = deque(_S(i, 1, 30, "a") for i in range(6))
    assert _ids(select_batch(w, 0, 1000, 2, False, lookahead=8, key=key)) == [0, 1]
    assert _ids(select_batch(w, 0, 1000, 8, False, max_work=64, lookahead=8, key=key)) == [2, 3]
    assert _ids(select_batch(w, 998, 1000, 8, False, lookahead=8, key=key)) == [4, 5]
    # the head waits when it does not fit and other work is in flight
    w = deque([_S(0, 50, 1, "a"), _S(1, 1, 1, "a")])
    assert _ids(select_batch(w, 60, 100, 8, True, lookahead=8, key=key)) == []
    assert _ids(w) == [0, 1]


def test_group_work_cap_grows_with_the_queue_only():
    assert group_work_cap(None, 10 ** 15, 2) is None
    assert group_work_cap(32, 0, 2) == 32            # short queue: the configured cap
    assert group_work_cap(32, 16 * 32, 2) == 32      # exactly eight groups per slot: unchanged
    assert group_work_cap(32, 16 * 320, 2) == 320    # a deep queue: an eighth of it per slot
