"""Host logic of particle neighbourhoods (ccl.jl:266-466), against the reference's tests and
the rules of its docstrings. CPU only."""
import math

from particleholography import ccl


def test_update_particle_neighborhoods3d_known_answer():
    """runtests.jl:235-244."""
    pbbs = ccl.gen_particle_neighborhoods([(1, 1, 3, 3)], 1)
    ccl.update_particle_neighborhoods3d_(pbbs, [(2, 2, 4, 4), (10, 10, 12, 12)], 2)
    assert len(pbbs) == 2
    vals = list(pbbs.values())
    assert [1, 1, 1, 4, 4, 2] in vals
    assert [10, 10, 2, 12, 12, 2] in vals


def test_overlap_is_strict():
    assert ccl.judge_overlap2d((1, 1, 3, 3), (2, 2, 4, 4))
    assert not ccl.judge_overlap2d((1, 1, 3, 3), (3, 3, 5, 5))  # touching corners
    assert not ccl.judge_overlap2d((5, 5, 5, 5), (5, 5, 5, 5))  # single pixel boxes never overlap
    assert ccl.judge_overlap3d((1, 1, 1, 4, 4, 3), (2, 2, 2, 5, 5, 4))
    assert not ccl.judge_overlap3d((1, 1, 1, 4, 4, 2), (2, 2, 2, 5, 5, 4))


def test_update_2d_extends_z_and_merges_all_overlaps():
    pbbs = ccl.gen_particle_neighborhoods([(1, 1, 4, 4), (10, 10, 14, 14)], 3)
    ccl.update_particle_neighborhoods_(pbbs, [(2, 2, 12, 12)], 1)  # overlaps both: both grow
    vals = sorted(pbbs.values())
    assert vals == [[1, 1, 1, 12, 12, 3], [2, 2, 1, 14, 14, 3]]
    ccl.update_particle_neighborhoods_(pbbs, [(30, 30, 31, 31)], 5)
    assert [30, 30, 5, 31, 31, 5] in pbbs.values() and len(pbbs) == 3


def test_unit_extent_neighbourhoods_are_dropped():
    """ccl.jl:356-360: |dx| = |dy| = |dz| = 1 after an update removes the neighbourhood."""
    pbbs = ccl.gen_particle_neighborhoods([(1, 1, 2, 2)], 1)
    ccl.update_particle_neighborhoods_(pbbs, [(1, 1, 2, 2)], 2)
    assert pbbs == {}


def test_finalize_rules():
    pn = {1: [1, 1, 1, 10, 10, 5],    # kept
          2: [20, 20, 1, 30, 30, 2],  # two slices deep (|dz| == 1): dropped
          3: [40, 40, 1, 60, 44, 5],  # x/y ratio 20/4 = 5 > 3: dropped
          4: [70, 70, 1, 72, 72, 5],  # area 2*2 = 4 < 10: dropped
          5: [80, 80, 1, 80, 90, 5],  # ratio 0/10 = 0 < 1/3: dropped
          6: [100, 100, 1, 100, 100, 5]}  # 0/0 = NaN passes the ratio test, area 0 drops it
    ccl.finalize_particle_neighborhoods_(pn)
    assert list(pn) == [1]


def test_delete_duplicates_merges_transitively():
    pn = {1: [1, 1, 1, 5, 5, 5], 2: [4, 4, 4, 9, 9, 9], 3: [8, 8, 8, 12, 12, 12]}
    ccl.delete_duplicates_(pn)
    assert list(pn.values()) == [[1, 1, 1, 12, 12, 12]]


def test_ratio_semantics():
    assert ccl._ratio(3, 0) == math.inf and math.isnan(ccl._ratio(0, 0))
