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


# scalar restatement of ccl.jl:314-466, the check for the vectorised implementation
def _ov2(a, b):
    return a[0] < b[2] and a[2] > b[0] and a[1] < b[3] and a[3] > b[1]


def _scalar_update(pn, brs, s, three_d):
    sub = [v for v in pn if v[5] == s - 1] if three_d else None
    for br in brs:
        hit = False
        for v in (sub if three_d else list(pn)):
            if _ov2(br, (v[0], v[1], v[3], v[4])):
                v[0], v[1], v[3], v[4] = min(br[0], v[0]), min(br[1], v[1]), max(br[2], v[3]), max(br[3], v[4])
                if three_d:
                    v[5] = s
                elif s < v[2]:
                    v[2] = s
                elif s > v[5]:
                    v[5] = s
                hit = True
        if not hit:
            pn.append([br[0], br[1], s, br[2], br[3], s])
    pn[:] = [v for v in pn if not (abs(v[0] - v[3]) == 1 and abs(v[1] - v[4]) == 1 and abs(v[2] - v[5]) == 1)]


def _scalar_dedup(pn):
    changed = True
    while changed:
        changed = False
        i = 0
        while i < len(pn):
            v = pn[i]
            for j, w in enumerate(pn):
                if j != i and ccl.judge_overlap3d(tuple(v), tuple(w)):
                    v[:] = [min(v[0], w[0]), min(v[1], w[1]), min(v[2], w[2]),
                            max(v[3], w[3]), max(v[4], w[4]), max(v[5], w[5])]
                    del pn[j]
                    if j < i:
                        i -= 1
                    changed = True
                    break
            i += 1


import numpy as np  # noqa: E402
import pytest  # noqa: E402


@pytest.mark.parametrize("three_d", [False, True])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_vectorised_linking_equals_scalar_loops(three_d, seed):
    rng = np.random.default_rng(seed)
    slices = [[tuple(int(v) for v in (x, y, x + w, y + h))
               for x, y, w, h in zip(rng.integers(1, 60, k), rng.integers(1, 60, k), rng.integers(0, 8, k),
                                     rng.integers(0, 8, k))] for k in rng.integers(0, 12, 15)]
    pn = ccl.gen_particle_neighborhoods(slices[0], 1)
    ref = [list(v) for v in pn.values()]
    upd = ccl.update_particle_neighborhoods3d_ if three_d else ccl.update_particle_neighborhoods_
    for z in range(2, len(slices) + 1):
        upd(pn, slices[z - 1], z)
        _scalar_update(ref, slices[z - 1], z, three_d)
        assert list(pn.values()) == ref
    ccl.delete_duplicates_(pn)
    _scalar_dedup(ref)
    assert list(pn.values()) == ref


def _per_call(slices, three_d):
    pn = ccl.gen_particle_neighborhoods(slices[0], 1)
    upd = ccl.update_particle_neighborhoods3d_ if three_d else ccl.update_particle_neighborhoods_
    for z in range(2, len(slices) + 1):
        upd(pn, slices[z - 1], z)
    ccl.finalize_particle_neighborhoods_(pn)
    return list(pn.values())


@pytest.mark.parametrize("three_d", [False, True])
@pytest.mark.parametrize("seed,nmax,wmin,extent", [(7, 10, 2, 80), (8, 6, 0, 40), (9, 14, 1, 120), (10, 3, 1, 20)])
def test_native_linking_equals_per_call_updates(three_d, seed, nmax, wmin, extent):
    """libholo's holo_link_rectangles (link_slices) against the reference's per-call sequence
    gen / update! per slice / finalize!, itself checked against the scalar loops above; widths
    from 0 or 1 make unit cubes and degenerate extents."""
    rng = np.random.default_rng(seed)
    slices = [[tuple(int(v) for v in (x, y, x + w, y + h))
               for x, y, w, h in zip(rng.integers(1, extent, k), rng.integers(1, extent, k), rng.integers(wmin, 9, k),
                                     rng.integers(wmin, 9, k))] for k in rng.integers(0, nmax, 60)]
    assert list(ccl.link_slices(slices, three_d).values()) == _per_call(slices, three_d)


@pytest.mark.parametrize("three_d", [False, True])
def test_native_linking_edge_cases(three_d):
    assert ccl.link_slices([], three_d) == {}
    assert ccl.link_slices([[], [], []], three_d) == {}
    one = [[(1, 1, 5, 4), (10, 10, 12, 12)]]
    assert list(ccl.link_slices(one, three_d).values()) == _per_call(one, three_d)
    tall = [[(3, 3, 7, 7)]] * 5
    assert list(ccl.link_slices(tall, three_d).values()) == [[3, 3, 1, 7, 7, 5]]


def test_native_linking_rejects_bad_offsets():
    rects = np.zeros((2, 4), np.int32)
    with pytest.raises(AssertionError, match="decrease"):
        ccl.link_arrays(rects, np.array([0, 2, 1], np.int64))
    with pytest.raises(AssertionError, match="offsets"):
        ccl.link_arrays(rects, np.array([1, 2], np.int64))
    with pytest.raises(AssertionError, match="fewer rectangles"):
        ccl.link_arrays(rects, np.array([0, 3], np.int64))


@pytest.mark.parametrize("three_d", [False, True])
def test_native_linking_on_the_reference_holograms_stream(three_d):
    """The 9634 per-slice rectangles of the reference's test holograms (GS, 1000 N0f8 slices,
    threshold 10/255; tests/golden/rects_reference_holograms.npz): native linking equals the
    per-call sequence."""
    from pathlib import Path

    d = np.load(Path(__file__).resolve().parent / "golden" / "rects_reference_holograms.npz")
    rects, off = d["rects"], d["off"]
    slices = [[tuple(int(v) for v in r) for r in rects[off[z]:off[z + 1]]] for z in range(off.size - 1)]
    got = list(ccl.link_arrays(rects, off, three_d).values())
    assert got == _per_call(slices, three_d)
    assert len(got) == (36 if three_d else 23)


from hypothesis import given, settings, strategies as st  # noqa: E402

_coord = st.integers(min_value=-3, max_value=25)
_rect = st.tuples(_coord, _coord, st.integers(min_value=-2, max_value=7), st.integers(min_value=-2, max_value=7)).map(
    lambda t: (t[0], t[1], t[0] + t[2], t[1] + t[3]))  # negative widths: inverted rectangles
_stream = st.lists(st.lists(_rect, max_size=6), min_size=1, max_size=9)


@settings(max_examples=300, deadline=None)
@given(_stream, st.booleans())
def test_native_linking_property(slices, three_d):
    """Random streams, degenerate and inverted rectangles included (the latter take the scan
    without the grid): native linking equals the per-call sequence."""
    assert list(ccl.link_slices(slices, three_d).values()) == _per_call(slices, three_d)


_gcoord = st.integers(min_value=0, max_value=120)
_grect = st.tuples(_gcoord, _gcoord, st.integers(min_value=0, max_value=40), st.integers(min_value=0, max_value=40)).map(
    lambda t: (t[0], t[1], t[0] + t[2], t[1] + t[3]))


@settings(max_examples=300, deadline=None)
@given(st.lists(st.lists(_grect, max_size=8), min_size=1, max_size=12), st.booleans())
def test_native_linking_property_grid(slices, three_d):
    """Non-inverted rectangles over several grid cells (the 2-D update and the finalize take
    their candidates from the grid): native linking equals the per-call sequence."""
    assert list(ccl.link_slices(slices, three_d).values()) == _per_call(slices, three_d)
