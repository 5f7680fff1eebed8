"""The device-resident DFTable on the GPU: the reference tutorial's flows
(doc/tutorial/tutorial.md sections 5.3-5.6, the data of
doc/tutorial/src/tut5.3-5.6; string columns are carried as integer codes)
with their printed results as known answers, NULL handling through an outer
join, and a longer random chain checked against the same chain of oracle
calls."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def _rows(t, names=None):
    d = t.to_numpy()
    nulls = {c: set(v.tolist()) for c, v in t.null_rows().items()}
    names = names or t.columns()
    return [tuple(None if i in nulls[c] else d[c][i].item() for c in names)
            for i in range(t.num_row())]


A, B, C, D, E, F, G, H = range(8)  # codes of the string column values a..h


def test_tutorial_filter(fa):
    # tut5.3: c1 int, c2 double, c3 dic_string, c4 double
    t = fa.DFTable({"c1": np.arange(1, 9, dtype=np.int32),
                    "c2": np.arange(10, 90, 10).astype(np.float64),
                    "c3": np.arange(8, dtype=np.int64),
                    "c4": np.arange(9, 1, -1).astype(np.float64)})
    assert _rows(t.filter("c1", "eq", "c4"), ["c1"]) == [(5,)]
    assert _rows(t.filter("c1", "lt", "c4"), ["c1"]) == [(1,), (2,), (3,), (4,)]
    assert _rows(t.filter("c2", "eq", np.float64(30.0)), ["c1"]) == [(3,)]
    assert _rows(t.filter("c2", "le", np.float64(30.0)).filter("c3", "neq", np.int64(C)),
                 ["c1"]) == [(1,), (2,)]


def test_tutorial_sort(fa):
    # tut5.4: t.sort("c2").sort("c1")
    t = fa.DFTable({"c1": np.array([1, 3, 3, 2, 2, 2, 1, 1], np.int64),
                    "c2": np.array([50, 20, 30, 20, 30, 40, 40, 50], np.int64),
                    "c3": np.array([A, B, C, D, E, F, G, H], np.int64)})
    t3 = t.sort("c2").sort("c1")
    assert _rows(t3) == [(1, 40, G), (1, 50, A), (1, 50, H), (2, 20, D), (2, 30, E),
                         (2, 40, F), (3, 20, B), (3, 30, C)]
    t4 = t.filter("c1", "eq", np.int64(2)).sort("c2")
    assert _rows(t4, ["c3"]) == [(D,), (E,), (F,)]
    assert _rows(t.sort_desc("c2"), ["c3"]) == [(A,), (H,), (F,), (G,), (C,), (E,), (B,), (D,)]
    m = t3.materialize()
    assert _rows(m) == _rows(t3)


def _tut55(fa):
    t1 = fa.DFTable({"t1c1": np.arange(1, 9, dtype=np.int64),
                     "t1c2": np.arange(10, 90, 10, dtype=np.int64),
                     "t1c3": np.array([3, 2, 1, 2, 3, 2, 4, 1], np.int64)})
    t2 = fa.DFTable({"t2c1": np.array([1, 2, 3, 2], np.int64),
                     "t2c2": np.array([A, B, C, D], np.int64)})
    t3 = fa.DFTable({"t3c1": np.array([2, 1, 3], np.int64),
                     "t3c2": np.array([20, 10, 30], np.int64)})
    return t1, t2, t3


def test_tutorial_joins(fa):
    t1, t2, t3 = _tut55(fa)
    j = t1.hash_join(t2, "t1c3", "t2c1")
    assert sorted(_rows(j)) == sorted([
        (3, 30, 1, 1, A), (8, 80, 1, 1, A), (2, 20, 2, 2, D), (2, 20, 2, 2, B),
        (4, 40, 2, 2, D), (4, 40, 2, 2, B), (6, 60, 2, 2, D), (6, 60, 2, 2, B),
        (1, 10, 3, 3, C), (5, 50, 3, 3, C)])
    o = t1.outer_hash_join(t2, "t1c3", "t2c1")
    rows = _rows(o)
    assert (7, 70, 4, None, None) in rows and len(rows) == 11
    # the filtered join of the tutorial: t1c1 >= 2, t2c2 not "d"
    f = t1.filter("t1c1", "ge", np.int64(2)).hash_join(t2.filter("t2c2", "neq", np.int64(D)),
                                                       "t1c3", "t2c1")
    assert sorted(_rows(f)) == sorted([(3, 30, 1, 1, A), (8, 80, 1, 1, A), (2, 20, 2, 2, B),
                                       (4, 40, 2, 2, B), (6, 60, 2, 2, B), (5, 50, 3, 3, C)])
    # star join with the filtered t2 and t3 (both unique on the join key)
    s = t1.star_join([t2.filter("t2c2", "neq", np.int64(D)), t3],
                     [("t1c3", "t2c1"), ("t1c3", "t3c1")])
    want = [(i + 1, 10 * (i + 1), k, k, [A, B, C][k - 1], k, 10 * k)
            for i, k in enumerate([3, 2, 1, 2, 3, 2, 4, 1]) if k != 4]
    assert _rows(s) == want
    with pytest.raises(fa.FhipError, match="not unique"):
        t1.star_join([t2], [("t1c3", "t2c1")])


def test_outer_join_nulls_flow_into_filter_and_group_by(fa):
    t1, t2, _ = _tut55(fa)
    o = t1.outer_hash_join(t2, "t1c3", "t2c1")
    assert sorted(_rows(o.filter("t2c1", "eq", np.int64(2)), ["t1c1"])) == \
        [(2,), (2,), (4,), (4,), (6,), (6,)]
    # NULL rows never pass a comparison, not even neq
    assert o.filter("t2c1", "ge", np.int64(0)).num_row() == 10
    assert o.filter("t2c1", "neq", np.int64(-1)).num_row() == 10
    g = o.group_by("t1c3").select(["t1c3"], [("sum", "t2c1", "s"), ("count", "t2c1", "n"),
                                             ("max", "t1c2", "m")])
    # key 4 only has the outer row: its sum is NULL (max() with a null
    # entry, dfcolumn_grouped_agg.cc:232-270), its count 0
    assert _rows(g) == [(1, 2, 2, 80), (2, 12, 6, 60), (3, 6, 2, 50), (4, None, 0, 70)]
    assert g.column("s").numpy()[3] == np.iinfo(np.int64).max


def test_tutorial_group_by(fa):
    # tut5.6
    c1 = np.array([1, 3, 3, 2, 2, 2, 1, 1], np.int64)
    t = fa.DFTable({"c1": c1, "c2": np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int64),
                    "c3": np.array([30, 20, 30, 20, 30, 40, 40, 50], np.int64)})
    aggs = [("sum", "c3", "sum"), ("count", "c3", "count"), ("avg", "c3", "avg"),
            ("max", "c3", "max"), ("min", "c3", "min")]
    g = t.group_by("c1").select(["c1"], aggs)
    assert _rows(g) == [(1, 120, 3, 40.0, 50, 30), (2, 90, 3, 30.0, 40, 20),
                        (3, 50, 2, 25.0, 30, 20)]
    m = t.group_by(["c1", "c2"]).select(["c1", "c2"], aggs)
    assert set(_rows(m)) == {(1, 0, 30, 1, 30.0, 30, 30), (2, 0, 20, 1, 20.0, 20, 20),
                             (3, 0, 50, 2, 25.0, 30, 20), (1, 1, 90, 2, 45.0, 50, 40),
                             (2, 1, 70, 2, 35.0, 40, 30)}


def test_random_chain_matches_oracle_chain(fa):
    """filter -> sort -> sort -> multi-column group_by with aggregates, on
    2^20 rows, against the same chain of oracle calls."""
    rng = np.random.default_rng(30)
    n = 1 << 20
    k1 = rng.integers(0, 50, n).astype(np.int64)
    k2 = rng.integers(-20, 20, n).astype(np.int32)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    w = rng.integers(0, 1 << 20, n).astype(np.uint32)
    t = fa.DFTable({"k1": k1, "k2": k2, "v": v, "w": w})
    view = t.filter("w", "lt", np.uint32(700000)).sort("v").sort_desc("k2")
    pos = oracle.filter(w, "lt", np.uint32(700000))
    _, o1 = oracle.sort_index(v, pos)
    _, o2 = oracle.sort_index(k2, o1, descending=True)
    np.testing.assert_array_equal(view._idx.numpy(), o2)
    g = view.group_by(["k2", "k1"]).select(["k2", "k1"], [("sum", "v", "s"), ("max", "w", "m")])
    gi, split, (rk2, rk1) = oracle.multi_group_by([k2, k1], o2)
    np.testing.assert_array_equal(g.column("k2").numpy(), rk2)
    np.testing.assert_array_equal(g.column("k1").numpy(), rk1)
    np.testing.assert_array_equal(g.column("s").numpy(), oracle.grouped_agg("sum", v, gi, split))
    np.testing.assert_array_equal(g.column("m").numpy(), oracle.grouped_agg("max", w, gi, split))
