"""Dataset preparation (parse, split, sort, qd indices, test filtering):
hand-derived expectations from the reference's rules, checked against both
the CPU oracle and the engine's host code (mc_dataset_* in libmclicks.so,
which runs on the CPU)."""
import json
import os

import numpy as np
import pytest

import oracle
from massiveclicks_amd import Dataset, synth
from massiveclicks_amd._lib import MCError

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "parse_fixture.txt")


def both_parse(path, max_sessions, test_share):
    o = oracle.parse_raw(path, max_sessions, test_share)
    d = Dataset.parse(path, max_sessions, test_share)
    return o, d


def test_parse_fixture_rules():
    """dataset.cpp:224-305, search.cu:93-122:
    * query lines (15 tokens) open a SERP; it is emitted when the next query
      line arrives, so the SERP of query 300 (last in the file) is dropped;
    * a click marks the first rank showing that document (doc 21 of query
      200 is at ranks 0 and 3: rank 0 is marked); clicks on documents not in
      the SERP (99, 5) change nothing;
    * lines whose first token is not a number are skipped; lines with other
      token counts are ignored (the 4-token line "3 1 X 5" is a click line).
    Split of 3 sessions at 0.2: int(3 * 0.8f - 0.5) + 1 = 2 train, 1 test."""
    o, d = both_parse(FIXTURE, -1, 0.2)
    assert o["n_parsed"] == d.n_parsed == 3
    assert len(d.train) == 2 and len(d.test) == 1
    np.testing.assert_array_equal(d.train.query, [100, 200])
    np.testing.assert_array_equal(d.train.clicks[0], [0, 0, 1, 0, 0, 0, 0, 0, 0, 0])
    np.testing.assert_array_equal(d.train.clicks[1], [1, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    # first-occurrence indices; the repeated document 21 shares index 10
    np.testing.assert_array_equal(d.train.pidx[0], np.arange(10))
    np.testing.assert_array_equal(d.train.pidx[1], [10, 11, 12, 10, 13, 14, 15, 16, 17, 18])
    assert d.n_qd == o["n_qd"] == 19
    np.testing.assert_array_equal(d.test.query, [100])
    np.testing.assert_array_equal(d.test.clicks[0], [0] * 9 + [1])
    np.testing.assert_array_equal(d.test.pidx[0], np.arange(10))
    for k in ("query", "docs", "clicks", "pidx"):
        np.testing.assert_array_equal(o["train"][k], getattr(d.train, k))
        np.testing.assert_array_equal(o["test"][k], getattr(d.test, k))


def test_parse_stops_at_max_sessions():
    """The read loop stops once max_sessions SERPs were emitted (dataset.cpp:250)."""
    o, d = both_parse(FIXTURE, 2, 0.2)
    assert o["n_parsed"] == d.n_parsed == 2
    np.testing.assert_array_equal(d.train.query, [100, 200])
    assert len(d.test) == 0


@pytest.mark.parametrize("n,share,n_train", [
    (1, 0.2, 1), (2, 0.2, 2), (3, 0.2, 2), (5, 0.2, 4), (10, 0.2, 8), (40000, 0.2, 32000),
    (6, 0.25, 5), (10, 0.3, 7), (3, 0.5, 2), (7, 0.0, 7),
])
def test_split_rounding(n, share, n_train):
    """n_train = int(n * (1 - share) - 0.5) + 1 with (1 - share) in float and
    the product in float (dataset.cpp:390); worked by hand per case."""
    q = np.arange(n, dtype=np.int32)
    docs = np.arange(n * 10, dtype=np.int32).reshape(n, 10)
    clicks = np.zeros((n, 10), np.uint8)
    d = Dataset.from_sessions(q, docs, clicks, share)
    o = oracle.from_sessions(q, docs, clicks, share)
    assert len(d.train) == len(o["train"]["query"]) == n_train


def test_test_set_filtering():
    """assign_test_parameter_indices (dataset.cpp:349-374): a test session
    with an unseen query, or with no document seen for its query, is
    dropped; unseen documents of a kept session get index -1; order kept."""
    q = np.array([7, 8, 999, 7, 8, 7], np.int32)
    docs = np.array([np.arange(10) + 100, np.arange(10) + 200, np.arange(10) + 100,
                     np.arange(10) + 500, np.r_[200, 201, np.arange(8) + 900], np.arange(10) + 105],
                    np.int32)
    clicks = np.zeros((6, 10), np.uint8)
    share = 0.6   # int(6 * 0.4f - 0.5) + 1 = 2 train sessions
    d = Dataset.from_sessions(q, docs, clicks, share)
    o = oracle.from_sessions(q, docs, clicks, share)
    assert len(d.train) == 2
    np.testing.assert_array_equal(d.test.query, [8, 7])
    np.testing.assert_array_equal(d.test.pidx[0], [10, 11] + [-1] * 8)
    np.testing.assert_array_equal(d.test.pidx[1], [5, 6, 7, 8, 9] + [-1] * 5)
    for k in ("query", "docs", "pidx"):
        np.testing.assert_array_equal(o["test"][k], getattr(d.test, k))


def test_test_lookup_in_queries_with_many_documents():
    """Queries whose train sessions show many distinct documents (more than
    the 32 a query group is scanned for) and few: every test pair resolves
    to the index the oracle assigns, unseen documents to -1."""
    rng = np.random.default_rng(11)
    n = 4000
    q = rng.integers(0, 6, n).astype(np.int32)
    # query 0 and 1 draw from 400 documents, the others from 12
    pool = np.where(q[:, None] < 2, rng.integers(0, 400, (n, 10)), rng.integers(0, 12, (n, 10)))
    docs = (pool + 1000 * q[:, None]).astype(np.int32)
    clicks = (rng.random((n, 10)) < 0.2).astype(np.uint8)
    d = Dataset.from_sessions(q, docs, clicks, 0.3)
    o = oracle.from_sessions(q, docs, clicks, 0.3)
    assert d.n_qd == o["n_qd"] and d.n_qd > 800
    for part in ("train", "test"):
        for k in ("query", "docs", "pidx"):
            np.testing.assert_array_equal(o[part][k], getattr(getattr(d, part), k))
    assert (d.test.pidx == -1).any() and (d.test.pidx >= 0).any()


def test_train_order_is_libstdcxx_sort_of_equal_keys():
    """The train set is ordered by std::sort on the query (dataset.cpp:19-20),
    which is not stable: the committed fixture holds the permutation
    libstdc++ produces for 3000 keys with ties (tests/golden/sort_ties.json,
    made by tests/golden/make_sort_fixture.py); it differs from a stable sort."""
    fx = json.load(open(os.path.join(HERE, "golden", "sort_ties.json")))
    keys = np.array(fx["keys"], np.int32)
    perm = np.array(fx["perm"])
    assert fx["differs_from_stable"]
    n = len(keys)
    docs = (np.arange(n)[:, None] * 10 + np.arange(10)[None, :]).astype(np.int32)
    clicks = np.zeros((n, 10), np.uint8)
    d = Dataset.from_sessions(keys, docs, clicks, 0.0)
    o = oracle.from_sessions(keys, docs, clicks, 0.0)
    np.testing.assert_array_equal(d.train.docs[:, 0] // 10, perm)
    np.testing.assert_array_equal(o["train"]["docs"][:, 0] // 10, perm)


def test_synthetic_log_round_trip_matches_oracle(tmp_path):
    """The 40k-session PBM workload written in the raw format and parsed by
    both implementations: identical train/test sets and indices."""
    s = synth.generate(synth.config_pbm_40k())
    path = str(tmp_path / "pbm40k.txt")
    synth.write_raw_log(s, path)
    o, d = both_parse(path, 40000, 0.2)
    assert d.n_parsed == 40000 and len(d.train) == 32000
    for k in ("query", "docs", "clicks", "pidx"):
        np.testing.assert_array_equal(o["train"][k], getattr(d.train, k))
        np.testing.assert_array_equal(o["test"][k], getattr(d.test, k))
    # sessions parsed back in log order = the generated ones
    n_tr = len(d.train)
    np.testing.assert_array_equal(np.sort(d.train.query), np.sort(s.query[:n_tr]))


def test_parse_errors_are_reported():
    with pytest.raises(MCError) as e:
        Dataset.parse("/nonexistent/log.txt", -1, 0.2)
    assert e.value.status == 4
    with pytest.raises(MCError) as e:
        Dataset.from_sessions(np.zeros(1, np.int32), np.zeros((1, 10), np.int32), np.zeros((1, 10), np.uint8), -0.5)
    assert e.value.status == 1


def test_lengths_of_ten_equal_the_fixed_length_preparation():
    """mc_dataset_from_sessions_len with every len = 10 is
    mc_dataset_from_sessions (and so the reference's preparation)."""
    s = synth.generate(synth.SynthConfig(n_sessions=3000, n_queries=300, vocab=6000, model="dbn", seed=5))
    a = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
    b = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2, length=np.full(3000, 10, np.uint8))
    assert a.n_qd == b.n_qd
    for part in ("train", "test"):
        for k in ("query", "docs", "clicks", "pidx"):
            np.testing.assert_array_equal(getattr(getattr(a, part), k), getattr(getattr(b, part), k))
        assert (getattr(b, part).length == 10).all()


def test_short_result_lists_index_only_shown_ranks():
    """Hand-worked: train = two sessions of query 7 (after the split of 3 at
    0.2): [1, 2, 3] (len 3) and [3, 4] (len 2); ranks past len hold docs that
    must be ignored.  Indices in first-occurrence order over shown ranks:
    1 -> 0, 2 -> 1, 3 -> 2, 4 -> 3.  The test session shows only doc 9
    (unseen; its rank 1 holds doc 1, which is not shown), so it is dropped."""
    q = np.array([7, 7, 7], np.int32)
    docs = np.full((3, 10), 1, np.int32)
    docs[0, :3] = [1, 2, 3]
    docs[1, :2] = [3, 4]
    docs[2, :1] = [9]
    clicks = np.zeros((3, 10), np.uint8)
    clicks[1, 1] = 1
    ln = np.array([3, 2, 1], np.uint8)
    d = Dataset.from_sessions(q, docs, clicks, 0.2, length=ln)
    assert (len(d.train), len(d.test), d.n_qd) == (2, 0, 4)
    np.testing.assert_array_equal(d.train.length, [3, 2])
    np.testing.assert_array_equal(d.train.pidx[0], [0, 1, 2] + [-1] * 7)
    np.testing.assert_array_equal(d.train.pidx[1], [2, 3] + [-1] * 8)
    # the same log with the test session's doc 1 at rank 0: kept, len 1
    docs[2, 0] = 1
    d = Dataset.from_sessions(q, docs, clicks, 0.2, length=ln)
    assert len(d.test) == 1
    np.testing.assert_array_equal(d.test.pidx[0], [0] + [-1] * 9)
    np.testing.assert_array_equal(d.test.length, [1])


def test_length_errors():
    q = np.zeros(2, np.int32)
    docs = np.zeros((2, 10), np.int32)
    clicks = np.zeros((2, 10), np.uint8)
    for bad in ([0, 10], [11, 3]):
        with pytest.raises(MCError) as e:
            Dataset.from_sessions(q, docs, clicks, 0.2, length=np.array(bad, np.uint8))
        assert e.value.status == 1
    clicks[1, 5] = 1
    with pytest.raises(MCError) as e:
        Dataset.from_sessions(q, docs, clicks, 0.2, length=np.array([10, 5], np.uint8))
    assert "past len" in str(e.value)


def _noisy_log(tmp_path, n_sessions, seed, error_at=None):
    """A synthetic log (about 85 bytes per session) with malformed lines
    spread through it: words, empty lines, lines of other token counts,
    clicks on absent documents, repeated session ids."""
    s = synth.generate(synth.SynthConfig(n_sessions=n_sessions, n_queries=n_sessions // 20,
                                         vocab=n_sessions, model="pbm", seed=seed))
    path = tmp_path / f"noisy{seed}.txt"
    synth.write_raw_log(s, str(path))
    lines = path.read_text().split("\n")
    rng = np.random.default_rng(seed)
    noise = ["garbage words here", "", "7 1 X 5", "12 0 Q 1 2", "3 0 C 123456789",
             "x 0 Q 1 0 1 2 3 4 5 6 7 8 9 10", "  ", "4 0 C"]
    for pos in sorted(rng.choice(len(lines), 2000, replace=False), reverse=True):
        lines.insert(int(pos), noise[int(rng.integers(len(noise)))])
    if error_at is not None:
        lines.insert(error_at, "5 0 Q 10 0 1 2 3 4 5 6 7 8 9 bad")
    path.write_text("\n".join(lines))
    return str(path)


def test_parallel_parse_of_a_large_noisy_log_matches_the_sequential_oracle(tmp_path):
    """A 35 MB log is tokenised in several chunks cut at line boundaries; the
    result equals the oracle's line-by-line read, with and without a
    session cap (including caps that end inside a later chunk)."""
    path = _noisy_log(tmp_path, 420_000, 21)
    assert os.path.getsize(path) > 32 << 20
    for cap in (-1, 400_000, 123_457, 5):
        o, d = both_parse(path, cap, 0.2)
        assert o["n_parsed"] == d.n_parsed
        for k in ("query", "docs", "clicks", "pidx"):
            np.testing.assert_array_equal(o["train"][k], getattr(d.train, k))
            np.testing.assert_array_equal(o["test"][k], getattr(d.test, k))


def test_malformed_id_is_reported_only_if_the_read_reaches_it(tmp_path):
    """A query line whose document id is not a number: an error when the
    session cap has not been reached before it (line number reported), none
    when the cap stops the read earlier (dataset.cpp:250)."""
    path = _noisy_log(tmp_path, 420_000, 22, error_at=700_000)
    with pytest.raises(MCError) as e:
        Dataset.parse(path, -1, 0.2)
    assert "bad document id at line 700001" in str(e.value)
    d = Dataset.parse(path, 1000, 0.2)
    assert d.n_parsed == 1000
