"""Synthetic workload generator: seeded, Zipf query popularity, pools of
distinct documents, ground-truth click rates."""
import numpy as np

from massiveclicks_amd import synth


def test_generator_is_seeded_and_shapes_hold():
    cfg = synth.SynthConfig(n_sessions=20000, n_queries=2000, vocab=50000, zipf_s=1.1, model="dbn", seed=11)
    a, b = synth.generate(cfg), synth.generate(cfg)
    np.testing.assert_array_equal(a.query, b.query)
    np.testing.assert_array_equal(a.clicks, b.clicks)
    assert a.docs.shape == (20000, 10) and a.clicks.dtype == np.uint8
    # documents of a session are distinct members of the query's pool
    assert all(len(set(row)) == 10 for row in a.docs[:500])


def test_zipf_head_share_matches_the_law():
    cfg = synth.SynthConfig(n_sessions=200000, n_queries=2000, vocab=50000, zipf_s=1.1, model="pbm", seed=12)
    s = synth.generate(cfg)
    k = np.arange(1, 2001)
    p = k ** -1.1 / np.sum(k ** -1.1)
    assert abs(s.stats["top_query_share"] - p[0]) < 0.01


def test_pbm_click_rates_follow_ground_truth():
    cfg = synth.SynthConfig(n_sessions=100000, n_queries=500, vocab=20000, model="pbm", seed=13)
    s = synth.generate(cfg)
    rate = s.clicks.mean(axis=0)
    expect = (2 / 7) / (np.arange(10) + 1.0) ** 0.8     # E[alpha] = 2/7
    np.testing.assert_allclose(rate, expect, rtol=0.05)


def test_variable_length_sessions():
    cfg = synth.SynthConfig(n_sessions=5000, n_queries=100, vocab=5000, model="dbn", variable_length=True, seed=14)
    s = synth.generate(cfg)
    assert s.length.min() == 1 and s.length.max() == 10
    for i in range(200):
        L = s.length[i]
        assert (s.docs[i, L:] == -1).all() and (s.clicks[i, L:] == 0).all()
