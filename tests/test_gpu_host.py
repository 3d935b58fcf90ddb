"""The command-line driver, the parameter CSV files and checkpoints on the
GPU (through libmclicks.so), against the CPU oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from massiveclicks_amd import CCM, DBN, PBM, UBM, ClickModel, Dataset, synth
from massiveclicks_amd._lib import MCError, check, lib

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "massiveclicks_amd", "lib", "mclicks")


def test_cli_trains_the_pbm_40k_log_like_the_oracle(tmp_path):
    """The 40,000-session PBM workload in the raw text format, run by the
    command line as `-r log -s 40000 -i 50 -m 0 -t 0.2 -o prefix`: held-out LL and
    per-rank perplexity match the oracle (1e-6 relative; the command line prints 8
    significant digits) and the attractiveness/examination CSV files hold the
    oracle's values within the 1e-5 parameter bar (plus print rounding)."""
    s = synth.generate(synth.config_pbm_40k())
    path = str(tmp_path / "pbm40k.txt")
    synth.write_raw_log(s, path)
    prefix = str(tmp_path / "pbm")
    print("case: cli pbm40k", flush=True)
    r = subprocess.run([CLI, "-r", path, "-s", "40000", "-i", "50", "-m", "0", "-t", "0.2", "-o", prefix],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "Itr: 49" in out and "Model type: PBM" in out
    llh = float(re.search(r"Total Log likelihood is: (\S+)", out).group(1))
    ppl = [float(x) for x in re.findall(r"Perplexity at rank \d is: (\S+)", out)]

    ds = oracle.parse_raw(path, 40000, 0.2)
    tr, te = ds["train"], ds["test"]
    om = oracle.OracleModel(PBM, tr["query"], tr["pidx"], tr["clicks"], ds["n_qd"], accum_fp64=True)
    om.iterate(50)
    oev = om.evaluate(te["pidx"], te["clicks"])
    n_test = len(te["query"])
    assert abs(llh - oev["llh_sum_f64"] / n_test) <= 1e-6 * abs(oev["llh_sum_f64"] / n_test)
    oppl = 2.0 ** (-oev["ppl_log2_sum_f64"] / oev["rank_count"])
    np.testing.assert_allclose(ppl, oppl, rtol=1e-6)

    _, _, atr = om.get_table(0)
    _, _, exm = om.get_table(1)
    rows = np.loadtxt(prefix + "_attractiveness.csv", delimiter=",")
    assert rows.shape == (len(tr["query"]) * 10, 3)
    np.testing.assert_array_equal(rows[:, 0], np.repeat(tr["query"], 10))
    np.testing.assert_array_equal(rows[:, 1], tr["docs"].reshape(-1))
    np.testing.assert_allclose(rows[:, 2], atr[tr["pidx"].reshape(-1)], atol=1e-5 + 5e-7)
    lines = open(prefix + "_examination.csv").read().splitlines()
    assert lines[0] == "rank, examination"
    vals = np.array([float(l.split(",")[1]) for l in lines[1:]])
    np.testing.assert_allclose(vals, exm, rtol=5e-6, atol=1e-5)


@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_checkpoint_resume_equals_an_uninterrupted_run(model, tmp_path):
    """20 iterations, save; a fresh model loads the checkpoint and runs 30
    more: every table equals that of one 50-iteration run bit for bit
    (deterministic mode; the iterates depend only on num/den)."""
    name = {PBM: "pbm", UBM: "ubm", DBN: "dbn", CCM: "ccm"}[model]
    s = synth.generate(synth.SynthConfig(n_sessions=6000, n_queries=600, vocab=20000, model=name, seed=11))
    ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
    ck = str(tmp_path / "ck.bin")
    print(f"case: checkpoint model={model}", flush=True)
    with ClickModel(model) as a:
        a.init_from_dataset(ds)
        a.em(20)
        check(lib().mc_save_parameters(a._h, ck.encode()))
    with ClickModel(model) as c:
        c.init_from_dataset(ds)
        c.em(50)
        want = [c.get_parameters(t) for t in range(c.n_tables())]
    with ClickModel(model) as b:
        b.init_from_dataset(ds)
        check(lib().mc_load_parameters(b._h, ck.encode()))
        b.em(30)
        for t in range(b.n_tables()):
            n, d = b.get_parameters(t)
            np.testing.assert_array_equal(n, want[t][0])
            np.testing.assert_array_equal(d, want[t][1])
    other = PBM if model != PBM else UBM
    with ClickModel(other) as x:
        x.init_from_dataset(ds)
        with pytest.raises(MCError):
            check(lib().mc_load_parameters(x._h, ck.encode()))


def test_cli_fits_with_adam(tmp_path):
    """`--fit adam`: Adam epochs on the log-likelihood gradient instead of EM
    iterations; the held-out LL ends within 1% of the oracle's EM result on
    the same log."""
    s = synth.generate(synth.config_pbm_40k())
    path = str(tmp_path / "pbm40k.txt")
    synth.write_raw_log(s, path)
    print("case: cli adam", flush=True)
    r = subprocess.run([CLI, "-r", path, "-s", "40000", "-i", "30", "-m", "0", "--fit", "adam", "--batches", "4",
                        "--lr", "0.1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Epoch: 29" in r.stdout and "Adam epochs" in r.stdout
    llh = float(re.search(r"Total Log likelihood is: (\S+)", r.stdout).group(1))
    ds = oracle.parse_raw(path, 40000, 0.2)
    tr, te = ds["train"], ds["test"]
    om = oracle.OracleModel(PBM, tr["query"], tr["pidx"], tr["clicks"], ds["n_qd"], accum_fp64=True)
    om.iterate(50)
    ll_em = om.evaluate(te["pidx"], te["clicks"])["llh_sum_f64"] / len(te["query"])
    assert abs(llh - ll_em) <= 0.01 * abs(ll_em), (llh, ll_em)
