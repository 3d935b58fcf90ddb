"""The EM autotuner (mc_autotune) on the GPU: it measures candidates, its
pick keeps parity with the oracle, and results are cached."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from massiveclicks_amd import DBN, UBM, ClickModel, Dataset, autotune, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "massiveclicks_amd", "lib", "mclicks")


@pytest.mark.parametrize("model", [DBN, UBM], ids=["DBN", "UBM"])
def test_autotuned_options_keep_parity_and_are_cached(model, tmp_path):
    name = "dbn" if model == DBN else "ubm"
    s = synth.generate(synth.SynthConfig(n_sessions=20000, n_queries=1500, vocab=60000, model=name, seed=77))
    ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
    cache = str(tmp_path / "tune.json")
    print(f"case: autotune model={model}", flush=True)
    opts, info = autotune(model, ds.train, ds.n_qd, iterations=3, cache=cache)
    assert not info["cached"] and info["trials"] >= 5 and info["ms_per_iter"] > 0
    again, info2 = autotune(model, ds.train, ds.n_qd, iterations=3, cache=cache)
    assert info2["cached"] and again == opts
    assert list(json.load(open(cache)))[0].startswith(name.upper())

    tr = ds.train
    om = oracle.OracleModel(model, tr.query, tr.pidx, tr.clicks, ds.n_qd, accum_fp64=True)
    om.iterate(50)
    with ClickModel(model, **opts) as cm:
        cm.init_from_dataset(ds)
        cm.em(50)
        for t in range(cm.n_tables()):
            _, _, ov = om.get_table(t)
            assert float(np.max(np.abs(cm.get_values(t) - ov))) < 1e-5, (opts, t)


def test_cli_autotune_flag(tmp_path):
    s = synth.generate(synth.config_pbm_40k())
    path = str(tmp_path / "pbm40k.txt")
    synth.write_raw_log(s, path)
    r = subprocess.run([CLI, "-r", path, "-s", "40000", "-i", "10", "-m", "0", "--autotune", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Autotune (" in r.stdout and "Itr: 9" in r.stdout
