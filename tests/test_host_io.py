"""Host-side pieces that need no GPU: the parameter CSV writer
(Communicate::output_parameters, communicator.cpp:415-516), the C++ header
include/mclicks.hpp compiled into a client program, and the command-line
driver's flag handling (main.cpp:222-288)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from massiveclicks_amd import Dataset, _lib
from massiveclicks_amd._lib import check, f32p

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIXTURE = os.path.join(HERE, "golden", "parse_fixture.txt")
CLI = os.path.join(ROOT, "massiveclicks_amd", "lib", "mclicks")


def write_csv(ds, prefix, model_type, tables):
    nums = [np.ascontiguousarray(n, np.float32) for n, _ in tables]
    dens = [np.ascontiguousarray(d, np.float32) for _, d in tables]
    pn = (f32p * len(tables))(*[n.ctypes.data_as(f32p) for n in nums])
    pd = (f32p * len(tables))(*[d.ctypes.data_as(f32p) for d in dens])
    check(_lib.lib().mc_write_parameters_csv(ds._h, str(prefix).encode(), model_type, pn, pd))


def test_csv_files_of_a_pbm_model(tmp_path):
    """The fixture's train set is query 100 (docs 11..20, indices 0..9) and
    query 200 (docs 21, 22, 23, 21, 25..30; indices 10, 11, 12, 10, 13..18).
    Attractiveness entry i is set to (i, 4 i) = 0.25 except entry 10, set to
    (3, 3), which Param::value clamps to 1 - 1e-6; examination rank r is
    (1, r + 2).  Per-qd lines use std::to_string (6 decimals) and cover
    every rank of every train session, the header being overwritten by the
    data; public lines use the ostream default (6 significant digits)."""
    ds = Dataset.parse(FIXTURE, -1, 0.2)
    n_qd = ds.n_qd
    an = np.arange(n_qd, dtype=np.float32)
    ad = 4 * np.arange(n_qd, dtype=np.float32)
    an[0], ad[0] = 1, 4
    an[10], ad[10] = 3, 3
    en = np.ones(10, np.float32)
    ed = np.arange(10, dtype=np.float32) + 2
    write_csv(ds, tmp_path / "out", 0, [(an, ad), (en, ed)])
    exm = (tmp_path / "out_examination.csv").read_text()
    assert exm.splitlines() == ["rank, examination", "0, 0.5", "1, 0.333333", "2, 0.25", "3, 0.2",
                                "4, 0.166667", "5, 0.142857", "6, 0.125", "7, 0.111111", "8, 0.1", "9, 0.0909091"]
    atr = (tmp_path / "out_attractiveness.csv").read_text().splitlines()
    assert len(atr) == 20
    assert atr[:10] == [f"100, {d}, 0.250000" for d in range(11, 21)]
    assert atr[10:] == ["200, 21, 0.999999", "200, 22, 0.250000", "200, 23, 0.250000", "200, 21, 0.999999"] + \
        [f"200, {d}, 0.250000" for d in range(25, 31)]


def test_csv_files_of_a_dbn_model(tmp_path):
    """DBN: two per-qd files and the public continuation file; entries at
    (1, 2) are written as 0.500000, gamma (1, 10) as 0.1."""
    q = np.array([5], np.int32)
    docs = np.arange(10, dtype=np.int32).reshape(1, 10)
    ds = Dataset.from_sessions(q, docs, np.zeros((1, 10), np.uint8), 0.2)
    # one session: int(1 * 0.8f - 0.5) + 1 = 1 train session
    n_qd = ds.n_qd
    t = (np.ones(n_qd, np.float32), 2 * np.ones(n_qd, np.float32))
    write_csv(ds, tmp_path / "dbn", 2, [t, t, (np.ones(1, np.float32), np.ones(1, np.float32) * 10)])
    assert (tmp_path / "dbn_continuation.csv").read_text() == "rank, continuation\n0, 0.1\n"
    sat = (tmp_path / "dbn_satisfaction.csv").read_text().splitlines()
    assert sat == [f"5, {d}, 0.500000" for d in range(10)]


def test_csv_writer_rejects_bad_arguments(tmp_path):
    ds = Dataset.parse(FIXTURE, -1, 0.2)
    L = _lib.lib()
    assert L.mc_write_parameters_csv(ds._h, str(tmp_path / "x").encode(), 7, None, None) == 1
    n = np.ones(ds.n_qd, np.float32)
    pn = (f32p * 2)(n.ctypes.data_as(f32p), n.ctypes.data_as(f32p))
    assert L.mc_write_parameters_csv(ds._h, b"/nonexistent/dir/x", 0, pn, pn) == 4


def test_cpp_header_client_builds_and_runs(tmp_path):
    """include/mclicks.hpp compiles with g++ alone and drives the library:
    dataset preparation works without a GPU, and failures surface as
    mclicks::Error with the C status."""
    exe = tmp_path / "client"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "cpp", "hpp_client.cpp"), "-L", libdir, "-lmclicks",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), FIXTURE], check=True, capture_output=True, text=True).stdout.splitlines()
    assert out[0] == "parsed 3 train 2 test 1 n_qd 19"
    assert out[1:3] == ["train 100", "train 200"]
    assert out[3] in ("model on device", f"model error 3")
    assert out[4] == "io error 4"


def test_cli_flags():
    """Help, unknown flags and the one-GPU / GPU-only limits of the driver."""
    r = subprocess.run([CLI, "-h", "x"], capture_output=True, text=True)
    assert r.returncode == 0 and "--raw-path" in r.stdout
    r = subprocess.run([CLI, "--bogus", "1"], capture_output=True, text=True)
    assert r.returncode == 0 and "Did not recognize flag '--bogus'" in r.stderr and "Usage" in r.stdout
    r = subprocess.run([CLI, "-g", "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "one GPU" in r.stderr
    r = subprocess.run([CLI, "-e", "1"], capture_output=True, text=True)
    assert r.returncode == 1 and "GPU-only" in r.stderr
    r = subprocess.run([CLI, "-m", "5"], capture_output=True, text=True)
    assert r.returncode == 1 and "Invalid model type" in r.stderr
    r = subprocess.run([CLI, "-i", "abc"], capture_output=True, text=True)
    assert "Invalid argument 'abc' for flag '-i'" in r.stderr
