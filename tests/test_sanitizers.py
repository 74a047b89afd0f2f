"""The host C++ and the oracle built with AddressSanitizer and
UndefinedBehaviorSanitizer (gcc/g++, host code only) and run: the
device-free host library (pattern compression, models, eigensystems,
discrete Gamma, workloads, the uninformative-pattern set, Brent and Newton)
through tests/sanitize/host_driver.cpp, the oracle through
tests/sanitize/oracle_driver.c on a workload, its results equal to the
ordinary build's."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from refmath import indicators

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc and g++")


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, **kw)
    assert r.returncode == 0, (cmd[0], r.stdout[-2000:], r.stderr[-4000:])
    return r.stdout


def test_host_library_under_sanitizers(tmp_path):
    host = os.path.join(ROOT, "iqtree2_amd", "csrc", "host")
    exe = str(tmp_path / "host_driver")
    srcs = [os.path.join(host, f) for f in ("alignment.cpp", "model.cpp", "synth.cpp")]
    run(["g++", "-std=c++17", *SAN, "-I" + os.path.join(ROOT, "include"), "-I" + host,
         os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"), *srcs, "-o", exe])
    assert "host ok" in run([exe])


def test_oracle_under_sanitizers(tmp_path):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=12, nsite=600, ncat=4, mean_len=0.2, alpha=0.6,
                    p_inv=0.1, gap_frac=0.05, seed=9)
    planner = ia.PhyloTree(None, w)
    ops, node, dad, length = planner.plan(*planner.branches()[2])
    planner.close()
    m = w.model_arrays()
    ind = indicators(ia.DNA, 4, w.state_unknown)
    path = tmp_path / "workload.bin"
    with open(path, "wb") as f:
        f.write(np.array([w.ntaxa, 4, 4, w.npatterns, w.ntaxa, w.state_unknown, len(ops)], np.int32).tobytes())
        f.write(np.ascontiguousarray(w.get("states"), np.uint8).tobytes())
        for a in (w.get("freq"), w.get("invar"), m["eval"], m["evec"], m["inv_evec"], m["tip"], m["rates"],
                  m["props"]):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(np.ascontiguousarray(ops).tobytes())
        f.write(np.array([node.is_tip, node.id, dad.is_tip, dad.id], np.int32).tobytes())
        f.write(np.array([length], np.float64).tobytes())
        for a in (ind, w.get("Q"), w.get("pi")):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
    exe = str(tmp_path / "oracle_driver")
    run(["gcc", "-std=gnu11", "-fopenmp", *SAN, "-I" + os.path.join(ROOT, "oracle"),
         os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"), os.path.join(ROOT, "oracle", "oracle.c"),
         "-lm", "-o", exe])
    out = [list(map(float, line.split())) for line in run([exe, str(path)]).strip().splitlines()]
    assert len(out) == 2
    for row in out:
        safe = bool(row[0])
        orc = OracleTree.from_workload(w, safe_numeric=safe, nthreads=2)
        lnl = orc.likelihood_branch(ops, node, dad, length)
        orc.likelihood_buffer(ops, node, dad)
        df, ddf = orc.likelihood_derv(length)
        assert abs(row[1] - lnl) <= 1e-12 * abs(lnl)
        assert abs(row[2] - df) <= 1e-10 * max(1.0, abs(df))
        assert abs(row[3] - ddf) <= 1e-10 * max(1.0, abs(ddf))
        assert row[9] == 0                                   # the state-basis model was accepted
        # the state-basis path on a reversible model: the same lnL (+I mass kept)
        assert abs(row[10] - lnl) <= 1e-10 * abs(lnl)
        assert abs(row[11] - df) <= 1e-7 * max(1.0, abs(df))
