"""The reference's own integration runs (test_occ_ri_k_h2o.sh:2 and the
corrected test_3c_transform_h2o.sh invocation) through the drop-in C++ drivers
on the device, checked against the h2o known answers."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, rel_fro

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(GOLDEN), "..", "strassen_amd", "bin")


def run(args, timeout=300):
    r = subprocess.run(args, cwd=GOLDEN, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_occ_ri_k_step_h2o():
    out = run([os.path.join(BIN, "occ_ri_k_step"), "24", "113", "5", "322", "3c_ints_PQm12_h2o.data",
               "occ_mo_h2o.data", "v2m_key_h2o.data", "3", "1e-20"])
    # the reference's own lines (occ_ri_k_step.cpp:70-130), formats unchanged
    assert re.search(r"L2 norm of output = \d\.\d{6}e\+01\n", out), out
    assert len(re.findall(r"  \[\d\] \d+\.\d{4} s  \(\d+\.\d{4} GFLOPs\)", out)) == 6, out
    dense, sparse = (float(x) for x in re.search(r"full-precision L2 norms: dense (\S+) sparse (\S+)", out).groups())
    assert abs(dense - 10.415071322740426) < 1e-12, out
    assert abs(sparse - 10.415071322740426) < 1e-12, out
    for v in re.findall(r"relative RMSD = (\S+)", out):
        assert float(v) < 1e-13, out


def test_3c_2c_transform_h2o(tmp_path):
    res = tmp_path / "pq.data"
    out = run([os.path.join(BIN, "3c_2c_transform"), "24", "113", "322", "3c_ints_h2o.data",
               "2c_ints_inv05_h2o.data", "3", "1e-20", "-o", str(res)])
    got = np.fromfile(res).reshape(322, 113, order="F")
    assert rel_fro(got, load_golden("3c_ints_PQm12_h2o.data", (322, 113))) < 1e-13
    rm = float(re.search(r"relative RMSD = (\S+)", out).group(1))
    assert rm < 1e-13, out  # sparse vs dense device products, FP64 rounding only


def test_strassen_bench_synthetic():
    out = run([os.path.join(BIN, "strassen_bench"), "64", "512", "1e-10", "1"])
    errs = [float(v) for v in re.findall(r"relative RMSD = (\S+)", out)]
    assert len(errs) == 2 and max(errs) < 1e-12, out
