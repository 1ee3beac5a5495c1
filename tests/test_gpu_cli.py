"""The reference-compatible CLI driver writes the reference CSV schema with costs equal
to the CPU oracle's."""
import csv
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from gpu4gst_1_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu4gst-1_amd", "gst_cli")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,D", [(2, None), (3, None), (2, 3), (3, 3)])
def test_cli_csv(tmp_path, rmat, rmat_inst, mode, D):
    datasets.write_workload(rmat, str(tmp_path))
    args = [CLI, str(mode), str(tmp_path) + "/", "rmat32k", "3"] + ([str(D)] if D is not None else []) + ["0", "9"]
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    tag = "GPU1" if mode == 2 else "GPU2"
    kind = "Hop" if D is not None else "nonHop"
    name = f"exp_{tag}_{kind}_rmat32k_T3" + (f"_D{D}" if D is not None else "") + "_0-9.csv"
    rows = [row for row in csv.reader(open(tmp_path / "result" / name)) if row]
    assert rows[0][0] == "task_ID" and rows[0][-2:] == ["counts", "process_num"]
    assert len(rows) == 11
    for row in rows[1:]:
        i = int(row[0])
        q = rmat.queries[3][i]
        assert row[1].split() == [str(x) for x in q]
        o = pyoracle.prunedpp(rmat_inst, q) if D is None else pyoracle.dprunedpp(rmat_inst, q, D)
        assert int(row[3]) == o["tree_cost"], (row, o["tree_cost"])
        assert int(row[5]) > 0 and int(row[6]) > 0
