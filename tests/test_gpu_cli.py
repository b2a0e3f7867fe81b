"""The G2H command line (main.cu:244-307 argument list) on the reference's
13-vertex fixture: runs every phase and appends the reference's CSV files."""
import os
import shutil
import subprocess

import pytest

import cases
import h2h

pytestmark = pytest.mark.gpu
G2H = os.path.join(os.path.dirname(h2h.LIB_PATH), "G2H")


@pytest.mark.parametrize("args", [["13", "4", "1", "20", "1", "1", "2", "1"],
                                  ["13", "4", "2", "8", "1", "0", "1", "1"],
                                  ["13", "3", "1", "4", "3", "1", "3", "0"]])
def test_g2h_runs_and_writes_csvs(tmp_path, args):
    os.makedirs(tmp_path / "Graph/USA")
    os.makedirs(tmp_path / "Partition/13")
    for f in ("13Edge.txt", "13Node.txt"):
        shutil.copy(os.path.join(cases.REF13, f), tmp_path / "Graph/USA" / f)
    for m in cases.REF13_METHODS:
        for suf in ("ID.txt", "Cut.txt"):
            shutil.copy(os.path.join(cases.REF13, m + suf), tmp_path / "Partition/13" / (m + suf))
    r = subprocess.run([G2H] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = (tmp_path / "ConstructInfo/13/constructInfo.csv").read_text().splitlines()
    assert info[0].startswith("graphName,PartitionTreeHeight,changeHeight") and len(info[0].split(",")) == 31
    assert info[1].split(",")[0] == "13" and len(info[1].split(",")) == 31
    rnd = (tmp_path / "QueryResult/13/random.csv").read_text().splitlines()
    assert len(rnd[0].split(",")) == 23 and len(rnd[1].split(",")) == 23
    assert (tmp_path / "ConstructInfo/13/contract.csv").exists()


def test_g2h_wrong_argc_exits_zero(tmp_path):
    r = subprocess.run([G2H, "13"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
