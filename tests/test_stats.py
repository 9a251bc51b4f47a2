"""The stats tool (swap_stats) against the reference's own stats tool: its
output on each committed input (tests/golden/stats_cases.json, written by the
reference build) byte for byte, including the short-contig leak."""
import json
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "swap-assembler_amd", "bin", "swap_stats")
CASES = json.load(open(os.path.join(REPO, "tests", "golden", "stats_cases.json")))


def run(tmp_path, text, ref_len):
    path = tmp_path / "c.fa"
    path.write_bytes(text.encode())
    p = subprocess.run([EXE, str(path), str(ref_len)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    return p.stdout


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_matches_reference_stats(tmp_path, case):
    assert run(tmp_path, case["fasta"], case["ref_len"]) == case["stdout"]


def test_header_only_noce_file_has_no_contig(tmp_path):
    # noCEcontigs.fasta keeps each sequence in its header line: nothing is recorded
    text = ">proc0: " + "A" * 500 + "\n>proc0: " + "C" * 300 + "\n"
    out = run(tmp_path, text, 1000)
    assert out.splitlines()[0] == "Input 0 contigs"
    assert out.splitlines()[-1] == "Max = 0, mean = 0"
