"""The drop-in `swap` driver (swap-assembler_amd/bin/swap) run as a program, in
each of the reference's modes, against what the unmodified reference wrote as
one rank for the same input (tests/golden/make_golden_np1.py): the set of files
in the output directory, their bytes, and the labels of logtime.txt
(graph.cpp:3458-3694; the times themselves differ from run to run)."""
import os
import re
import subprocess
import tempfile

import pytest

from test_oracle import NP1, np1_input, same_as_np1

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SWAP = os.path.join(REPO, "swap-assembler_amd", "bin", "swap")

# logtime.txt of the reference in each mode, numbers replaced by N (taken from
# its own runs): lines written through MPIEnviroment::print carry the process
# and memory prefix and end in a blank line (mympi.cpp:30-40)
_HEAD = [
    "Proc:0 [Mem=N MB] ->Time spend in Geting Reads N", "",
    "Proc:0 [Mem=N MB] ->Time spend in I/O part in MPIcontrol N", "",
    "Proc:0 [Mem=N MB] ->Time spend in locate function N", "",
    "Proc:0 [Mem=N MB] ->Time spend in read function N", "",
    "Proc:0 [Mem=N MB] ->Time spend in Constructing Graph N (cutTimetot=N, communication time=N, "
    "communicationtimetot=N)", "",
    "Proc:0 [Mem=N MB] ->Time spend in cutoffGraph N", "",
    "Proc:0 [Mem=N MB] ->Time spend in Constructing DistGraph N", "",
    "Proc:0 [Mem=N MB] ->Time spend in Simplify DistGraph N (workTime=N(N), totalTime=N(N))", "",
]
LOG = {
    "default": _HEAD + ["Time spend in printDistGraph N", "Time spend in reading DistGraph  N",
                        "Time spend in Contig Extension N", "Time spend in All step N"],
    "s": _HEAD + ["Time spend in printDistGraph N"],
    "p": _HEAD + ["Time spend in printDistGraph N", "Time spend in All step N"],
}

# mode flags -> files the reference leaves in the output directory, and the
# fixture each one is compared with (None: compared otherwise)
MODES = {
    "default": ([], {"noCEcontigs.fasta": "noce", "CEContig.fasta": "cecontig", "logtime.txt": None}, "default"),
    "d_j": (["-d", "-j"], {"contigGraph.txt": "contigGraph", "noCEcontigs.fasta": "noce",
                           "CEContig.fasta": "cecontig", "JungGraph_arc.txt": "jungarc",
                           "JungGraph_mul.txt": "jungmul", "logtime.txt": None}, "default"),
    "s": (["-s"], {"kmerGraph.txt": "kmerGraph", "logtime.txt": None}, "s"),
    "s_d": (["-s", "-d"], {"kmerGraph.txt": "kmerGraph", "logtime.txt": None}, "s"),  # -d is ignored under -s
    "p": (["-p"], {"noCEcontigs.fasta": "perf", "logtime.txt": None}, "p"),
    "s_p": (["-s", "-p"], {"noCEcontigs.fasta": None, "logtime.txt": None}, "p"),  # unmerged table: empty
}

CASES = ["foldlock_k19_c1", "cycles_k21_c1", "mixed_k31_c2"]


def run_swap(args, cwd):
    print("swap %s" % " ".join(args), flush=True)
    return subprocess.run([SWAP] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("name", CASES)
def test_driver_modes_match_single_rank_reference(name, mode):
    flags, files, log = MODES[mode]
    case = NP1[name]
    with tempfile.TemporaryDirectory() as tmp:
        fa = os.path.join(tmp, "reads.fa")
        with open(fa, "wb") as f:
            f.write(np1_input(name))
        out = os.path.join(tmp, "out")
        r = run_swap(["-k", str(case["k"]), "-c", str(case["c"]), "-i", fa, "-o", out] + flags, tmp)
        assert r.returncode == 0, r.stdout + r.stderr
        assert sorted(os.listdir(out)) == sorted(files)
        for fname, suffix in files.items():
            data = open(os.path.join(out, fname), "rb").read()
            if suffix is not None:
                assert same_as_np1(name, suffix, data), fname
        if mode == "s_p":
            assert open(os.path.join(out, "noCEcontigs.fasta"), "rb").read() == b""
        text = open(os.path.join(out, "logtime.txt")).read()
        masked = re.sub(r"(?<=[ =(])\d+(\.\d+)?", "N", text).split("\n")
        assert masked[-1] == ""
        assert masked[:-1] == LOG[log]
