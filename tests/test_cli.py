"""The drop-in `swap` driver's argument checks, which end before any device
work (no GPU needed): an existing output directory and a k outside the
reference's range (sequence.cpp:44, 128-133) are refused with the reference's
messages, and nothing is written."""
import os
import subprocess
import tempfile

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SWAP = os.path.join(REPO, "swap-assembler_amd", "bin", "swap")


def run_swap(args, cwd):
    return subprocess.run([SWAP] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_existing_output_directory_is_refused():
    with tempfile.TemporaryDirectory() as tmp:
        fa = os.path.join(tmp, "reads.fa")
        with open(fa, "w") as f:
            f.write(">r0\nACGTACGTACGTACGTACGTACGTACGTACGTACGT\n")
        out = os.path.join(tmp, "out")
        os.mkdir(out)
        r = run_swap(["-k", "21", "-c", "2", "-i", fa, "-o", out], tmp)
        assert r.returncode != 0
        assert "Error: Creating directory %s" % out in r.stdout
        assert os.listdir(out) == []


def test_k_outside_the_reference_range_is_refused():
    with tempfile.TemporaryDirectory() as tmp:
        for k in ("17", "20", "33"):
            r = run_swap(["-k", k, "-c", "2", "-i", "x.fa", "-o", os.path.join(tmp, "o")], tmp)
            assert r.returncode != 0
            assert "odd number between 18 and 32" in r.stdout
            assert not os.path.exists(os.path.join(tmp, "o"))


def test_missing_paths_print_the_usage():
    with tempfile.TemporaryDirectory() as tmp:
        r = run_swap(["-k", "21"], tmp)
        assert r.returncode != 0
        assert r.stdout.startswith("Usage: ")
