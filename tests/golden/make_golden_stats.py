"""Writes tests/golden/stats_cases.json: inputs and the output of the reference's
stats tool (stats.cpp, built from /root/reference into oracle/_ref/stats by
`make -C oracle ref`) on each.  Run here, where the reference exists; the GPU
box only reads the JSON."""
import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(REPO, "oracle", "_ref", "stats")


def seq(rng, n):
    return "".join(rng.choice("ACGT") for _ in range(n))


def cases():
    rng = random.Random(7)
    out = []
    # CEContig-like: one header line + one sequence line per contig, lengths
    # around the 200 threshold (short ones leak into the next contig)
    lens = [450, 120, 199, 200, 201, 1000, 80, 80, 80, 650, 3000, 150, 220]
    text = "".join(">contig_%d\n%s\n" % (i, seq(rng, n)) for i, n in enumerate(lens))
    out.append(("threshold_and_leak", text, 5000))
    # sequences over several lines; equal lengths (ties)
    parts = []
    for i in range(9):
        s = seq(rng, 300 + 100 * (i % 3))
        parts.append(">c%d\n" % i + "\n".join(s[j:j + 70] for j in range(0, len(s), 70)) + "\n")
    out.append(("multiline_ties", "".join(parts), 2000))
    # CR line endings count as bases; last record without a final newline
    text = ">a\r\n%s\r\n>b\r\n%s" % (seq(rng, 260), seq(rng, 900))
    out.append(("crlf_no_final_newline", text, 1000))
    # genome length above the contig total: N50_abs stays 0
    text = "".join(">x%d\n%s\n" % (i, seq(rng, 210 + 37 * i)) for i in range(15))
    out.append(("abs_not_reached", text, 10 ** 7))
    # a single long contig
    out.append(("single", ">only\n%s\n" % seq(rng, 5000), 4000))
    return out


def main():
    recs = []
    with tempfile.TemporaryDirectory() as d:
        for name, text, ref_len in cases():
            path = os.path.join(d, name + ".fa")
            with open(path, "w") as f:
                f.write(text)
            p = subprocess.run([EXE, path, str(ref_len)], capture_output=True, text=True, timeout=60)
            assert p.returncode == 0, (name, p.stderr)
            recs.append({"name": name, "fasta": text, "ref_len": ref_len, "stdout": p.stdout})
    with open(os.path.join(HERE, "stats_cases.json"), "w") as f:
        json.dump(recs, f, indent=1)
    print("wrote %d cases" % len(recs))


if __name__ == "__main__":
    main()
