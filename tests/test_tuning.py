"""The tuning table (swap-assembler_amd/tune/swap_tune.txt, written by the
autotuner) as the library reads it: row choice by input size, clamping, the
built-in defaults without a table.  No device needed: swap_tuning_for only
reads the table.  The library reads its table once per process, so each case
runs in a child process with SWAP_TUNE_FILE set."""
import json
import os
import subprocess
import sys
import tempfile

import swapamd

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "swap-assembler_amd")
DEFAULT = {"part_recs": 8192, "fill_recs": 4096, "index_load": 67, "ruler_stride": 32, "occ_groups": 1,
           "order_bin_keys": 2048}


def lookup(table_text, sizes):
    with tempfile.TemporaryDirectory() as tmp:
        env = dict(os.environ)
        if table_text is None:
            env["SWAP_TUNE_FILE"] = os.path.join(tmp, "missing.txt")
        else:
            path = os.path.join(tmp, "t.txt")
            with open(path, "w") as f:
                f.write(table_text)
            env["SWAP_TUNE_FILE"] = path
        code = ("import sys, json; sys.path.insert(0, %r); import swapamd; "
                "print(json.dumps([swapamd.tuning_for(n) for n in %r]))" % (PKG, list(sizes)))
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        return json.loads(r.stdout)


def test_rows_are_chosen_by_input_size():
    table = ("# comment line\n"
             "1000 4096 2048 50 8 2 1024   # first class\n"
             "100000 12288 8192 85 64 1\n")  # no order_bin_keys column: its default
    got = lookup(table, [1, 1000, 1001, 100000, 10 ** 9])
    first = {"part_recs": 4096, "fill_recs": 2048, "index_load": 50, "ruler_stride": 8, "occ_groups": 2,
             "order_bin_keys": 1024}
    second = {"part_recs": 12288, "fill_recs": 8192, "index_load": 85, "ruler_stride": 64, "occ_groups": 1,
              "order_bin_keys": 2048}
    assert got == [first, first, second, second, second]  # above every bound: the last row


def test_values_are_clamped_and_bad_lines_skipped():
    got = lookup("not a row\n5000 1 1 99 1000 9 1\n6000 1 1 1 1 1 4294967295\n", [10, 5500])
    assert got == [{"part_recs": 1024, "fill_recs": 512, "index_load": 85, "ruler_stride": 256, "occ_groups": 4,
                    "order_bin_keys": 256},
                   {"part_recs": 1024, "fill_recs": 512, "index_load": 50, "ruler_stride": 4, "occ_groups": 1,
                    "order_bin_keys": 1 << 30}]


def test_defaults_without_a_table():
    assert lookup(None, [0, 10 ** 12]) == [DEFAULT, DEFAULT]


def test_committed_table_is_read():
    rows = []
    with open(os.path.join(PKG, "tune", "swap_tune.txt")) as f:
        for line in f:
            line = line.split("#")[0].split()
            if len(line) == 7:
                rows.append([int(x) for x in line])
    assert len(rows) >= 3 and rows == sorted(rows)
    for bound, pr, fr, il, rs, og, ob in rows:
        want = {"part_recs": pr, "fill_recs": fr, "index_load": il, "ruler_stride": rs, "occ_groups": og,
                "order_bin_keys": ob}
        assert swapamd.tuning_for(min(bound, 2 ** 63)) == want
    report = json.load(open(os.path.join(PKG, "tune", "swap_tune_report.json")))
    assert set(report) == {"tiny", "small", "medium", "large"}
    assert all(r["best_ms"] > 0 and r["trials"] for r in report.values())
