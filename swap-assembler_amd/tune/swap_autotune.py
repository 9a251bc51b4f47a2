"""Autotuner of the k-mer graph path's performance parameters.

Run on the GPU (it times builds and merges):

    python swap-assembler_amd/tune/swap_autotune.py [--reps 7] [--out swap-assembler_amd/tune]

For each input-size class it takes representative seeded inputs and tunes,
by coordinate descent from the library's defaults, the parameters of
swap_tuning (include/swap_hip.h): the records per count partition (the load of
the on-chip counting table), the records of a shared table fill, the load
factor of the survivor index (0.50-0.85), the ruler stride of the list ranking,
the workgroups per CU of the k-mer passes and the keys per bin of the
visit-order sort.  A candidate is timed as the median of --reps builds + merges of the input already in device memory (the
two small classes: as the time per sample of a server batch, in a child
process that reads the candidate from a one-row table), and is only accepted
if every output file is byte-identical to the defaults' (the parameters change
timings, never results).

Writes <out>/swap_tune.txt, the table the library reads (one row per class:
the largest input in bytes it covers, then the six parameters), and
<out>/swap_tune_report.json with every timing.
"""
import argparse
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import swapamd  # noqa: E402

DEFAULTS = {"part_recs": 8192, "fill_recs": 4096, "index_load": 67, "ruler_stride": 32, "occ_groups": 1,
            "order_bin_keys": 2048}
CANDIDATES = [
    ("part_recs", [4096, 6144, 8192, 12288]),
    ("fill_recs", [2048, 4096, 8192]),
    ("index_load", [50, 60, 67, 75, 85]),
    ("ruler_stride", [8, 16, 32, 64, 128, 256]),
    ("occ_groups", [1, 2]),
    ("order_bin_keys", [256, 512, 1024, 2048, 4096, 8192]),
]
# size classes: (name, largest input in bytes, input generator arguments, k, c,
# objective).  Inputs of the two small classes come in batches (the many-sample
# workload's samples average ~15 MB), so they are tuned for the samples/s of a
# server assembling 48 of them on 12 slots; the larger ones for the time of one
# build + merge.
CLASSES = [
    ("tiny", 2 << 20, dict(genome_len=20000, read_len=150, coverage=60.0, error_rate=0.005, seed=901), 31, 4,
     "batch"),
    ("small", 16 << 20, dict(genome_len=150000, read_len=150, coverage=50.0, error_rate=0.005, seed=902), 31, 4,
     "batch"),
    ("medium", 128 << 20, dict(genome_len=1000000, read_len=100, coverage=30.0, error_rate=0.005, seed=903), 31, 3,
     "single"),
    ("large", (1 << 64) - 1, dict(genome_len=4600000, read_len=100, coverage=50.0, error_rate=0.01, seed=904), 31, 5,
     "single"),
]
BATCH_JOBS, BATCH_SLOTS = 48, 12
FILES = (swapamd.SWAP_FILE_CONTIG_GRAPH, swapamd.SWAP_FILE_NOCE)


def run(a, buf, reps):
    times = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        a.build_device(buf.ptr.value, buf.n)
        a.merge()
        times.append(time.perf_counter() - t0)
    digest = hashlib.sha256(b"".join(a.file_bytes(w) for w in FILES)).hexdigest()
    return statistics.median(times[1:]) * 1e3, digest


def batch_run(bound, gen, k, c, params, reps):
    """ms per sample of a server batch, with params in force for inputs up to
    bound bytes: a child process reads a one-row table (the library reads its
    table once per process)."""
    table = os.path.join(os.environ.get("TMPDIR", "/tmp"), "swap_tune_candidate.txt")
    with open(table, "w") as f:
        f.write("%d %d %d %d %d %d %d\n" % (bound, params["part_recs"], params["fill_recs"], params["index_load"],
                                             params["ruler_stride"], params["occ_groups"], params["order_bin_keys"]))
    env = dict(os.environ, SWAP_TUNE_FILE=table)
    arg = json.dumps({"gen": gen, "k": k, "c": c, "reps": reps})
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--batch-child", arg], env=env,
                       capture_output=True, text=True, timeout=600)
    if r.returncode:
        raise SystemExit("batch child failed: " + r.stderr[-2000:])
    out = json.loads(r.stdout.strip().splitlines()[-1])
    return out["ms_per_sample"], out["digest"]


def batch_child(arg):
    a = json.loads(arg)
    samples = [swapamd.generate_params(**dict(a["gen"], seed=a["gen"]["seed"] * 1000 + j)) for j in range(BATCH_JOBS)]
    bufs = [swapamd.DeviceBuffer(fa) for fa in samples]
    srv = swapamd.Server(slots=BATCH_SLOTS)
    jobs = [swapamd.Server.job(len(fa), a["k"], a["c"], outputs=FILES, device_ptr=b.ptr.value)
            for fa, b in zip(samples, bufs)]
    times, digest = [], None
    for _ in range(a["reps"] + 1):
        t0 = time.perf_counter()
        ids = srv.submit_batch(jobs)
        srv.wait_all()
        times.append(time.perf_counter() - t0)
        h = hashlib.sha256()
        for j in ids:
            for w in FILES:
                h.update(srv.output(j, w))
            srv.release(j)
        digest = h.hexdigest()
    srv.close()
    for b in bufs:
        b.close()
    print(json.dumps({"ms_per_sample": statistics.median(times[1:]) * 1e3 / BATCH_JOBS, "digest": digest}))


def tune_class(name, bound, gen, k, c, objective, reps, report):
    if objective == "batch":
        fa = swapamd.generate_params(**dict(gen, seed=gen["seed"] * 1000))
        measure = lambda p: batch_run(bound, gen, k, c, p, max(2, reps // 3))  # noqa: E731
    else:
        fa = swapamd.generate_params(**gen)
        buf = swapamd.DeviceBuffer(fa)
        a = swapamd.Assembler(k, c)

        def measure(p):
            a.set_tuning(p)
            return run(a, buf, reps)
    best = dict(DEFAULTS)
    best_ms, want = measure(best)
    entry = {"bytes": len(fa), "k": k, "c": c, "objective": objective, "default_ms": best_ms, "trials": []}
    print("%s: %d bytes, defaults %.3f ms" % (name, len(fa), best_ms), flush=True)
    for pname, values in CANDIDATES:
        for v in values:
            if v == best[pname]:
                continue
            cand = dict(best)
            cand[pname] = v
            ms, got = measure(cand)
            if got != want:
                raise SystemExit("%s: %s=%d changed the output files" % (name, pname, v))
            entry["trials"].append({"params": cand, "ms": ms})
            print("  %s=%d: %.3f ms" % (pname, v, ms), flush=True)
            if ms < best_ms * 0.98:  # keep a change only for a clear gain
                best, best_ms = cand, ms
        # re-time the incumbent so later comparisons use a fresh figure
        best_ms, _ = measure(best)
    # changes kept for a gain that noise made: the defaults, timed afresh, win
    # unless the incumbent still beats them
    if best != DEFAULTS:
        default_ms, _ = measure(dict(DEFAULTS))
        entry["default_retimed_ms"] = default_ms
        if default_ms <= best_ms:
            best, best_ms = dict(DEFAULTS), default_ms
    entry["best"] = best
    entry["best_ms"] = best_ms
    entry["gain"] = entry["default_ms"] / best_ms
    report[name] = entry
    print("%s: best %s %.3f ms (%.2fx over defaults)" % (name, best, best_ms, entry["gain"]), flush=True)
    if objective != "batch":
        a.close()
        buf.close()
    return bound, best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=7)
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--batch-child", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.batch_child:
        return batch_child(args.batch_child)
    report = {}
    rows = [tune_class(n, b, g, k, c, o, args.reps, report) for n, b, g, k, c, o in CLASSES]
    os.makedirs(args.out, exist_ok=True)
    with open(os.path.join(args.out, "swap_tune.txt"), "w") as f:
        f.write("# Written by swap_autotune.py on %s: per input-size class, the largest input in\n"
                "# bytes, then part_recs fill_recs index_load ruler_stride occ_groups order_bin_keys.\n" %
                time.strftime("%Y-%m-%d"))
        for bound, t in rows:
            f.write("%d %d %d %d %d %d %d\n" % (bound, t["part_recs"], t["fill_recs"], t["index_load"],
                                                 t["ruler_stride"], t["occ_groups"], t["order_bin_keys"]))
    with open(os.path.join(args.out, "swap_tune_report.json"), "w") as f:
        json.dump(report, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
