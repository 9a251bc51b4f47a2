"""Benchmark: k-mer graph build + k-molecule merge on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bacterial]
                    [--hash-order] [--no-cpu-baseline]

A step is one pass of the path over one batch: swap_build_device on FASTA bytes
already resident in HBM, then swap_merge.  Default workload: the bacterial-size
configuration (seeded random 4.6 Mbp genome at 50% GC, 100 bp reads from both
strands to 50x = 2.3M reads, 1% substitutions, k=31, c=5), synthetic data from
the project's generator.  Prints one JSON line.

roofline: the dominant device kernel of the step, its algorithmic bytes per
launch (the words it must read and write once) over its device time measured
with HIP events recorded on the library's own stream around the launch.
cpu_baseline: the reference itself (built from its sources into oracle/_ref,
where it was built) under mpirun on the host's cores, wall clock on the first
quarter of the reads; without it, the single-threaded CPU oracle on a bounded
prefix.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "swap-assembler_amd"))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

# phase index -> kernel name (see swap_stats.phase_ms)
PHASES = {0: "record_scan", 1: "pack_2bit", 2: "partition_hist", 3: "partition_level1", 4: "count_lds_cutoff",
          5: "count_global_overflow", 6: "growth_rank_select", 7: "partition_level2", 8: "survivor_index",
          9: "list_ranking", 10: "kept_edges", 11: "edge_strings", 12: "cartesian_trees", 13: "visit_order",
          14: "classify", 15: "survivor_keys"}


# phase index -> device kernel of that phase whose hardware byte counters are kept
# (the bacterial workload: one partition per count table fill, 2^7 level-2 sub-buckets)
# in profiles/traffic.json (written from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)
PHASE_KERNEL = {4: "swapk::k_count_lds<%s, false, false>", 3: "swapk::k_part1<%s>", 7: "swapk::k_part2<%s, 128, false, 4096, 512>",
                2: "swapk::k_part_hist", 1: "swapk::k_pack"}


def measured_traffic(phase, ref_order, workload):
    """HBM bytes per launch of the phase's kernel from the committed counter summary, or None."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    if phase not in PHASE_KERNEL or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("workload") != workload:
        return None
    name = PHASE_KERNEL[phase]
    if "%s" in name:
        # level 1 stages rounds of 11 records (reference order) or 16
        name = name % (("true, 11" if ref_order else "false, 16") if phase == 3 else ("true" if ref_order else "false"))
    k = t.get("kernels", {}).get(name)
    return k["bytes"] if k else None


def algorithmic_bytes(phase, st, words, ref_order):
    """Bytes a phase's kernel must read and write once, per launch.

    Records are 8 bytes (key bits + edge mask), plus a 4-byte input index in
    reference order; a packed word comes with an 8-byte occurrence descriptor.
    """
    N, S, D, R = st["kmers"], st["survivors"], st["distinct"], st["reads"]
    occ = 4 if ref_order else 0
    if phase == 3:   # level-1 partition: packed words + descriptors in, records out
        return 16 * words + N * (8 + occ)
    if phase == 7:   # level-2 partition: records in, records out
        return 2 * N * (8 + occ)
    if phase == 4:   # LDS count: records in, survivors out, first occurrence of every distinct key
        return N * (8 + occ) + S * (8 + 8 + occ) + (4 * D if ref_order else 0)
    if phase == 2:   # partition histogram: packed words + descriptors in
        return 16 * words
    if phase == 1:   # 2-bit pack: sequence bytes in, packed words + descriptors out
        return 100 * R + 16 * words
    return 0


BUILD_PHASES = (0, 1, 2, 3, 4, 5, 6, 7, 15)


def build_bytes(st, fasta_bytes, read_len, k):
    """Table-build algorithmic bytes F + 8N + 8E + 48D + 16S, E = 2(L-k) per read."""
    N, S, D, R = st["kmers"], st["survivors"], st["distinct"], st["reads"]
    E = 2 * R * max(read_len - k, 0)
    return fasta_bytes + 8 * N + 8 * E + 48 * D + 16 * S


def path_bytes(st, fasta_bytes, read_len, k, edge_bases):
    """Whole-path algorithmic bytes of the build+merge (word-granular unique traffic):
    F + 8N + 8E + 48D + 16S (build) + 16S + edge bases (merge), E = 2(L-k) per read."""
    N, S, D, R = st["kmers"], st["survivors"], st["distinct"], st["reads"]
    E = 2 * R * max(read_len - k, 0)
    return fasta_bytes + 8 * N + 8 * E + 48 * D + 16 * S + 16 * S + edge_bases


def sample_prefix(fa: bytes, frac: float) -> bytes:
    """The first records of the FASTA, about frac of its bytes, ending on a record boundary."""
    cut = int(len(fa) * frac)
    pos = fa.rfind(b"\n>", 0, cut)
    return fa[:pos + 1] if pos > 0 else fa


def count_kmers(fa: bytes, k: int) -> int:
    """Input k-mer occurrences of a one-line-per-sequence FASTA (records shorter than k give none)."""
    lines = fa.split(b"\n")
    n = 0
    for i in range(1, len(lines), 2):
        n += max(0, len(lines[i]) - k + 1)
    return n


def find_mpirun():
    for cand in (shutil.which("mpirun"), "/opt/conda/bin/mpirun"):
        if cand and os.path.exists(cand):
            return cand
    return None


def host_cores():
    """CPU cores this process may run on (the GPU box gives a share of its host)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def reference_baseline(fa: bytes, k: int, c: int, frac=0.25):
    """The reference itself (built from its sources into oracle/_ref) under mpirun
    on the host cores, on the first `frac` of the reads, in its -p mode (the
    process ends after the merge; graph.cpp:3591-3611).

    value: input k-mers per second of wall-clock time around the whole mpirun
    (process start, MPI start-up and file reading included).  Also reported: the
    sum of the reference's own build+merge phase clocks from logtime.txt
    (Constructing Graph + cutoffGraph + Constructing DistGraph + Simplify
    DistGraph: rank-0 CPU time after barriers).  One rank per core, at least
    4 (at 2 and 3 ranks the reference's merge overruns an MPI_Status array,
    graph.cpp:1215/1230).  Returns None if the binary or mpirun is absent."""
    exe = os.path.join(REPO, "oracle", "_ref", "swap")
    mpirun = find_mpirun()
    if not os.path.exists(exe) or not mpirun:
        return None
    nprocs = max(4, min(16, host_cores()))
    tmpdir = tempfile.mkdtemp(prefix="ref_base_")
    sample = sample_prefix(fa, frac)
    path = os.path.join(tmpdir, "s.fa")
    with open(path, "wb") as f:
        f.write(sample)
    out = os.path.join(tmpdir, "out")
    t0 = time.perf_counter()
    p = subprocess.run([mpirun, "-np", str(nprocs), exe, "-k", str(k), "-c", str(c), "-i", path, "-o", out, "-p"],
                       capture_output=True, text=True, timeout=600)
    wall = time.perf_counter() - t0
    log = os.path.join(out, "logtime.txt")
    if p.returncode != 0 or not os.path.exists(log):
        print("reference baseline failed (rc %d): %s" % (p.returncode, p.stderr[-400:]), file=sys.stderr)
        return None
    secs = 0.0
    for line in open(log):
        m = re.search(r"Time spend in (Constructing Graph|cutoffGraph|Constructing DistGraph|Simplify DistGraph) "
                      r"([0-9.]+)", line)
        if m:
            secs += float(m.group(2))
    kmers = count_kmers(sample, k)
    shutil.rmtree(tmpdir, ignore_errors=True)
    return {"value": kmers / wall, "unit": "input k-mers/s", "cores": nprocs, "kind": "reference",
            "host_cores": host_cores(), "wall_s": wall, "phase_clock_s": secs,
            "phase_clock_value": kmers / secs if secs > 0 else None,
            "sample": "first %.0f%% of the reads (%d input k-mers); reference swap -p under mpirun -np %d, "
                      "wall clock around mpirun %.1f s (its own build+merge phase clocks: %.1f s)" % (
                          100 * len(sample) / len(fa), kmers, nprocs, wall, secs)}


def cpu_baseline(fa: bytes, k: int, c: int, target_s=15.0):
    """Times the oracle CLI on growing prefixes of the reads until ~target_s."""
    exe = os.path.join(REPO, "oracle", "swap_oracle")
    if not os.path.exists(exe):
        return None
    tmpdir = tempfile.mkdtemp(prefix="cpu_base_")
    frac = 0.02
    res = None
    while True:
        cut = int(len(fa) * frac)
        # end the sample on a record boundary
        pos = fa.rfind(b"\n>", 0, cut)
        sample = fa[:pos + 1] if pos > 0 else fa
        path = os.path.join(tmpdir, "s.fa")
        with open(path, "wb") as f:
            f.write(sample)
        out = os.path.join(tmpdir, "o%d" % int(frac * 1000))
        t0 = time.time()
        p = subprocess.run([exe, "-k", str(k), "-c", str(c), "-i", path, "-o", out], capture_output=True, text=True,
                           timeout=600)
        wall = time.time() - t0
        info = json.loads(p.stdout.strip().splitlines()[-1])
        res = {"value": info["kmers"] / info["seconds"], "unit": "input k-mers/s", "cores": 1, "kind": "port",
               "sample": "first %.1f%% of the reads (%d input k-mers), oracle build+cutoff+merge, %.1f s" % (
                   100 * len(sample) / len(fa), info["kmers"], info["seconds"])}
        if wall >= target_s / 3 or frac >= 1.0:
            break
        frac = min(1.0, frac * max(2.0, target_s / max(wall, 0.05) * 0.8))
    return res


def copy_probe(nbytes):
    """One hipMemcpy device-to-device of nbytes: reads and writes nbytes once."""
    import ctypes
    import swapamd
    H = swapamd.hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if H.hipMalloc(ctypes.byref(a), nbytes) or H.hipMalloc(ctypes.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    if H.hipMemset(a, 1, nbytes) or H.hipMemcpy(b, a, nbytes, 3):  # 3 = hipMemcpyDeviceToDevice
        raise RuntimeError("copy probe failed")
    H.hipDeviceSynchronize()
    H.hipFree(a)
    H.hipFree(b)


def batched_vs_per_sample(n_samples, slots, group_mib=256, fused_slots=8, warmup=True, reps=3):
    """Metric 2: the many-sample workload (seeded genomes log-uniform in
    [20 kbp, 2 Mbp], 150 bp reads at 20-80x, 0.5% errors, k=31, c=4), inputs
    resident in HBM, noCEcontigs.fasta produced for every sample.  Plain path:
    one context assembling the samples one after another.  Batched path: one
    server's fused submission (groups of up to group_mib MiB of FASTA built by
    one batch build each, the samples merged by `slots` concurrent contexts);
    beside it, the same server with every sample built by its own slot.
    Each path is timed over `reps` passes (the median pass is reported).
    Returns samples/s of each and whether every sample's output is
    byte-identical between them."""
    import concurrent.futures as cf
    import swapamd
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        gen = list(ex.map(lambda i: swapamd.generate("batch", i), range(n_samples)))
    fas = [g[0] for g in gen]
    k, c = gen[0][1], gen[0][2]
    bufs = [swapamd.DeviceBuffer(fa) for fa in fas]
    order = sorted(range(n_samples), key=lambda i: -len(fas[i]))
    noce = swapamd.SWAP_FILE_NOCE

    a = swapamd.Assembler(k, c)
    if warmup:  # grow the context's buffers to the largest sample once
        a.build_device(bufs[order[0]].ptr.value, len(fas[order[0]]))
        a.merge()
    # each path timed over `reps` passes; the median pass is reported
    t_runs = []
    for _ in range(reps):
        swapamd.hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        plain, kmers = [None] * n_samples, 0
        for i in range(n_samples):
            a.build_device(bufs[i].ptr.value, len(fas[i]))
            a.merge()
            plain[i] = a.file_bytes(noce)
            kmers += a.stats()["kmers"]
        swapamd.hip().hipDeviceSynchronize()
        t_runs.append(time.perf_counter() - t0)
    t_plain = sorted(t_runs)[len(t_runs) // 2]
    a.close()

    srv = swapamd.Server(slots=slots)
    if warmup:  # grow every slot's buffers with the largest samples once
        ids = srv.submit_batch([swapamd.Server.job(len(fas[i]), k, c, outputs=(noce,), device_ptr=bufs[i].ptr.value)
                                for i in order[:slots]])
        srv.wait_all()
        for j in ids:
            srv.release(j)
    jobs = [swapamd.Server.job(len(fas[i]), k, c, outputs=(noce,), device_ptr=bufs[i].ptr.value)
            for i in range(n_samples)]
    t_runs = []
    for r in range(reps):
        swapamd.hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        ids = srv.submit_batch(jobs)
        srv.wait_all()
        t_runs.append(time.perf_counter() - t0)
        if r + 1 < reps:
            for j in ids:
                srv.release(j)
    t_batch = sorted(t_runs)[len(t_runs) // 2]
    same = all(srv.output(ids[i], noce) == plain[i] for i in range(n_samples))
    srv.close()

    # fused: the same jobs built in groups by batch builds
    srv = swapamd.Server(slots=fused_slots)
    gb = group_mib << 20
    if warmup:  # grow the batch contexts and the slots once
        ids = srv.submit_fused(jobs, group_bytes=gb)
        srv.wait_all()
        for j in ids:
            srv.release(j)
    t_runs = []
    for r in range(reps):
        swapamd.hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        ids = srv.submit_fused(jobs, group_bytes=gb)
        srv.wait_all()
        t_runs.append(time.perf_counter() - t0)
        if r + 1 < reps:
            for j in ids:
                srv.release(j)
    t_fused = sorted(t_runs)[len(t_runs) // 2]
    fused_ok = all(srv.info(j)["status"] == 0 for j in ids)
    same_fused = fused_ok and all(srv.output(ids[i], noce) == plain[i] for i in range(n_samples))
    srv.close()
    for b in bufs:
        b.close()
    # every sample's noCEcontigs.fasta against what the reference wrote for it
    # running alone as one rank (digests from tests/golden/make_golden_batch.py)
    ref_ok, ref_n = 0, 0
    dpath = os.path.join(REPO, "tests", "golden", "np1_batch.json")
    if os.path.exists(dpath):
        import hashlib
        with open(dpath) as f:
            ref = json.load(f)["batch"]
        for i in range(n_samples):
            r = ref.get(str(i))
            if r:
                ref_n += 1
                ref_ok += hashlib.sha256(plain[i]).hexdigest() == r["files"]["noce"]["sha256"]
    return {"workload": "batch", "samples": n_samples, "input_bytes": sum(len(f) for f in fas),
            "input_kmers": kmers, "k": k, "c": c, "slots": slots,
            "per_sample_samples_per_s": n_samples / t_plain, "batched_samples_per_s": n_samples / t_fused,
            "factor": t_plain / t_fused, "per_sample_s": t_plain, "batched_s": t_fused,
            "batched_path": "fused: groups of up to %d MiB of FASTA, one batch build each, merged by %d slots"
                            % (group_mib, fused_slots),
            "identical_outputs": same_fused and same, "timed_passes": reps,
            "concurrent_slots": {"samples_per_s": n_samples / t_batch, "factor": t_plain / t_batch, "s": t_batch,
                                 "identical_outputs": same},
            "reference_checked": ref_n, "reference_identical": ref_ok}


def plasmid_reads(genome_bp=100000, read_len=100, coverage=50, error_rate=0.01, seed=7):
    """FASTA reads of a seeded circular genome (a plasmid): reads wrap around
    its end, so its unitig closes into a cycle; both strands, uniform
    substitution errors."""
    import numpy as np
    rng = np.random.default_rng(seed)
    g = rng.integers(0, 4, genome_bp, dtype=np.uint8)
    n = genome_bp * coverage // read_len
    starts = rng.integers(0, genome_bp, n)
    idx = (starts[:, None] + np.arange(read_len)[None, :]) % genome_bp
    r = g[idx]
    rev = rng.random(n) < 0.5
    r[rev] = 3 - r[rev, ::-1]
    err = rng.random(r.shape) < error_rate
    r = np.where(err, (r + rng.integers(1, 4, r.shape, dtype=np.uint8)) % 4, r).astype(np.uint8)
    seqs = np.frombuffer(b"ACGT", dtype=np.uint8)[r]
    out = bytearray()
    for i in range(n):
        out += b">p%d\n" % i
        out += seqs[i].tobytes()
        out += b"\n"
    return bytes(out)


def with_cycle(fa, k, c, steps, base_ms):
    """The bacterial reads plus a 100 kbp circular plasmid's at 50x: the plasmid's
    unitig is a cycle, merged by the host replay of cycles and fold-backs
    (reference order).  Step time of build + merge with inputs in HBM."""
    import swapamd
    fa2 = fa + plasmid_reads()
    dev = swapamd.DeviceBuffer(fa2)
    a = swapamd.Assembler(k, c)
    for _ in range(2):
        a.build_device(dev.ptr.value, dev.n)
        a.merge()
    swapamd.hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a.build_device(dev.ptr.value, dev.n)
        a.merge()
    swapamd.hip().hipDeviceSynchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    st = a.stats()
    a.close()
    dev.close()
    return {"workload": "bacterial + 100 kbp circular plasmid at 50x", "ms_per_step": ms,
            "bacterial_ms_per_step": base_ms, "special": st["special"], "kept": st["kept"],
            "survivors": st["survivors"]}


def serving(n_requests, rate, slots, seed=5):
    """The request-stream workload: requests arrive as a Poisson stream (mean
    `rate` per second), each a seeded genome log-uniform in [50 kbp, 5 Mbp] with
    reads of 100/150/250 bp at 20-60x, 0.5% errors, k from {21, 25, 31}, c=3,
    FASTA bytes in host memory (the transfer to the device is part of each
    request).  A server with `slots` contexts admits each request as a slot and
    the HBM budget allow; its noCE contigs are produced.  Latency = submit to
    end of the request."""
    import concurrent.futures as cf
    import random
    import swapamd
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        reqs = list(ex.map(lambda i: swapamd.generate("request", i), range(n_requests)))
    rng = random.Random(seed)
    arrivals, t = [], 0.0
    for _ in range(n_requests):
        t += rng.expovariate(rate)
        arrivals.append(t)
    noce = swapamd.SWAP_FILE_NOCE
    srv = swapamd.Server(slots=slots)
    # warm every slot once (device buffers grow to the largest request)
    big = sorted(range(n_requests), key=lambda i: -len(reqs[i][0]))[:slots]
    warm = srv.submit_batch([swapamd.Server.job(reqs[i][0], reqs[i][1], reqs[i][2], outputs=(noce,)) for i in big])
    srv.wait_all()
    for j in warm:
        srv.release(j)
    ids = []
    t0 = time.perf_counter()
    for i, (fa, k, c) in enumerate(reqs):
        delay = arrivals[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(swapamd.Server.job(fa, k, c, outputs=(noce,))))
    srv.wait_all()
    wall = time.perf_counter() - t0
    infos = [srv.info(j) for j in ids]
    lat = sorted(x["end_ms"] - x["submit_ms"] for x in infos)
    queue = sorted(x["start_ms"] - x["submit_ms"] for x in infos)
    ok = all(x["status"] == 0 for x in infos)
    srv.close()

    def pct(v, q):
        return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]

    return {"workload": "request", "requests": n_requests, "rate_per_s": rate, "slots": slots,
            "input_bytes": sum(len(r[0]) for r in reqs), "input_kmers": sum(x["kmers"] for x in infos),
            "p50_ms": pct(lat, 0.50), "p99_ms": pct(lat, 0.99), "mean_ms": sum(lat) / len(lat),
            "queue_p50_ms": pct(queue, 0.50), "queue_p99_ms": pct(queue, 0.99),
            "samples_per_s": n_requests / wall, "all_ok": ok}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="bacterial")
    ap.add_argument("--hash-order", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-samples", type=int, default=256,
                    help="samples of the many-sample workload for the batched-vs-per-sample figure (0 skips it)")
    ap.add_argument("--batch-slots", type=int, default=12,
                    help="concurrent contexts of the batched and serving figures (12: the best of 4/8/12/16 "
                         "on the batch workload)")
    ap.add_argument("--fuse-group-mib", type=int, default=256,
                    help="FASTA MiB per batch build of the batched figure's fused submission (256: the best of "
                         "128/256/512)")
    ap.add_argument("--fuse-slots", type=int, default=8,
                    help="merge slots of the fused submission (8: the best of 8/12/16 with 256 MiB groups)")
    ap.add_argument("--serve-requests", type=int, default=60,
                    help="requests of the serving figure (Poisson arrivals at --serve-rate; 0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a known-size kernel that calibrates "
                         "hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the k-mer graph path runs on one device and has no multi-GPU mode",
              file=sys.stderr)
        sys.exit(2)

    import swapamd

    t0 = time.time()
    fa, k, c = swapamd.generate(args.workload)
    print("generated %s: %d bytes in %.1f s" % (args.workload, len(fa), time.time() - t0), file=sys.stderr,
          flush=True)
    dev = swapamd.DeviceBuffer(fa)
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
    ref_order = not args.hash_order
    a = swapamd.Assembler(k, c, hash_order=args.hash_order)
    has_merge = True

    def step():
        nonlocal has_merge
        a.build_device(dev.ptr.value, dev.n)
        if has_merge:
            try:
                a.merge()
            except swapamd.SwapError as e:
                if e.code != -8:
                    raise
                has_merge = False

    for i in range(args.warmup):
        step()
        print("warmup %d done" % i, file=sys.stderr, flush=True)
    swapamd.hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    swapamd.hip().hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    st = a.stats()
    tuning = a.tuning()  # from the autotuned table (swap-assembler_amd/tune/swap_tune.txt)
    ms = 1000.0 * dt / args.steps
    N = st["kmers"]
    value = N / (dt / args.steps)

    # PCIe-inclusive figure: FASTA bytes from host memory (swap_build stages them
    # through pinned buffers), then the merge; after the timed steps, one untimed
    # step (the context's pinned staging buffers and device input buffer are
    # allocated on its first host build), then 3 timed ones
    a.build(fa)
    if has_merge:
        a.merge()
    host_steps = 3
    swapamd.hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(host_steps):
        a.build(fa)
        if has_merge:
            a.merge()
    swapamd.hip().hipDeviceSynchronize()
    t_host = (time.perf_counter() - t0) / host_steps
    t_path = dt / args.steps
    # a stream of host inputs with read-ahead: input i+1's transfer (swap_prefetch)
    # overlaps input i's build and merge; the first transfer (pipeline fill) and
    # every step's whole work are inside the timed region
    pipe_steps = 8
    swapamd.hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    a.prefetch(fa)
    for i in range(pipe_steps):
        if i + 1 < pipe_steps:
            a.prefetch(fa)
        a.build(fa)
        if has_merge:
            a.merge()
    swapamd.hip().hipDeviceSynchronize()
    t_pipe = (time.perf_counter() - t0) / pipe_steps
    host_input = {"ms_per_step": 1000.0 * t_host, "value": a.stats()["kmers"] / t_host, "unit": "input k-mers/s",
                  "input_gb_per_s": len(fa) / 1e9 / t_host,
                  # the step beyond the device-resident one: staging and the host-to-device copy
                  "transfer_ms": 1000.0 * (t_host - t_path),
                  "transfer_gb_per_s": len(fa) / 1e9 / max(t_host - t_path, 1e-9),
                  "read_ahead": {"steps": pipe_steps, "ms_per_step": 1000.0 * t_pipe,
                                 "value": a.stats()["kmers"] / t_pipe,
                                 "input_gb_per_s": len(fa) / 1e9 / t_pipe}}

    # dominant kernel of the last step
    reads = st["reads"]
    words = reads * ((len(fa.split(b"\n", 2)[1]) + 31) // 32)
    phases = {i: st["phase_ms"][i] for i in PHASES if st["phase_ms"][i] > 0}
    dom = max(phases, key=lambda i: phases[i])
    alg = algorithmic_bytes(dom, st, words, ref_order)
    achieved = alg / (phases[dom] * 1e-3) if alg else None
    roof = {"bound": "hbm", "kernel": PHASES[dom], "achieved": achieved / 1e9 if achieved else None,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": (achieved / HBM_PEAK) if achieved else None,
            "traffic": measured_traffic(dom, ref_order, args.workload), "algorithmic_bytes": alg,
            "kernel_ms": phases[dom]}
    # whole-path figure: edge-string bases measured as the bases of the contig file
    edge_bases = 0
    if has_merge:
        noce = a.file_bytes(swapamd.SWAP_FILE_NOCE)
        edge_bases = len(noce) - noce.count(b"\n") * len(b">proc0: ")
    read_len = len(fa.split(b"\n", 2)[1])
    pb = path_bytes(st, len(fa), read_len, k, edge_bases)
    roof["path"] = {"algorithmic_bytes": pb, "achieved": pb / (dt / args.steps) / 1e9, "unit": "GB/s",
                    "frac": pb / (dt / args.steps) / HBM_PEAK}
    # the table build alone: F + 8N + 8E + 48D + 16S over the device time of the
    # build phases (record scan through survivor keys)
    bb = build_bytes(st, len(fa), read_len, k)
    build_ms = sum(phases.get(i, 0.0) for i in BUILD_PHASES)
    roof["build"] = {"algorithmic_bytes": bb, "ms": build_ms, "achieved": bb / (build_ms * 1e-3) / 1e9,
                     "unit": "GB/s", "frac": bb / (build_ms * 1e-3) / HBM_PEAK}
    cycle = None
    if ref_order and has_merge and args.workload == "bacterial":
        cycle = with_cycle(fa, k, c, max(3, args.steps // 2), ms)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = reference_baseline(fa, k, c) or cpu_baseline(fa, k, c)
    batched = None
    if args.batch_samples > 0:
        a.close()
        dev.close()
        batched = batched_vs_per_sample(args.batch_samples, args.batch_slots, args.fuse_group_mib, args.fuse_slots)
    served = None
    if args.serve_requests > 0:
        a.close()
        dev.close()
        served = serving(args.serve_requests, args.serve_rate, args.batch_slots)
    line = {
        "metric": "input k-mers/s (build+merge) at 4.6Mbp/50x",
        "value": value,
        "unit": "input k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": args.workload, "genome_bp": 4600000, "read_len": 100, "coverage": 50,
                   "reads": reads, "k": k, "c": c, "order": "hash" if args.hash_order else "reference",
                   "merge": has_merge, "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stats": {kk: st[kk] for kk in ("kmers", "distinct", "survivors", "kept", "overflow", "special")},
        "tuning": tuning,
        "phase_ms": {PHASES[i]: round(v, 4) for i, v in phases.items()},
        "host_input": host_input,
        "with_cycle": cycle,
        "batched": batched,
        "serving": served,
    }
    print(json.dumps(line))
    a.close()
    dev.close()


if __name__ == "__main__":
    main()
