"""ctypes binding of libswaphip.so (include/swap_hip.h) for Python callers.

The library is the product: every call goes through its C ABI.  There is no
Python fallback -- if the library or a gfx950 device is missing, the calls
raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libswaphip.so")
GEN_PATH = os.path.join(HERE, "lib", "libswapgen.so")

SWAP_FLAG_KMER_GRAPH = 1
SWAP_FLAG_HASH_ORDER = 2
SWAP_FLAG_TABLE_SPILL = 4
SWAP_FLAG_WIDE_PARTITIONS = 8
SWAP_FLAG_SINGLE_FILL = 16
SWAP_FLAG_COARSE_RENUMBER = 32
SWAP_E_INVAL = -1  # bad argument (include/swap_hip.h)
SWAP_E_STATE = -8  # call out of order
SWAP_FILE_KMER_GRAPH = 0
SWAP_FILE_CONTIG_GRAPH = 1
SWAP_FILE_NOCE = 2
SWAP_FILE_PERF_CONTIGS = 3
SWAP_FILE_CE_CONTIGS = 4
SWAP_FILE_JUNG_ARC = 5
SWAP_FILE_JUNG_MUL = 6

# exported entry points and their signatures (mirrors include/swap_hip.h)
EXPORTS = [
    "swap_create", "swap_destroy", "swap_build", "swap_build_device", "swap_prefetch", "swap_merge", "swap_assemble_static",
    "swap_build_batch", "swap_batch_take", "swap_server_submit_fused",
    "swap_get_stats", "swap_tuning_for", "swap_set_tuning", "swap_get_tuning",
    "swap_export_kmer_graph", "swap_format_file", "swap_write_file", "swap_strerror", "swap_last_error",
    "swap_table_order", "swap_visit_order", "swap_table_epochs", "swap_reconfigure", "swap_export_compacted",
    "swap_server_create", "swap_server_submit", "swap_server_submit_batch", "swap_server_wait",
    "swap_server_wait_all", "swap_server_job_info", "swap_server_job_error", "swap_server_output",
    "swap_server_release", "swap_server_destroy", "swap_contig_extension",
]

SWAP_JOB_QUEUED = 1
SWAP_JOB_RUNNING = 2
SWAP_JOB_DEVICE_INPUT = 1


class SwapCfg(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("c", ctypes.c_int), ("device", ctypes.c_int), ("flags", ctypes.c_uint)]


class SwapStats(ctypes.Structure):
    _fields_ = [("reads", ctypes.c_uint64), ("kmers", ctypes.c_uint64), ("distinct", ctypes.c_uint64),
                ("survivors", ctypes.c_uint64), ("kept", ctypes.c_uint64), ("overflow", ctypes.c_uint64),
                ("special", ctypes.c_uint64), ("build_ms", ctypes.c_double), ("merge_ms", ctypes.c_double),
                ("phase_ms", ctypes.c_double * 16)]


class SwapJob(ctypes.Structure):
    _fields_ = [("fasta", ctypes.c_void_p), ("n", ctypes.c_size_t), ("k", ctypes.c_int), ("c", ctypes.c_int),
                ("flags", ctypes.c_uint), ("outputs", ctypes.c_uint), ("job_flags", ctypes.c_uint)]


class SwapTuning(ctypes.Structure):
    _fields_ = [("part_recs", ctypes.c_uint32), ("fill_recs", ctypes.c_uint32), ("index_load", ctypes.c_uint32),
                ("ruler_stride", ctypes.c_uint32), ("occ_groups", ctypes.c_uint32),
                ("order_bin_keys", ctypes.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SwapJobInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("end_ms", ctypes.c_double), ("stats", SwapStats)]


_lib = None


def lib():
    """Loads libswaphip.so (raises OSError if it was not built)."""
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        L.swap_create.argtypes = [ctypes.POINTER(SwapCfg), ctypes.POINTER(ctypes.c_void_p)]
        L.swap_create.restype = ctypes.c_int
        L.swap_destroy.argtypes = [ctypes.c_void_p]
        L.swap_destroy.restype = None
        L.swap_build.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.swap_build.restype = ctypes.c_int
        L.swap_build_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.swap_build_device.restype = ctypes.c_int
        L.swap_prefetch.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.swap_prefetch.restype = ctypes.c_int
        L.swap_merge.argtypes = [ctypes.c_void_p]
        L.swap_merge.restype = ctypes.c_int
        L.swap_build_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.c_size_t]
        L.swap_build_batch.restype = ctypes.c_int
        L.swap_batch_take.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.swap_batch_take.restype = ctypes.c_int
        L.swap_assemble_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.swap_assemble_static.restype = ctypes.c_int
        L.swap_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(SwapStats)]
        L.swap_get_stats.restype = ctypes.c_int
        L.swap_tuning_for.argtypes = [ctypes.c_uint64, ctypes.POINTER(SwapTuning)]
        L.swap_tuning_for.restype = ctypes.c_int
        L.swap_set_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(SwapTuning)]
        L.swap_set_tuning.restype = ctypes.c_int
        L.swap_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(SwapTuning)]
        L.swap_get_tuning.restype = ctypes.c_int
        L.swap_export_kmer_graph.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.swap_export_kmer_graph.restype = ctypes.c_int64
        L.swap_format_file.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        L.swap_format_file.restype = ctypes.c_int64
        L.swap_write_file.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        L.swap_write_file.restype = ctypes.c_int64
        L.swap_strerror.argtypes = [ctypes.c_int]
        L.swap_strerror.restype = ctypes.c_char_p
        L.swap_last_error.argtypes = [ctypes.c_void_p]
        L.swap_last_error.restype = ctypes.c_char_p
        L.swap_table_order.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64,
                                       ctypes.c_int, ctypes.c_void_p]
        L.swap_table_order.restype = ctypes.c_int64
        L.swap_visit_order.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
        L.swap_visit_order.restype = ctypes.c_int64
        L.swap_table_epochs.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        L.swap_table_epochs.restype = ctypes.c_int64
        L.swap_reconfigure.argtypes = [ctypes.c_void_p, ctypes.POINTER(SwapCfg)]
        L.swap_reconfigure.restype = ctypes.c_int
        L.swap_export_compacted.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_uint64)]
        L.swap_export_compacted.restype = ctypes.c_int64
        L.swap_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
        L.swap_server_create.restype = ctypes.c_int
        L.swap_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(SwapJob)]
        L.swap_server_submit.restype = ctypes.c_int64
        L.swap_server_submit_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(SwapJob), ctypes.c_size_t,
                                               ctypes.c_void_p]
        L.swap_server_submit_batch.restype = ctypes.c_int
        L.swap_server_submit_fused.argtypes = [ctypes.c_void_p, ctypes.POINTER(SwapJob), ctypes.c_size_t,
                                               ctypes.c_size_t, ctypes.c_void_p]
        L.swap_server_submit_fused.restype = ctypes.c_int
        L.swap_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.swap_server_wait.restype = ctypes.c_int
        L.swap_server_wait_all.argtypes = [ctypes.c_void_p]
        L.swap_server_wait_all.restype = ctypes.c_int
        L.swap_server_job_info.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(SwapJobInfo)]
        L.swap_server_job_info.restype = ctypes.c_int
        L.swap_server_job_error.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.swap_server_job_error.restype = ctypes.c_char_p
        L.swap_server_output.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_size_t]
        L.swap_server_output.restype = ctypes.c_int64
        L.swap_server_release.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.swap_server_release.restype = ctypes.c_int
        L.swap_contig_extension.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_size_t,
                                            ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        L.swap_contig_extension.restype = ctypes.c_int64
        L.swap_server_destroy.argtypes = [ctypes.c_void_p]
        L.swap_server_destroy.restype = None
        _lib = L
    return _lib


class SwapError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        text = lib().swap_strerror(code).decode()
        super().__init__("%s (%d)%s" % (text, code, (": " + msg) if msg else ""))


class Assembler:
    """One assembly context on one device (a thin wrapper of swap_ctx)."""

    def __init__(self, k, c, device=0, kmer_graph_only=False, hash_order=False, table_spill=False,
                 wide_partitions=False, single_fill=False, coarse_renumber=False):
        self.k, self.c = k, c
        cfg = SwapCfg(k, c, device, (SWAP_FLAG_KMER_GRAPH if kmer_graph_only else 0)
                      | (SWAP_FLAG_HASH_ORDER if hash_order else 0)
                      | (SWAP_FLAG_TABLE_SPILL if table_spill else 0)
                      | (SWAP_FLAG_WIDE_PARTITIONS if wide_partitions else 0)
                      | (SWAP_FLAG_SINGLE_FILL if single_fill else 0)
                      | (SWAP_FLAG_COARSE_RENUMBER if coarse_renumber else 0))
        h = ctypes.c_void_p()
        rc = lib().swap_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise SwapError(rc)
        self._h = h
        self._ahead = []  # prefetched inputs, kept alive until their build

    def close(self):
        if self._h:
            lib().swap_destroy(self._h)
            self._h = None
        self._ahead = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise SwapError(int(rc), lib().swap_last_error(self._h).decode())
        return rc

    def build(self, fasta: bytes):
        try:
            self._check(lib().swap_build(self._h, fasta, len(fasta)))
        finally:
            # the library consumed the oldest prefetch if it was this input,
            # else dropped them all
            if self._ahead and self._ahead[0] is fasta:
                self._ahead.pop(0)
            else:
                self._ahead = []

    def prefetch(self, fasta: bytes):
        """Starts the host-to-device copy of a later build's input (swap_prefetch):
        the next build() of the same bytes object uses the copied bytes."""
        self._check(lib().swap_prefetch(self._h, fasta, len(fasta)))
        self._ahead.append(fasta)

    def build_device(self, ptr: int, n: int):
        self._check(lib().swap_build_device(self._h, ctypes.c_void_p(ptr), n))

    def merge(self):
        self._check(lib().swap_merge(self._h))

    def build_batch(self, bufs):
        """One build of many samples (swap_build_batch): bufs is a list of
        device-resident FASTA inputs, each a DeviceBuffer or a (pointer, size) pair."""
        pairs = [(b.ptr.value, b.n) if isinstance(b, DeviceBuffer) else (int(b[0]), int(b[1])) for b in bufs]
        n = len(pairs)
        ptrs = (ctypes.c_void_p * n)(*[p for p, _ in pairs])
        sizes = (ctypes.c_size_t * n)(*[m for _, m in pairs])
        self._check(lib().swap_build_batch(self._h, ptrs, sizes, n))

    def take(self, batch, i):
        """This context takes sample i of batch's last build_batch (swap_batch_take)."""
        self._check(lib().swap_batch_take(batch._h, i, self._h))

    def assemble_static(self, ptr: int, n: int):
        """Build (and merge, unless kmer_graph_only) device-resident FASTA bytes
        in the library's static scratch: no device allocation in the call."""
        self._check(lib().swap_assemble_static(self._h, ctypes.c_void_p(ptr), n))

    def set_tuning(self, tuning=None):
        """Fixes the performance parameters of the next builds (a dict with the
        SwapTuning fields); None goes back to the autotuned table."""
        if tuning is None:
            self._check(lib().swap_set_tuning(self._h, None))
        else:
            t = SwapTuning(**tuning)
            self._check(lib().swap_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self):
        t = SwapTuning()
        self._check(lib().swap_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def stats(self):
        s = SwapStats()
        self._check(lib().swap_get_stats(self._h, ctypes.byref(s)))
        d = {f: getattr(s, f) for f, _ in SwapStats._fields_}
        d["phase_ms"] = list(s.phase_ms)
        return d

    def kmer_graph(self):
        """(keys u64[S], mult u8[S,8]) sorted by key."""
        n = self._check(lib().swap_export_kmer_graph(self._h, None, None, 0))
        keys = np.zeros(n, dtype=np.uint64)
        mult = np.zeros((n, 8), dtype=np.uint8)
        if n:
            self._check(lib().swap_export_kmer_graph(self._h, keys.ctypes.data, mult.ctypes.data, n))
        return keys, mult

    def file_bytes(self, which):
        n = self._check(lib().swap_format_file(self._h, which, None, 0))
        buf = ctypes.create_string_buffer(max(n, 1))
        self._check(lib().swap_format_file(self._h, which, buf, n))
        return buf.raw[:n]

    def write_file(self, which, path):
        return self._check(lib().swap_write_file(self._h, which, path.encode()))


    def reconfigure(self, k, c, kmer_graph_only=False, hash_order=False, device=0):
        cfg = SwapCfg(k, c, device, (SWAP_FLAG_KMER_GRAPH if kmer_graph_only else 0)
                      | (SWAP_FLAG_HASH_ORDER if hash_order else 0))
        self._check(lib().swap_reconfigure(self._h, ctypes.byref(cfg)))
        self.k, self.c = k, c

    def compacted(self):
        """(keys u64[n], edge_off u64[n,8], edge_len u32[n,8], edge_mult u8[n,8], pool bytes)."""
        plen = ctypes.c_uint64()
        n = self._check(lib().swap_export_compacted(self._h, None, None, None, None, 0, None, 0, ctypes.byref(plen)))
        keys = np.zeros(n, dtype=np.uint64)
        off = np.zeros((n, 8), dtype=np.uint64)
        ln = np.zeros((n, 8), dtype=np.uint32)
        mu = np.zeros((n, 8), dtype=np.uint8)
        pool = ctypes.create_string_buffer(max(plen.value, 1))
        self._check(lib().swap_export_compacted(self._h, keys.ctypes.data, off.ctypes.data, ln.ctypes.data,
                                                mu.ctypes.data, n, pool, plen.value, ctypes.byref(plen)))
        return keys, off, ln, mu, pool.raw[:plen.value]


class Server:
    """Many assemblies on one GPU: swap_server with `slots` concurrent contexts."""

    def __init__(self, slots=4, device=0, hbm_budget=0):
        h = ctypes.c_void_p()
        rc = lib().swap_server_create(device, slots, hbm_budget, ctypes.byref(h))
        if rc:
            raise SwapError(rc)
        self._h = h
        self._keep = {}  # job id -> input bytes kept alive until the job ends

    def close(self):
        if self._h:
            lib().swap_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def job(fasta, k, c, outputs=(SWAP_FILE_NOCE,), kmer_graph_only=False, hash_order=False, device_ptr=None):
        """A SwapJob for host bytes (fasta) or device memory (device_ptr, len(fasta) ignored if fasta is int)."""
        j = SwapJob()
        if device_ptr is not None:
            j.fasta = ctypes.c_void_p(device_ptr)
            j.n = fasta if isinstance(fasta, int) else len(fasta)
            j.job_flags = SWAP_JOB_DEVICE_INPUT
        else:
            j.fasta = ctypes.cast(ctypes.c_char_p(fasta), ctypes.c_void_p)
            j.n = len(fasta)
            j.job_flags = 0
        j.k, j.c = k, c
        j.flags = (SWAP_FLAG_KMER_GRAPH if kmer_graph_only else 0) | (SWAP_FLAG_HASH_ORDER if hash_order else 0)
        j.outputs = 0
        for w in outputs:
            j.outputs |= 1 << w
        j._src = fasta  # the input stays alive as long as the job object
        return j

    def submit(self, job, keep=None):
        jid = lib().swap_server_submit(self._h, ctypes.byref(job))
        if jid < 0:
            raise SwapError(int(jid))
        self._keep[jid] = (job, keep)
        return jid

    def submit_batch(self, jobs, keep=None):
        arr = (SwapJob * len(jobs))(*jobs)
        ids = np.zeros(len(jobs), dtype=np.int64)
        rc = lib().swap_server_submit_batch(self._h, arr, len(jobs), ids.ctypes.data)
        if rc:
            raise SwapError(rc)
        for i, jid in enumerate(ids):
            self._keep[int(jid)] = (jobs[i], keep[i] if keep else None)
        return [int(x) for x in ids]

    def submit_fused(self, jobs, group_bytes=0, keep=None):
        """Merge jobs of one k and c with device-resident inputs, built in
        groups by one batch build each (swap_server_submit_fused)."""
        arr = (SwapJob * len(jobs))(*jobs)
        ids = np.zeros(len(jobs), dtype=np.int64)
        rc = lib().swap_server_submit_fused(self._h, arr, len(jobs), group_bytes, ids.ctypes.data)
        if rc:
            raise SwapError(rc)
        for i, jid in enumerate(ids):
            self._keep[int(jid)] = (jobs[i], keep[i] if keep else None)
        return [int(x) for x in ids]

    def wait(self, jid):
        rc = lib().swap_server_wait(self._h, jid)
        if rc < 0:
            raise SwapError(rc, lib().swap_server_job_error(self._h, jid).decode())
        return rc

    def wait_all(self):
        rc = lib().swap_server_wait_all(self._h)
        if rc < 0:
            raise SwapError(rc)

    def info(self, jid):
        i = SwapJobInfo()
        rc = lib().swap_server_job_info(self._h, jid, ctypes.byref(i))
        if rc:
            raise SwapError(rc)
        return {"status": i.status, "submit_ms": i.submit_ms, "start_ms": i.start_ms, "end_ms": i.end_ms,
                "kmers": i.stats.kmers, "survivors": i.stats.survivors, "kept": i.stats.kept}

    def output(self, jid, which):
        n = lib().swap_server_output(self._h, jid, which, None, 0)
        if n < 0:
            raise SwapError(int(n))
        buf = ctypes.create_string_buffer(max(n, 1))
        n = lib().swap_server_output(self._h, jid, which, buf, n)
        if n < 0:
            raise SwapError(int(n))
        return buf.raw[:n]

    def error(self, jid):
        """Error text of an ended job ("" if none)."""
        return lib().swap_server_job_error(self._h, jid).decode()

    def release(self, jid):
        lib().swap_server_release(self._h, jid)
        self._keep.pop(jid, None)


def tuning_for(input_bytes):
    """The tuning a build of that many FASTA bytes takes from the autotuned table."""
    t = SwapTuning()
    rc = lib().swap_tuning_for(input_bytes, ctypes.byref(t))
    if rc:
        raise SwapError(rc)
    return t.as_dict()


def contig_extension(k, survivors, contig_graph: bytes, which):
    """One output file of the contig extension run on contigGraph.txt bytes (host only)."""
    n = lib().swap_contig_extension(k, survivors, contig_graph, len(contig_graph), which, None, 0)
    if n < 0:
        raise SwapError(int(n))
    buf = ctypes.create_string_buffer(max(n, 1))
    n = lib().swap_contig_extension(k, survivors, contig_graph, len(contig_graph), which, buf, n)
    if n < 0:
        raise SwapError(int(n))
    return buf.raw[:n]


def table_order(keys, ranks, n_total, extra_resize=False):
    """Iteration order of the reference's hash table for a subset of keys."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    ranks = np.ascontiguousarray(ranks, dtype=np.uint64)
    out = np.zeros(len(keys), dtype=np.uint32)
    rc = lib().swap_table_order(keys.ctypes.data, ranks.ctypes.data, len(keys), n_total, int(extra_resize),
                                out.ctypes.data)
    if rc < 0:
        raise SwapError(int(rc))
    return out


def table_epochs(n_total, extra_resize=False):
    out = np.zeros(64, dtype=np.uint64)
    n = lib().swap_table_epochs(n_total, int(extra_resize), out.ctypes.data, 64)
    return out[:n]


def visit_order(keys, first, epoch_first, distinct, tail_lookup=False):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    first = np.ascontiguousarray(first, dtype=np.uint64)
    ef = np.ascontiguousarray(epoch_first, dtype=np.uint64)
    out = np.zeros(len(keys), dtype=np.uint32)
    rc = lib().swap_visit_order(keys.ctypes.data, first.ctypes.data, len(keys), ef.ctypes.data, len(ef), distinct,
                                int(tail_lookup), out.ctypes.data)
    if rc < 0:
        raise SwapError(int(rc))
    return out


def generate(workload, index=0):
    """FASTA bytes of a named workload (small, bacterial, repeat, batch, request)
    from the seeded generator; returns (bytes, k, c)."""
    g = ctypes.CDLL(GEN_PATH)
    g.swapgen_workload_size.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
    g.swapgen_workload_size.restype = ctypes.c_uint64
    g.swapgen_workload.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64]
    g.swapgen_workload.restype = ctypes.c_uint64
    k, c = ctypes.c_int(), ctypes.c_int()
    n = g.swapgen_workload_size(workload.encode(), index, ctypes.byref(k), ctypes.byref(c))
    if n == 0:
        raise ValueError("unknown workload " + workload)
    buf = ctypes.create_string_buffer(n)
    m = g.swapgen_workload(workload.encode(), index, buf, n)
    assert m == n
    return buf.raw[:n], k.value, c.value


def generate_params(genome_len, read_len, coverage, error_rate, seed, repeat_frac=0.0, families=0):
    g = ctypes.CDLL(GEN_PATH)
    g.swapgen_fasta.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]
    g.swapgen_fasta.restype = ctypes.c_uint64
    n = g.swapgen_fasta(genome_len, read_len, coverage, error_rate, seed, repeat_frac, families, None, 0)
    buf = ctypes.create_string_buffer(n)
    m = g.swapgen_fasta(genome_len, read_len, coverage, error_rate, seed, repeat_frac, families, buf, n)
    assert m == n
    return buf.raw[:n]


# ---------------------------------------------------------------------------
# minimal HIP runtime access (device-resident inputs for benchmarks), no torch

_hip = None


def hip():
    global _hip
    if _hip is None:
        H = ctypes.CDLL("libamdhip64.so")
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipDeviceSynchronize.argtypes = []
        H.hipSetDevice.argtypes = [ctypes.c_int]
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        H.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipSetDevice", "hipGetDeviceCount",
                  "hipMemset"):
            getattr(H, f).restype = ctypes.c_int
        _hip = H
    return _hip


class DeviceBuffer:
    """Bytes copied once into device memory (hipMalloc), freed on close()."""

    def __init__(self, data: bytes, device=0):
        H = hip()
        if H.hipSetDevice(device):
            raise RuntimeError("hipSetDevice failed")
        self.ptr = ctypes.c_void_p()
        if H.hipMalloc(ctypes.byref(self.ptr), max(len(data), 1)):
            raise RuntimeError("hipMalloc failed")
        if H.hipMemcpy(self.ptr, data, len(data), 1):  # hipMemcpyHostToDevice
            raise RuntimeError("hipMemcpy failed")
        self.n = len(data)

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None


def gpu_count():
    n = ctypes.c_int(0)
    try:
        if hip().hipGetDeviceCount(ctypes.byref(n)):
            return 0
    except OSError:
        return 0
    return n.value
