"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure
only; the checker for the GPU path, never the thing measured."""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle.so")

_L = None


def _lib():
    global _L
    if _L is None:
        L = ctypes.CDLL(ORACLE_LIB)
        L.oracle_run.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_run.restype = ctypes.c_void_p
        L.oracle_free.argtypes = [ctypes.c_void_p]
        for f in ("oracle_num_occurrences", "oracle_num_reads", "oracle_num_distinct", "oracle_num_survivors",
                  "oracle_num_kept"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_uint64
        for f in ("oracle_table", "oracle_survivors"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
            getattr(L, f).restype = ctypes.c_uint64
        for f in ("oracle_survivor_first", "oracle_survivor_rank", "oracle_visit_order"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
            getattr(L, f).restype = ctypes.c_uint64
        L.oracle_first_of_rank.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.oracle_first_of_rank.restype = ctypes.c_uint64
        L.oracle_file.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64]
        L.oracle_file.restype = ctypes.c_uint64
        _L = L
    return _L


class Oracle:
    # planted mistakes (oracle flags bits 1 and 2), for tests that show a
    # fixture detects them
    PLANT_EAGER_UNLOCK = 2
    PLANT_DOUBLE_LOCK = 4

    def __init__(self, fasta: bytes, k: int, c: int, kmer_graph_only=False, plant=0):
        self._h = _lib().oracle_run(fasta, len(fasta), k, c, (1 if kmer_graph_only else 0) | plant)
        if not self._h:
            raise ValueError("oracle rejected the input")

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().oracle_free(self._h)

    def stats(self):
        L = _lib()
        return {"kmers": L.oracle_num_occurrences(self._h), "reads": L.oracle_num_reads(self._h),
                "distinct": L.oracle_num_distinct(self._h), "survivors": L.oracle_num_survivors(self._h),
                "kept": L.oracle_num_kept(self._h)}

    def _table(self, fn):
        n = fn(self._h, None, None, 0)
        keys = np.zeros(n, dtype=np.uint64)
        mult = np.zeros((n, 8), dtype=np.uint8)
        if n:
            fn(self._h, keys.ctypes.data, mult.ctypes.data, n)
        return keys, mult

    def table(self):
        return self._table(_lib().oracle_table)

    def survivors(self):
        return self._table(_lib().oracle_survivors)

    def survivor_first(self):
        n = _lib().oracle_survivor_first(self._h, None, 0)
        out = np.zeros(n, dtype=np.uint64)
        if n:
            _lib().oracle_survivor_first(self._h, out.ctypes.data, n)
        return out

    def survivor_rank(self):
        n = _lib().oracle_survivor_rank(self._h, None, 0)
        out = np.zeros(n, dtype=np.uint64)
        if n:
            _lib().oracle_survivor_rank(self._h, out.ctypes.data, n)
        return out

    def first_of_rank(self, r):
        return _lib().oracle_first_of_rank(self._h, r)

    def visit_order(self):
        n = _lib().oracle_visit_order(self._h, None, 0)
        out = np.zeros(n, dtype=np.uint64)
        if n:
            _lib().oracle_visit_order(self._h, out.ctypes.data, n)
        return out

    def file(self, which):
        n = _lib().oracle_file(self._h, which, None, 0)
        buf = ctypes.create_string_buffer(max(n, 1))
        _lib().oracle_file(self._h, which, buf, n)
        return buf.raw[:n]
