"""CPU-side checks of the product library: it loads, exports every entry point
declared in include/swap_hip.h, refuses to run without a gfx950 device, and
its host-side replay of the reference's hash-table order agrees with the
oracle (which uses the reference's own container)."""
import ctypes
import gzip
import json
import os
import re

import numpy as np
import pytest

import swapamd
from oracle_binding import Oracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")


def declared_symbols():
    text = open(os.path.join(REPO, "include", "swap_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(swap_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(swapamd.LIB_PATH)
    syms = declared_symbols()
    assert "swap_build" in syms and "swap_merge" in syms
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(swapamd.EXPORTS) == syms


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_create_without_gpu_fails_loudly():
    with pytest.raises(swapamd.SwapError) as e:
        swapamd.Assembler(31, 3)
    assert e.value.code in (-6,)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_server_without_gpu_fails_loudly():
    with pytest.raises(swapamd.SwapError) as e:
        swapamd.Server(slots=2)
    assert e.value.code in (-6,)


def test_server_rejects_bad_slot_count():
    h = ctypes.c_void_p()
    assert swapamd.lib().swap_server_create(0, 0, 0, ctypes.byref(h)) == -1
    assert swapamd.lib().swap_server_create(0, 65, 0, ctypes.byref(h)) == -1


def test_binding_constants_match_the_header():
    text = open(os.path.join(REPO, "include", "swap_hip.h")).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define (SWAP_\w+) \(?(-?\d+)u?\)?", text, re.M)}
    names = [n for n in dir(swapamd) if re.fullmatch(r"SWAP_(FLAG|FILE|E|JOB)_\w+", n)]
    assert len(names) >= 16
    for n in names:
        assert defs[n] == getattr(swapamd, n), n


def test_batch_entry_points_reject_bad_arguments():
    L = swapamd.lib()
    ptrs = (ctypes.c_void_p * 1)(None)
    sizes = (ctypes.c_size_t * 1)(10)
    assert L.swap_build_batch(None, ptrs, sizes, 1) == swapamd.SWAP_E_INVAL
    assert L.swap_batch_take(None, 0, None) == swapamd.SWAP_E_INVAL
    jobs = (swapamd.SwapJob * 1)()
    ids = (ctypes.c_int64 * 1)()
    assert L.swap_server_submit_fused(None, jobs, 1, 0, ids) == swapamd.SWAP_E_INVAL


def test_create_rejects_bad_k():
    cfg = swapamd.SwapCfg(20, 3, 0, 0)
    h = ctypes.c_void_p()
    assert swapamd.lib().swap_create(ctypes.byref(cfg), ctypes.byref(h)) == -1


def _chain(keys, rank, distinct):
    o1 = swapamd.table_order(keys, rank, distinct, False)
    r2 = np.empty(len(keys), dtype=np.uint64)
    r2[o1] = np.arange(len(keys), dtype=np.uint64)
    o2 = swapamd.table_order(keys, r2, len(keys), False)
    r3 = np.empty(len(keys), dtype=np.uint64)
    r3[o2] = np.arange(len(keys), dtype=np.uint64)
    return swapamd.table_order(keys, r3, len(keys), False)


@pytest.mark.parametrize("name", ["uniform_k31_c3", "dirty_k21_c2", "repeats_k25_c3"])
def test_visit_order_replay_matches_oracle(name):
    case = json.load(open(os.path.join(GOLD, "cases.json")))[name]
    with gzip.open(os.path.join(GOLD, name + ".fa.gz"), "rb") as f:
        fa = f.read()
    o = Oracle(fa, case["k"], case["c"])
    keys, _ = o.survivors()
    rank = o.survivor_rank()
    order = _chain(keys, rank, o.stats()["distinct"])
    assert np.array_equal(keys[order], o.visit_order())


def test_table_order_small_cases_match_container():
    # first-insert-order ranks over a table that grows through several primes
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 62, size=5000, dtype=np.uint64)
    # a subset of 700 keys out of 5000 inserts
    sub = np.sort(rng.choice(5000, size=700, replace=False))
    order = swapamd.table_order(keys[sub], sub.astype(np.uint64), 5000, False)
    # the oracle's container is the reference's: rebuild the same insert sequence
    # there through a tiny FASTA is awkward, so check the structural facts instead:
    # (1) a permutation, (2) grouped by final bucket (h % 6151), ascending
    assert sorted(order.tolist()) == list(range(700))
    h = keys[sub][order]
    hv = (h >> np.uint64(32)) ^ (h & np.uint64(0xFFFFFFFF)).astype(np.uint32).astype(np.int32).astype(np.int64).astype(np.uint64)
    b = hv % np.uint64(6151)
    assert np.all(np.diff(b.astype(np.int64)) >= 0)


@pytest.mark.parametrize("name", sorted(json.load(open(os.path.join(GOLD, "cases.json")))))
def test_visit_order_from_first_occurrences(name):
    """The replay used by the library (first-occurrence indices + the first
    occurrence of each growth rank) reproduces the reference container's order."""
    case = json.load(open(os.path.join(GOLD, "cases.json")))[name]
    with gzip.open(os.path.join(GOLD, name + ".fa.gz"), "rb") as f:
        fa = f.read()
    o = Oracle(fa, case["k"], case["c"])
    keys, _ = o.survivors()
    first = o.survivor_first()
    D = o.stats()["distinct"]
    starts = swapamd.table_epochs(D, False)
    ef = [o.first_of_rank(int(t)) for t in starts]
    order = swapamd.visit_order(keys, first, ef, D, False)
    assert np.array_equal(keys[order], o.visit_order())
