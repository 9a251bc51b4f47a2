"""Chain-level model of the k-molecule merge used by the GPU kernels, written
in plain Python: unitig walks between kept nodes, and merged multiplicities
from the Cartesian tree of the chain keyed by visit time.  Checked against
the oracle (a restatement of the reference's sequential SWAP merge) so that
the model the device implements is pinned independently of the device."""


def rcv(x, k):
    r = 0
    for _ in range(k):
        r = (r << 2) | (3 - (x & 3))
        x >>= 2
    return r


def splice(a, b):
    """(m, len) pairs merged with the reference's int arithmetic."""
    ma, la = int(a[0]), int(a[1])
    mb, lb = int(b[0]), int(b[1])
    s = ma * la + mb * lb
    s &= 0xFFFFFFFF
    pre = s - (1 << 32) if s >= (1 << 31) else s
    q = (pre % (1 << 64)) // (la + lb)
    return (q & 0xFF, la + lb)


def cartesian(units, prios):
    """units: n+1 (m, 1) unit edges; prios: n chain-node priorities."""
    n = len(prios)
    if n == 0:
        return units[0]
    # iterative evaluation over the tree built with a stack
    parent = [-1] * n
    stack = []
    for i in range(n):
        last = -1
        while stack and prios[stack[-1]] < prios[i]:
            last = stack.pop()
        if stack:
            parent[i] = stack[-1]
        if last != -1:
            parent[last] = i
        stack.append(i)
    left = [-1] * n
    right = [-1] * n
    root = -1
    for i in range(n):
        p = parent[i]
        if p == -1:
            root = i
        elif i < p:
            left[p] = i
        else:
            right[p] = i

    def ev(i, lo, hi):
        # node i spans unit edges lo..hi (inclusive); its left part lo..i, right i+1..hi
        lv = ev(left[i], lo, i) if left[i] != -1 else units[lo]
        rv = ev(right[i], i + 1, hi) if right[i] != -1 else units[hi]
        return splice(lv, rv)

    import sys
    sys.setrecursionlimit(max(10000, 4 * n))
    return ev(root, 0, n)


def compacted(keys, mult, prio, k):
    """{(node, slot): (string, m)} for kept nodes' edges into normal chains and
    kept-to-kept edges; chains touching a node that meets itself are left out."""
    mask = (1 << (2 * k)) - 1
    idx = {int(x): i for i, x in enumerate(keys)}
    deg = []
    for m in mult:
        deg.append((sum(1 for s in range(4) if m[s]), sum(1 for s in range(4, 8) if m[s])))
    one = [d == (1, 1) for d in deg]

    def hop(X, s):
        o = X if s < 4 else rcv(X, k)
        t = ((o << 2) | (s & 3)) & mask
        r = rcv(t, k)
        Y = max(t, r)
        return Y, (0 if t == Y else 1)

    out = {}
    for i, X in enumerate(keys):
        X = int(X)
        if one[i]:
            continue
        for s in range(8):
            if not mult[i][s]:
                continue
            units = [(int(mult[i][s]), 1)]
            prios = []
            bases = ["ACGT"[s & 3]]
            Y, o = hop(X, s)
            seen = set()
            bad = False
            while one[idx[Y]]:
                j = idx[Y]
                if j in seen or j == i:
                    bad = True
                    break
                seen.add(j)
                m = mult[j]
                slot = [t for t in range(4) if m[t]][0] if o == 0 else [t for t in range(4, 8) if m[t]][0]
                units.append((int(m[slot]), 1))
                prios.append(int(prio[j]))
                Y, o = hop(Y, slot)
                bases.append("ACGT"[slot & 3])
            if bad:
                continue
            out[(X, s)] = ("".join(bases), cartesian(units, prios)[0])
    return out


NIL = 0xFFFFFFFF


def chain_bookkeeping(keys, mult, k):
    """The device merge's chain bookkeeping (swap_merge.hip: k_classify, the
    list ranking's walks -- here sequential --, k_kept_edges, k_fill_nodes'
    string and chain-array positions), restated on the CPU with every bound
    checked.  Returns (problems, stats): problems lists every inconsistency a
    kernel would trip over (an edge whose two ends disagree, a chain that starts
    and ends at one edge, a position outside its chain array, a chain array
    whose positions are not each written once)."""
    keys = [int(x) for x in keys]
    S = len(keys)
    idx = {x: i for i, x in enumerate(keys)}
    km = (1 << (2 * k)) - 1
    m = [list(map(int, r)) for r in mult]

    def hop(X, s):
        ox = rcv(X, k) if s >= 4 else X
        t = ((ox << 2) | (s & 3)) & km
        r = rcv(t, k)
        Y = max(t, r)
        j = idx.get(Y)
        if j is None:
            return NIL
        return (j << 1) | (0 if t == Y else 1)

    is11 = [False] * S
    nxt = [NIL] * (2 * S)
    selfadj = [False] * S
    for i in range(S):
        nr = sum(1 for t in range(4) if m[i][t])
        nl = sum(1 for t in range(4, 8) if m[i][t])
        if nr == 1 and nl == 1:
            is11[i] = True
            r = [t for t in range(4) if m[i][t]][-1]
            l = [t for t in range(4, 8) if m[i][t]][-1]
            a, b = hop(keys[i], r), hop(keys[i], l)
            if a == NIL or b == NIL:
                return [("missing neighbour", i)], {}
            nxt[2 * i], nxt[2 * i + 1] = a, b
            selfadj[i] = (a >> 1) == i or (b >> 1) == i
    # sequential walks (what the ruling set + jumping computes)
    walk = {}
    sp = {}
    for v in range(2 * S):
        if not is11[v >> 1]:
            continue
        u, dist, s_ = v, 1, selfadj[v >> 1]
        seen = {v}
        term, last, succ = NIL, v, NIL
        while True:
            t = nxt[u]
            if t == NIL or not is11[t >> 1]:
                term, last = t, u
                break
            if t in seen:
                succ = t  # cycle
                break
            seen.add(t)
            s_ = s_ or selfadj[t >> 1]
            u = t
            dist += 1
        walk[v] = (succ, dist, term, last)
        sp[v] = s_
    kept = [i for i in range(S) if not is11[i]]
    kept_of = {i: n for n, i in enumerate(kept)}

    def end_slot(term, last):
        a = term & 1
        L = keys[last >> 1]
        first = (3 - (L & 3)) if (last & 1) else ((L >> (2 * (k - 1))) & 3)
        return (0 if a else 4) + (3 - first)

    elen, eother, csize, espec = {}, {}, {}, {}
    for kap, i in enumerate(kept):
        for s in range(8):
            g = 8 * kap + s
            elen[g], eother[g], csize[g], espec[g] = 0, NIL, 0, 0
            if not m[i][s]:
                continue
            t = hop(keys[i], s)
            if t == NIL:
                continue
            j = t >> 1
            if not is11[j]:
                elen[g] = 1
                continue
            succ, dist, term, last = walk[t]
            if succ != NIL or sp[t] or term == NIL:
                espec[g] = 1
                continue
            elen[g] = dist + 1
            other = 8 * kept_of[term >> 1] + end_slot(term, last)
            eother[g] = other
            if g < other:
                csize[g] = dist + 2
    problems = []
    # symmetry of chain edges
    for g, oth in eother.items():
        if oth == NIL:
            continue
        if eother.get(oth) != g:
            problems.append(("asymmetric chain edge", g, oth, eother.get(oth)))
        elif elen[oth] != elen[g]:
            problems.append(("unequal lengths", g, oth, elen[g], elen[oth]))
        if g == oth:
            problems.append(("edge is its own other end", g))
    # chain positions written by k_fill_nodes
    writes = {}
    for v in range(2 * S):
        j = v >> 1
        if not is11[j]:
            continue
        vr = v ^ 1
        bs, bd, bt, bl = walk[vr]
        if bs != NIL or bt == NIL or sp[v] or sp[vr]:
            continue
        start = 8 * kept_of[bt >> 1] + end_slot(bt, bl)
        a_s, a_d, a_t, a_l = walk[v]
        if a_s != NIL or a_t == NIL:
            problems.append(("reverse walk of a normal node does not end", v))
            continue
        end = 8 * kept_of[a_t >> 1] + end_slot(a_t, a_l)
        if elen.get(start, 0) < 2:
            problems.append(("node on an edge without chain", v, start, elen.get(start)))
            continue
        if not (1 <= bd <= elen[start] - 1):
            problems.append(("string position out of range", v, bd, elen[start]))
        if start < end:
            if csize.get(start, 0) == 0:
                problems.append(("owner edge without chain array", v, start, end))
                continue
            if not (1 <= bd <= csize[start] - 2):
                problems.append(("chain position out of range", v, bd, csize[start]))
            writes.setdefault(start, []).append(bd)
        elif start == end:
            problems.append(("chain starts and ends at one edge", v, start))
    for g, cs in csize.items():
        if cs == 0:
            continue
        got = sorted(writes.get(g, []))
        if got != list(range(1, cs - 1)):
            problems.append(("chain positions not a permutation", g, cs, got[:5], len(got)))
    stats = {"S": S, "n11": sum(is11), "kept": len(kept), "special_edges": sum(espec.values()),
             "chains": sum(1 for v in csize.values() if v)}
    return problems, stats
