"""The tower's saddle-reduction rule, checked on the CPU against the full
level-ordered contraction.

The device (e-hgp_amd/csrc/tower_gpu.hip, k_attach / k_edge_flags /
k_msf_* / k_reduce_saddles) drops, per order, the saddles whose effect on
the contraction is already decided: with first[m] the least saddle touching
minimum m, a carrier m of saddle s is fresh when first[m] == s (order one has
none); fresh carriers attach to the saddle's first owned carrier (or its first
fresh one); a Boruvka spanning forest over star edges between the
representatives of owned carriers counts the components each saddle joins; a
saddle at an unshared level with an owned carrier and no forest edge is
dropped.  This test restates that rule in Python and compares the forests it
leaves the contraction with the forests of the full contraction (the
semantics of reference direct_morse_forest_journal.cpp:1042, as in
oracle/tower.py direct_forests) on random hypergraphs with shared levels; it
pins the rule, the GPU tests pin the kernels against the oracles.
"""

import random
from collections import Counter


def contract(M, saddles, order_one):
    """Level-ordered contraction: births, silent growth, multifusions."""
    parent = list(range(M))
    owner = [-1] * M
    nodes = []

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    if order_one:
        for p in range(M):
            nodes.append(("b", -1, ()))
            owner[p] = len(nodes) - 1
    i = 0
    while i < len(saddles):
        j = i
        while j < len(saddles) and saddles[j][0] == saddles[i][0]:
            j += 1
        bp, first_saddle = {}, {}

        def bfind(x):
            while bp[x] != x:
                bp[x] = bp[bp[x]]
                x = bp[x]
            return x

        for e in range(i, j):
            head = None
            for c in saddles[e][1]:
                r = find(c)
                if r not in bp:
                    bp[r] = r
                    first_saddle[r] = e
                if head is None:
                    head = r
                else:
                    a, b = bfind(head), bfind(r)
                    if a != b:
                        bp[b] = a
        groups = {}
        for r in bp:
            g = groups.setdefault(bfind(r), [len(saddles), []])
            g[0] = min(g[0], first_saddle[r])
            g[1].append(r)
        for _, members in sorted(groups.values()):
            members.sort()
            roots = sorted({owner[r] for r in members if owner[r] >= 0})
            level = saddles[i][0]
            if not roots:
                nodes.append(("b", level, ()))
                node = len(nodes) - 1
            elif len(roots) == 1:
                node = roots[0]
            else:
                nodes.append(("m", level, tuple(roots)))
                node = len(nodes) - 1
            for r in members[1:]:
                parent[find(r)] = find(members[0])
            owner[find(members[0])] = node
        i = j
    return nodes


def signatures(nodes):
    memo = {}

    def sig(x):
        if x not in memo:
            kind, level, kids = nodes[x]
            memo[x] = (kind, level, tuple(sorted(sig(c) for c in kids)))
        return memo[x]

    return Counter(sig(x) for x in range(len(nodes)))


def reduce_saddles(M, saddles, order_one, flip=None):
    """The device rule; `flip` plants a mistake for the negative checks."""
    S = len(saddles)
    first = [S] * M
    for s, (_, cs) in enumerate(saddles):
        for c in cs:
            first[c] = min(first[c], s)

    def owned(c, s):
        return order_one or (first[c] < s if flip != "fresh" else first[c] < s - 1)

    attach = [None] * M
    for s, (_, cs) in enumerate(saddles):
        own = [c for c in cs if owned(c, s)]
        fresh = [c for c in cs if not owned(c, s)]
        target = own[0] if own else fresh[0]
        for c in fresh:
            if c != target:
                attach[c] = target
    rep = []
    for m in range(M):
        x, steps = m, 0
        while attach[x] is not None:
            x, steps = attach[x], steps + 1
            if steps > M:  # a cycle: the device reports an error here
                return None
        rep.append(x)
    edges = []  # (u, v, saddle) in arm order
    for s, (_, cs) in enumerate(saddles):
        own = [c for c in cs if owned(c, s)]
        if not own:
            continue
        seen = {rep[own[0]]}
        for c in own[1:]:
            if rep[c] not in seen:
                seen.add(rep[c])
                edges.append((rep[own[0]], rep[c], s))
    # Boruvka with the device's hooking rule (least edge index per component)
    comp = list(range(M))
    in_forest = [False] * len(edges)
    for _ in range(64):
        best = {}
        for i, (u, v, _) in enumerate(edges):
            cu, cv = comp[u], comp[v]
            if cu != cv:
                best[cu] = min(best.get(cu, i), i)
                best[cv] = min(best.get(cv, i), i)
        hooks = {}
        for m, i in best.items():
            u, v, _ = edges[i]
            other = comp[v] if comp[u] == m else comp[u]
            in_forest[i] = True
            if best.get(other) == i and m < other:
                continue
            hooks[m] = other
        if not hooks:
            break
        for m, o in hooks.items():
            comp[m] = o
        for m in range(M):
            x, steps = m, 0
            while comp[x] != x:
                x, steps = comp[x], steps + 1
                if steps > M:
                    return None
            comp[m] = x
    forest_edges = Counter(edges[i][2] for i in range(len(edges)) if in_forest[i])
    kept = []
    for s, (level, cs) in enumerate(saddles):
        shared = (s > 0 and saddles[s - 1][0] == level) or (s + 1 < S and saddles[s + 1][0] == level)
        threshold = 0 if flip != "merge" else 1
        if not shared and any(owned(c, s) for c in cs) and forest_edges[s] <= threshold:
            continue
        kept.append(s)
    if order_one:
        return M, [saddles[s] for s in kept]
    used = sorted({rep[c] for s in kept for c in saddles[s][1]})
    dense = {r: i for i, r in enumerate(used)}
    return len(used), [(saddles[s][0], [dense[rep[c]] for c in saddles[s][1]]) for s in kept]


def random_case(rng):
    M = rng.randint(1, 24)
    S = rng.randint(1, 48)
    order_one = rng.random() < 0.2
    levels = sorted(rng.randint(0, rng.choice([S, S // 2 + 1, 4, 1])) for _ in range(S))
    saddles = []
    for s in range(S):
        centre = rng.randrange(M)
        saddles.append((levels[s], [min(M - 1, max(0, centre + rng.randint(-3, 3)))
                                    for _ in range(rng.randint(1, 4))]))
    return M, saddles, order_one


def mismatches(seed, trials, flip=None):
    rng = random.Random(seed)
    bad = 0
    for _ in range(trials):
        M, saddles, order_one = random_case(rng)
        want = signatures(contract(M, saddles, order_one))
        got = reduce_saddles(M, saddles, order_one, flip)
        bad += got is None or signatures(contract(got[0], got[1], order_one)) != want
    return bad


def test_reduction_rule_matches_full_contraction():
    assert mismatches(seed=11, trials=1500) == 0


def test_reduction_rule_checks_catch_planted_mistakes():
    # a wrong freshness test and a wrong multifusion threshold both show up
    assert mismatches(seed=12, trials=600, flip="fresh") > 0
    assert mismatches(seed=13, trials=600, flip="merge") > 0
