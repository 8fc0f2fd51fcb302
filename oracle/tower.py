"""Small-n tower oracles in exact rationals (test infrastructure only).

Two independent constructions of the per-order merge forests:

* ``gamma_forests``: the exhaustive Gamma_k definition.  Vertices are all
  k-subsets at their miniball level, relations are all (k+1)-subsets at their
  miniball level, and equal-level batches are contracted simultaneously
  (reference/morsehgp3d_oracle/gamma.py build_gamma_filtration and
  hierarchy.py build_merge_forest).  Exponential; n <= 10.
* ``direct_forests``: the direct Morse construction used by the product
  (reference morsehgp3d/src/cpu/hierarchy/direct_morse_forest_journal.cpp:1042):
  order-k minima are rank-k critical spheres, order-k saddles are rank-(k+1)
  spheres whose strict arms (closed label minus one shell point) are resolved
  to their minimum carriers by exact facet descent, and saddle groups with no
  prior reduced root create a reduced birth, two or more roots a multifusion.

Both use the exact predicates of ``liboracle.so`` on the cloud loaded by the
last ``oracle.catalog`` call.
"""

from __future__ import annotations

from fractions import Fraction
from itertools import combinations

from . import catalog, support_info, support_side


class DescentError(RuntimeError):
    pass


class _Geometry:
    def __init__(self, n: int):
        self.n = n
        self._ball = {}
        self._info = {}

    def info(self, sup):
        if sup not in self._info:
            self._info[sup] = support_info(sup)
        return self._info[sup]

    def miniball(self, label):
        """(level, support) of the minimum enclosing ball of ``label``."""
        label = tuple(sorted(label))
        if label in self._ball:
            return self._ball[label]
        cands = []
        for m in range(1, min(4, len(label)) + 1):
            for sup in combinations(label, m):
                inf = self.info(sup)
                if inf is None:
                    continue
                cands.append((inf[1], m, sup))
        cands.sort()
        for level, _, sup in cands:
            if all(support_side(sup, x) <= 0 for x in label if x not in sup):
                self._ball[label] = (level, sup)
                return self._ball[label]
        raise DescentError("no enclosing circumball found")

    def closed_ball(self, sup):
        return tuple(x for x in range(self.n) if x in sup or support_side(sup, x) <= 0)


def _new_node(nodes, order, kind, level, children, covered=None):
    nodes.append({"order": order, "kind": kind, "level": level,
                  "children": tuple(sorted(children)), "covered": covered})
    return len(nodes) - 1


def gamma_forests(points, k_max: int, profile: str = "hgp_reduced", vertical: bool = False):
    """Exhaustive Gamma_k forests for orders 1..min(k_max, n).  With
    vertical=True also returns the vertical maps: for every node of order
    k + 1 >= 2 (index local to its order) the node of order k onto which its
    component maps at the node's own level, closed cut (reference
    reference/morsehgp3d_oracle/hierarchy.py build_vertical_map_family and
    serialization.py _serialize_vertical_maps: the image of a facet F of
    Gamma_{k+1} is the Gamma_k component of its k-subsets, which the coface
    F joins at its own level)."""
    n = len(points)
    catalog(points, k_max, max_support=1)  # loads the exact cloud
    geo = _Geometry(n)
    forests = {}
    history = {}   # order -> facet -> [(level, root)] in level order
    witness = {}   # order -> node -> a facet of its component at its level
    for k in range(1, min(k_max, n) + 1):
        hist = history.setdefault(k, {})
        wit = witness.setdefault(k, {})
        facets = {f: geo.miniball(f)[0] for f in combinations(range(n), k)}
        cofaces = {}
        if k < n:
            for c in combinations(range(n), k + 1):
                cofaces[c] = geo.miniball(c)[0]
        births = {}
        if profile == "full_pi0" or k == 1:
            for f, lev in facets.items():
                births.setdefault(lev, []).append(f)
        rels = {}
        for c, lev in cofaces.items():
            rels.setdefault(lev, []).append(c)
        nodes = []
        comp_facets = {}  # root -> set of facets
        root_of = {}      # facet -> root
        for level in sorted(set(births) | set(rels)):
            parent = {}

            def tok(f):
                t = ("r", root_of[f]) if f in root_of else ("f", f)
                parent.setdefault(t, t)
                return t

            def find(t):
                while parent[t] != t:
                    parent[t] = parent[parent[t]]
                    t = parent[t]
                return t

            for f in births.get(level, []):
                tok(f)
            for c in rels.get(level, []):
                ts = [tok(tuple(x for x in c if x != r)) for r in c]
                for t in ts[1:]:
                    a, b = find(ts[0]), find(t)
                    if a != b:
                        parent[b] = a
            groups = {}
            for t in list(parent):
                groups.setdefault(find(t), []).append(t)
            for members in groups.values():
                roots = sorted(t[1] for t in members if t[0] == "r")
                new = [t[1] for t in members if t[0] == "f"]
                allf = set(new)
                for r in roots:
                    allf |= comp_facets[r]
                covered = tuple(sorted({x for f in allf for x in f}))
                if not roots:
                    root = _new_node(nodes, k, "birth", level, (), covered)
                elif len(roots) == 1:
                    root = roots[0]
                else:
                    root = _new_node(nodes, k, "merge", level, roots, covered)
                for r in roots:
                    comp_facets.pop(r, None)
                comp_facets[root] = allf
                for f in allf:
                    root_of[f] = root
                    hist.setdefault(f, []).append((level, root))
                if root not in wit:
                    wit[root] = min(allf)
        forests[k] = nodes
    if not vertical:
        return forests

    def root_at(k, f, level):
        out = None
        for lev, root in history[k].get(f, []):
            if lev <= level:
                out = root
        return out

    maps = {}
    for k in range(1, min(k_max, n)):
        targets = []
        for i, nd in enumerate(forests[k + 1]):
            face = witness[k + 1][i]
            t = root_at(k, tuple(face[:-1]), nd["level"])
            if t is None:
                raise DescentError("a vertical image outside the order-k forest")
            # every k-subset of the facet lands in the same component
            assert all(root_at(k, tuple(x for x in face if x != r), nd["level"]) == t for r in face)
            targets.append(t)
        maps[k + 1] = targets
    return forests, maps


def _descend(geo, label, minima, stats):
    """Follow strictly decreasing miniballs from a facet to a minimum."""
    cur = tuple(sorted(label))
    for _ in range(10000):
        if cur in minima:
            return cur
        level, sup = geo.miniball(cur)
        ball = geo.closed_ball(sup)
        extra = [x for x in ball if x not in cur]
        if not extra:
            raise DescentError(f"facet {cur} is a ball minimum but not a catalogued event")
        x = min(extra)
        stats["steps"] += 1
        nxt = None
        for y in sorted(sup):
            cand = tuple(sorted((set(cur) | {x}) - {y}))
            if geo.miniball(cand)[0] < level:
                nxt = cand
                break
        if nxt is None:
            raise DescentError(f"no strict descent from {cur}")
        cur = nxt
    raise DescentError("descent did not terminate")


def direct_forests(points, k_max: int, events=None) -> dict:
    """Direct Morse hgp_reduced forests from the exhaustive catalogue."""
    n = len(points)
    if events is None:
        events = catalog(points, k_max)
    else:
        catalog(points, k_max, max_support=1)
    if any(e.degenerate for e in events):
        raise DescentError("degenerate catalogue: extra-shell supports present")
    geo = _Geometry(n)
    keff = min(k_max, n)
    forests = {}
    stats = {"steps": 0}
    by_rank = {}
    for e in events:
        by_rank.setdefault(e.rank, []).append(e)
    for k in range(1, keff + 1):
        nodes = []
        minima = {}
        if k == 1:
            for p in range(n):
                minima[(p,)] = _new_node(nodes, 1, "birth", Fraction(0), ())
        else:
            for e in by_rank.get(k, []):
                minima[tuple(sorted(e.support + e.interior))] = None
        # carrier DSU over minima; each class may own a root node
        parent = {m: m for m in minima}
        owner = {m: minima[m] for m in minima}

        def find(m):
            while parent[m] != m:
                parent[m] = parent[parent[m]]
                m = parent[m]
            return m

        saddles = {}
        for e in by_rank.get(k + 1, []):
            if len(e.support) < 2:
                continue
            saddles.setdefault(e.level, []).append(e)
        for level in sorted(saddles):
            bp = {}

            def bfind(t):
                while bp[t] != t:
                    bp[t] = bp[bp[t]]
                    t = bp[t]
                return t

            for e in saddles[level]:
                closed = set(e.support) | set(e.interior)
                cls = []
                for u in e.support:
                    arm = tuple(sorted(closed - {u}))
                    carrier = _descend(geo, arm, minima, stats)
                    c = find(carrier)
                    bp.setdefault(c, c)
                    cls.append(c)
                for c in cls[1:]:
                    a, b = bfind(cls[0]), bfind(c)
                    if a != b:
                        bp[b] = a
            groups = {}
            for c in bp:
                groups.setdefault(bfind(c), []).append(c)
            for members in groups.values():
                roots = sorted({owner[c] for c in members if owner[c] is not None})
                if not roots:
                    root = _new_node(nodes, k, "birth", level, ())
                elif len(roots) == 1:
                    root = roots[0]
                else:
                    root = _new_node(nodes, k, "merge", level, roots)
                base = members[0]
                for c in members[1:]:
                    parent[find(c)] = find(base)
                owner[find(base)] = root
        forests[k] = nodes
    forests["descent_steps"] = stats["steps"]
    return forests


def signatures(nodes, with_covered: bool = False) -> list:
    """Canonical multiset of node signatures (order, kind, level, children)."""
    memo = {}

    def sig(i):
        if i not in memo:
            nd = nodes[i]
            kids = tuple(sorted(sig(c) for c in nd["children"]))
            cov = nd["covered"] if with_covered else None
            memo[i] = (nd["order"], nd["kind"], nd["level"], kids, cov)
        return memo[i]

    return sorted(sig(i) for i in range(len(nodes)))
