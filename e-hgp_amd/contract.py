"""Contract v2 records of the backend's results, with canonical SHA-256 ids.

Record shapes follow schemas/morsehgp3d-contract-v2.schema.json
(`CriticalEvent` :30, `MergeForest` :818, `VerticalMap` :904); canonical ids
follow the reference's scheme (canonical JSON with sorted keys and no
whitespace, record versions removed, SHA-256 under the domain prefix
"MorseHGP3D/v2/<RecordType>/"; reference/morsehgp3d_oracle/serialization.py
canonical_id and morsehgp3d/src/cpu/contract/canonical_id.cpp:147).

CriticalEvent ids depend only on the mathematical content of an event
(support, interior, shell, exact centre and level), so they are identical to
the reference's for the same cloud.  MergeNode ids of the reference also hash
its equal-level batches, which list the exhaustive Gamma cofaces of that
level; the direct construction has no such list, so this module's node ids
hash the node's order, profile, kind, exact level, child ids and the ids of
the critical events of its level instead, under the record type
"DirectMergeNode".  The exact geometry (centres, levels, shells) is
recomputed here from the binary64 coordinates with rational arithmetic, so
serialising a large catalogue is slow: it is meant for results that are
exported, not for the timed path.
"""

from __future__ import annotations

import hashlib
import json
from fractions import Fraction
from math import comb, gcd

SCHEMA_VERSION = "2.0.0"
CONTRACT_DOMAIN = "MorseHGP3D/v2"
COORDINATE_UNIT = "input_coordinate_unit"
SQUARED_LEVEL_UNIT = "input_coordinate_unit_squared"
BACKEND = "hip_mi355x"


def canonical_json(value) -> str:
    return json.dumps(value, ensure_ascii=False, allow_nan=False, sort_keys=True, separators=(",", ":"))


def _strip_versions(value):
    if isinstance(value, dict):
        return {k: _strip_versions(v) for k, v in value.items() if k != "schema_version"}
    if isinstance(value, (list, tuple)):
        return [_strip_versions(v) for v in value]
    return value


def canonical_id(record_type: str, projection) -> str:
    if not record_type or "/" in record_type:
        raise ValueError("record_type must be a non-empty path component")
    h = hashlib.sha256()
    h.update(f"{CONTRACT_DOMAIN}/{record_type}/".encode("ascii"))
    h.update(canonical_json(_strip_versions(projection)).encode("utf-8"))
    return h.hexdigest()


def exact_level(value) -> dict:
    v = Fraction(value)
    if v < 0:
        raise ValueError("a squared level cannot be negative")
    return {"schema_version": SCHEMA_VERSION, "numerator": str(v.numerator), "denominator": str(v.denominator),
            "unit": SQUARED_LEVEL_UNIT}


def exact_rational3(values) -> dict:
    """Homogeneous coordinates: one positive common denominator, common gcd 1."""
    vals = [Fraction(v) for v in values]
    den = 1
    for v in vals:
        den = den * v.denominator // gcd(den, v.denominator)
    nums = [v.numerator * (den // v.denominator) for v in vals]
    g = den
    for x in nums:
        g = gcd(g, abs(x))
    return {"schema_version": SCHEMA_VERSION, "x_numerator": str(nums[0] // g), "y_numerator": str(nums[1] // g),
            "z_numerator": str(nums[2] // g), "denominator": str(den // g), "unit": COORDINATE_UNIT}


def _sub(a, b):
    return [a[i] - b[i] for i in range(3)]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def circumsphere(pts):
    """Exact centre (in the affine hull) and barycentric coordinates of the
    circumsphere of 1..4 affinely independent points (rational 3-vectors)."""
    m = len(pts)
    p0 = pts[0]
    if m == 1:
        return list(p0), [Fraction(1)]
    a = _sub(pts[1], p0)
    if m == 2:
        return [p0[i] + a[i] / 2 for i in range(3)], [Fraction(1, 2), Fraction(1, 2)]
    b = _sub(pts[2], p0)
    aa, bb, ab = _dot(a, a), _dot(b, b), _dot(a, b)
    if m == 3:
        D = aa * bb - ab * ab
        if D == 0:
            raise ValueError("dependent support")
        l1, l2 = bb * (aa - ab) / (2 * D), aa * (bb - ab) / (2 * D)
        return [p0[i] + l1 * a[i] + l2 * b[i] for i in range(3)], [1 - l1 - l2, l1, l2]
    d = _sub(pts[3], p0)
    bxd, dxa, axb = _cross(b, d), _cross(d, a), _cross(a, b)
    V = _dot(a, bxd)
    if V == 0:
        raise ValueError("dependent support")
    dd = _dot(d, d)
    N = [aa * bxd[i] + bb * dxa[i] + dd * axb[i] for i in range(3)]
    lam = [_dot(N, bxd) / (2 * V * V), _dot(N, dxa) / (2 * V * V), _dot(N, axb) / (2 * V * V)]
    return [p0[i] + N[i] / (2 * V) for i in range(3)], [1 - sum(lam)] + lam


def critical_events(points, catalog, k_max: int, singletons: bool = True) -> list:
    """CriticalEvent records of a catalogue (mhgp.Catalog, canonical PointIds
    of the canonical `points`), singletons included, sorted by event id."""
    P = [[Fraction(float(x)) for x in p] for p in points]
    n = len(P)
    k_eff = min(k_max, n)
    events = []
    if singletons:
        for i in range(n):
            events.append(((i,), (), 1))
    for r in range(len(catalog.size)):
        m = int(catalog.size[r])
        events.append((tuple(int(x) for x in catalog.ids[r, :m]), tuple(catalog.interior[r]), int(catalog.rank[r])))
    out = []
    for sup, interior, rank in events:
        centre, bary = circumsphere([P[i] for i in sup])
        level = sum((centre[t] - P[sup[0]][t]) ** 2 for t in range(3))
        shell = sorted(set(sup) | {x for x in range(n) if x not in sup and x not in interior
                                   and sum((centre[t] - P[x][t]) ** 2 for t in range(3)) == level}) \
            if len(sup) > 1 else list(sup)
        m, s = len(sup), rank
        birth = s if s <= k_eff else None
        saddle = s - 1 if 1 <= s - 1 <= k_eff else None
        roles = []
        if saddle is not None:
            roles.append({"schema_version": SCHEMA_VERSION, "order": saddle, "morse_index": 1,
                          "local_multiplicity": comb(m - 1, s - saddle), "arm_count": len(shell)})
        if birth is not None:
            roles.append({"schema_version": SCHEMA_VERSION, "order": birth, "morse_index": 0,
                          "local_multiplicity": 1, "arm_count": 0})
        roles.sort(key=lambda x: x["order"])
        witness = exact_rational3(centre)
        lev = exact_level(level)
        projection = {"interior_ids": list(interior), "shell_ids": shell, "minimal_support_ids": list(sup),
                      "center_witness_homogeneous": witness, "squared_level_exact": lev}
        out.append({
            "schema_version": SCHEMA_VERSION,
            "event_id": canonical_id("CriticalEvent", projection),
            "closed_rank": s,
            "interior_ids": list(interior),
            "shell_ids": shell,
            "minimal_support_ids": list(sup),
            "center_witness_homogeneous": witness,
            "squared_level_exact": lev,
            "barycentric_signs": [1 if v > 0 else -1 if v < 0 else 0 for v in bary],
            "birth_order": birth,
            "saddle_order": saddle,
            "morse_roles": roles,
            "degeneracy_class": "cospherical_shell" if tuple(shell) != tuple(sup) else "general_position",
            "predicate_status": "certified_exact",
        })
    out.sort(key=lambda r: r["event_id"])
    return out


def forests_and_vertical_maps(tower: dict, level_of_class, events: list, profile: str = "hgp_reduced"):
    """MergeForest records (one per order) and VerticalMap records (one per
    adjacent pair of orders) of a fetched tower (mhgp.Context.tower), with
    exact node levels from level_of_class (level class -> Fraction)."""
    order = [int(x) for x in tower["order"]]
    kinds = {0: "birth", 1: "birth", 2: "merge"}
    off = tower["child_offset"]
    n = len(order)
    events_at = {}
    for e in events:
        lv = Fraction(int(e["squared_level_exact"]["numerator"]), int(e["squared_level_exact"]["denominator"]))
        events_at.setdefault(lv, []).append(e["event_id"])
    ids = [None] * n
    level = [level_of_class[int(c)] for c in tower["level_class"]]
    # children come before their parents inside an order
    for i in range(n):
        kids = sorted(ids[int(c)] for c in tower["children"][off[i]:off[i + 1]])
        projection = {"order": order[i], "profile": profile, "kind": kinds[int(tower["kind"][i])],
                      "squared_level": exact_level(level[i]), "child_ids": kids,
                      "event_ids": sorted(events_at.get(level[i], []))}
        ids[i] = canonical_id("DirectMergeNode", projection)
    parent = {}
    for i in range(n):
        for c in tower["children"][off[i]:off[i + 1]]:
            parent[int(c)] = i
    forests = []
    for k in sorted(set(order)):
        nodes = [i for i in range(n) if order[i] == k]
        recs = [{"schema_version": SCHEMA_VERSION, "node_id": ids[i], "kind": kinds[int(tower["kind"][i])],
                 "squared_level": exact_level(level[i]),
                 "child_ids": sorted(ids[int(c)] for c in tower["children"][off[i]:off[i + 1]])} for i in nodes]
        recs.sort(key=lambda r: r["node_id"])
        roots = sorted(ids[i] for i in nodes if i not in parent)
        projection = {"order": k, "profile": profile, "node_ids": [r["node_id"] for r in recs], "root_ids": roots}
        forests.append({"schema_version": SCHEMA_VERSION, "forest_id": canonical_id("MergeForest", projection),
                        "order": k, "profile": profile, "backend": BACKEND, "nodes": recs, "root_ids": roots,
                        "complete": True})
    vmaps = []
    vert = tower["vertical"]
    for k in sorted(set(order)):
        if k == 1:
            continue
        assign = sorted(({"schema_version": SCHEMA_VERSION, "source_node_id": ids[i],
                          "target_node_id": ids[int(vert[i])], "at_squared_level": exact_level(level[i]),
                          "method": "direct_carrier_descent"} for i in range(n) if order[i] == k),
                        key=lambda a: a["source_node_id"])
        projection = {"source_order": k, "target_order": k - 1, "assignments": assign}
        vmaps.append({"schema_version": SCHEMA_VERSION, "map_id": canonical_id("VerticalMap", projection),
                      "source_order": k, "target_order": k - 1, "assignments": assign, "complete": True})
    return forests, vmaps
