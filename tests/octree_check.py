"""Independent checks for octree topology (test helper, no reference code).

reference_octree builds the node set from its definition rather than from
the level-wise flag rule of tree.cu:50-78: the root holds every body; a node
at level L < 21 holding two or more bodies has one child per distinct
level-(L+1) cell among its bodies.  Nodes are then ordered level-major and,
within a level, by first body.
"""
import numpy as np

from oracle.binding import NODE_DTYPE


def reference_octree(keys):
    keys = np.asarray(keys, dtype=np.uint64)
    n = keys.shape[0]
    # (level, first, count, path, parent_tmp_id)
    recs = [(0, 0, n, 0, -1)]
    frontier = [0]
    while frontier:
        nxt = []
        for rid in frontier:
            level, first, count, path, _ = recs[rid]
            if count < 2 or level == 21:
                continue
            sh = np.uint64(63 - 3 * (level + 1))
            cells = keys[first:first + count] >> sh
            starts = np.flatnonzero(np.r_[True, cells[1:] != cells[:-1]])
            ends = np.r_[starts[1:], count]
            for s, e in zip(starts, ends):
                recs.append((level + 1, first + int(s), int(e - s), int(cells[s]), rid))
                nxt.append(len(recs) - 1)
        frontier = nxt
    order = sorted(range(len(recs)), key=lambda r: (recs[r][0], recs[r][1]))
    new_id = {old: i for i, old in enumerate(order)}
    out = np.zeros(len(recs), NODE_DTYPE)
    out["child"] = -1
    for i, old in enumerate(order):
        level, first, count, path, parent = recs[old]
        out[i]["level"] = level
        out[i]["first"] = first
        out[i]["count"] = count
        out[i]["path"] = path
        out[i]["parent"] = new_id[parent] if parent >= 0 else -1
        if parent >= 0:
            out[new_id[parent]]["child"][path & 7] = i
    return out


def check_invariants(nodes, n):
    """Structural invariants of a node array (child ranges partition the parent)."""
    assert nodes[0]["count"] == n
    leaf_bodies = 0
    for i, nd in enumerate(nodes):
        kids = [c for c in nd["child"] if c != -1]
        if not kids:
            leaf_bodies += int(nd["count"])
            continue
        # children partition the parent's body range, in octant order
        pos = int(nd["first"])
        for slot, c in enumerate(nd["child"]):
            if c == -1:
                continue
            ch = nodes[c]
            assert ch["parent"] == i
            assert int(ch["path"]) & 7 == slot
            assert int(ch["path"]) >> 3 == int(nd["path"])
            assert ch["level"] == nd["level"] + 1
            assert ch["first"] == pos
            pos += int(ch["count"])
        assert pos == int(nd["first"]) + int(nd["count"])
    assert leaf_bodies == n
