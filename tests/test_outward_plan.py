"""Link numbering of the random-tree workload (mr_builder.cpp
depth_first_order). The response stage's outward pass visits links in index
order and keeps a link's spatial acceleration in an LDS slot until its last
(highest-numbered) child has read it, so the slot count, and with it the LDS
each world needs, follows from the numbering. The builder numbers every tree
depth-first with the child that needs the most slots last.

Checked on the 8192 trees of the heterogeneous benchmark: parents precede
their children, the index-order slot plan (the rule the world uses,
restated here) never drops a value a later child reads and needs at most 3
slots (numbering the same trees in generation order needed 11), and the
DoF range stays 6-40."""

import numpy as np

import metalrobo as mr

INVALID = 0xFFFFFFFF


def index_order_slots(parent):
    """Greedy slot plan over links in index order; returns (slot per link, peak), and
    checks while replaying that every parent's value is still in its slot when read."""
    n = len(parent)
    last = [None] * n
    for i in range(1, n):
        last[parent[i]] = i
    used, peak, slot, owner = set(), 0, [None] * n, {}

    def take():
        s = 0
        while s in used:
            s += 1
        used.add(s)
        return s

    if last[0] is not None:
        slot[0] = take()
        owner[slot[0]] = 0
    peak = len(used)
    for i in range(1, n):
        p = parent[i]
        assert owner.get(slot[p]) == p, (i, p)  # the parent's acceleration is still there
        if last[p] == i:
            used.discard(slot[p])
            owner.pop(slot[p])
        if last[i] is not None:
            slot[i] = take()
            owner[slot[i]] = i
        peak = max(peak, len(used))
    return slot, peak


def test_random_trees_are_numbered_for_few_outward_slots():
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_RANDOM_TREES, 8192, 2024)
    headers, bodies = model.models(), model.bodies()
    worst = 0
    dofs = []
    for h in headers:
        n = int(h["link_count"])
        b = bodies[int(h["body_offset"]):int(h["body_offset"]) + n]
        parent = [0] + [int(x) for x in b["parent"][1:]]
        assert all(parent[i] < i for i in range(1, n))
        assert all(int(b[i]["v_index"]) == i - 1 for i in range(1, n))
        _, peak = index_order_slots(parent)
        worst = max(worst, peak)
        dofs.append(int(h["nv"]))
    assert worst <= 3, worst
    assert min(dofs) == 6 and max(dofs) == 40
