"""Gate scheduling before execution (the position of src/braket/snuqs/transpile.py).

* ``pseudo_sort_operations_descending`` -- the reordering the reference applies
  on its CPU/CUDA path (transpile.py:33-89, 184-192): an insertion sort whose
  comparison puts zero-target gates first, keeps gates that share a qubit in
  program order and otherwise puts the gate with the larger min(targets)
  first.  The comparison is not a strict weak order, so the insertion
  procedure itself is kept to reproduce the reference's order exactly.
* ``fuse_operations`` -- new (the reference's optimize_operations is a stub,
  transpile.py:233-234): fusion into dense blocks of at most ``max_qubits``
  qubits by two strategies, keeping the one with fewer blocks:
  - greedy placement: a gate joins an existing block when every block
    created after it is disjoint from the gate's qubits (so moving the gate
    there only commutes it past disjoint gates) and the union stays within
    ``max_qubits``; among those blocks the one that grows least is chosen;
  - frontier: blocks are built one at a time from the gates whose
    predecessors on every qubit are placed ("ready"); a block absorbs every
    ready gate inside its qubit set, then grows by the ready gate whose
    qubits let it absorb the most further gates (one step of lookahead).
    Layered circuits with random pairings fuse into ~14% fewer blocks this
    way; shallow ansatz circuits fuse better greedily.
  Both emit a topological order of the circuit's gate DAG (only gates on
  disjoint qubits are reordered).  Block matrices are accumulated in
  complex128.  A block holding a single gate is emitted as that gate, so
  cheap kernel classes (phase, diagonal, permutation) are kept; multi-gate
  blocks become Unitary gates on the block's qubits in ascending order.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._C.operation import GateOperation
from ._C.operation import gate_operations as g


def _gt(op1, i, op2, j) -> bool:
    t1, t2 = op1.targets, op2.targets
    if not t1 and not t2:
        return i < j
    if not t1:
        return True
    if not t2:
        return False
    if set(t1) & set(t2):
        return i < j
    s1, s2 = op1.sliceable(), op2.sliceable()
    if s1 and s2:
        return i < j
    if s1 != s2:
        return s1
    return min(t1) > min(t2)


def pseudo_sort_operations_descending(operations):
    ops = list(operations)
    for i in range(1, len(ops)):
        cur = ops[i]
        j = i - 1
        while j >= 0 and _gt(cur, i, ops[j], j):
            ops[j + 1] = ops[j]
            j -= 1
        ops[j + 1] = cur
    return ops


def _apply_to_block(mat: np.ndarray, qubits, gate_mat: np.ndarray, targets) -> np.ndarray:
    """Left-multiply the block matrix (local index MSB = qubits[0]) by a gate
    acting on `targets` (subset of `qubits`)."""
    k, t = len(qubits), len(targets)
    pos = [qubits.index(q) for q in targets]
    m = mat.reshape([2] * k + [1 << k])
    gm = gate_mat.reshape([2] * (2 * t))
    out = np.tensordot(gm, m, axes=(list(range(t, 2 * t)), pos))
    rest = [a for a in range(k) if a not in pos]
    # out axes: gate outputs (for pos order), then the remaining row axes, then columns
    order = [0] * (k + 1)
    for a_out, p in enumerate(pos):
        order[p] = a_out
    for i, a in enumerate(rest):
        order[a] = t + i
    order[k] = k
    return np.transpose(out, order).reshape(1 << k, 1 << k)


@dataclass
class _Block:
    qubits: list
    gates: list = field(default_factory=list)


def _greedy_blocks(ops, max_qubits):
    blocks: list[_Block] = []
    last: dict[int, int] = {}
    for op in ops:
        qs = set(op.targets)
        lo = max((last[x] for x in qs if x in last), default=0)
        best, best_grow = None, None
        for b in range(len(blocks) - 1, lo - 1, -1):
            union = qs | set(blocks[b].qubits)
            if len(union) > max_qubits:
                continue
            grow = len(union) - len(blocks[b].qubits)
            if best is None or grow < best_grow:
                best, best_grow = b, grow
            if grow == 0:
                break
        if best is None:
            blocks.append(_Block(sorted(qs), [op]))
            best = len(blocks) - 1
        else:
            blk = blocks[best]
            blk.qubits = sorted(set(blk.qubits) | qs)
            blk.gates.append(op)
        for x in qs:
            last[x] = max(last.get(x, -1), best)
    return blocks


def _frontier_blocks(ops, max_qubits):
    qs = [frozenset(op.targets) for op in ops]
    perq: dict[int, list] = {}
    for i, q in enumerate(qs):
        for x in q:
            perq.setdefault(x, []).append(i)
    ptr = {x: 0 for x in perq}

    def head(p, x):
        return perq[x][p[x]] if p[x] < len(perq[x]) else None

    def ready(p, i):
        return all(perq[x][p[x]] == i for x in qs[i])

    def take(p, i, out=None):
        for x in qs[i]:
            p[x] += 1
        if out is not None:
            out.append(i)

    def absorb(p, qset, out=None):
        """Place every ready gate inside qset; returns how many."""
        n, progress = 0, True
        while progress:
            progress = False
            for x in qset:
                i = head(p, x)
                if i is not None and qs[i] <= qset and ready(p, i):
                    take(p, i, out)
                    n += 1
                    progress = True
        return n

    blocks: list[_Block] = []
    remaining = len(ops)
    while remaining:
        heads = {head(ptr, x) for x in perq} - {None}
        seed = min(i for i in heads if ready(ptr, i))
        qset, gates = set(qs[seed]), []
        take(ptr, seed, gates)
        absorb(ptr, qset, gates)
        while True:
            best, best_score = None, None
            for x in list(qset):
                i = head(ptr, x)
                if i is None or not ready(ptr, i) or len(qset | qs[i]) > max_qubits:
                    continue
                trial = dict(ptr)
                take(trial, i)
                score = (absorb(trial, qset | qs[i]), -len(qset | qs[i]))
                if best is None or score > best_score:
                    best, best_score = i, score
            if best is None:
                break
            qset |= qs[best]
            take(ptr, best, gates)
            absorb(ptr, qset, gates)
        blocks.append(_Block(sorted(qset), [ops[i] for i in gates]))
        remaining -= len(gates)
    return blocks


def fusion_plan(operations, max_qubits: int = 5):
    """The grouping ``fuse_operations`` uses, as gate indices: a list of
    (block qubits, [indices of its gates in order]) and the indices of the
    zero-target gates (folded into one global phase).  It depends on the
    gates' targets only, so circuits that differ in parameters alone share
    it (parameter batches)."""
    ops = list(operations)
    idx = [i for i, op in enumerate(ops) if list(op.targets)]
    phases = [i for i, op in enumerate(ops) if not list(op.targets)]
    sub = [ops[i] for i in idx]
    blocks = _greedy_blocks(sub, max_qubits)
    if sub:
        alt = _frontier_blocks(sub, max_qubits)
        if len(alt) < len(blocks):
            blocks = alt
    pos = {id(op): i for op, i in zip(sub, idx)}
    return [(list(b.qubits), [pos[id(op)] for op in b.gates]) for b in blocks], phases


def fuse_with_plan(operations, plan):
    """Apply a ``fusion_plan`` to ``operations`` (the circuit it was made
    from, or one with the same targets and other parameters)."""
    ops = list(operations)
    blocks, phases = plan
    out = []
    for qubits, members in blocks:
        if len(members) == 1:
            out.append(ops[members[0]])
            continue
        mat = np.eye(1 << len(qubits), dtype=np.complex128)
        for i in members:
            mat = _apply_to_block(mat, qubits, np.array(ops[i].matrix()), list(ops[i].targets))
        out.append(g.Unitary(qubits, mat))
    if phases:
        ph = complex(np.prod([np.array(ops[i].matrix())[0, 0] for i in phases]))
        out.append(g.GPhase([], float(np.angle(ph))))
    return out


def fuse_operations(operations, max_qubits: int = 5):
    """Fuse gates into blocks of at most max_qubits qubits (see module doc)."""
    return fuse_with_plan(operations, fusion_plan(operations, max_qubits))


def plan_host_passes(operations, qubit_count: int, local_qubits: int, contiguous_bits: int):
    """Passes of the host-memory tier (the position of the reference's
    localize_operations / partition_operations, transpile.py:209-298, which
    insert Swap gates and split the circuit at every non-local gate).

    A pass makes ``local_qubits`` index bits local: the lowest
    ``contiguous_bits`` bits (each slice is gathered in contiguous chunks of
    2^contiguous_bits amplitudes) plus up to ``local_qubits -
    contiguous_bits`` others, chosen greedily: gates are taken in an order
    that respects every qubit's gate sequence, each gate whose high bits fit
    the pass's set is taken (gates that do not grow the set first), and a
    new pass starts when no ready gate fits.  Returns a list of
    (local_bits, [operation indices]) with local_bits in device order
    (0..contiguous_bits-1, then the chosen bits ascending, topped up with
    unused bits to local_qubits bits)."""
    n, m, k = qubit_count, local_qubits, contiguous_bits
    if not 1 <= k <= m <= n:
        raise ValueError("need 1 <= contiguous_bits <= local_qubits <= qubit_count")
    ops = list(operations)
    high = [frozenset(b for b in (n - 1 - q for q in op.targets) if b >= k) for op in ops]
    for i, h in enumerate(high):
        if len(h) > m - k:
            raise ValueError(f"gate {i} has {len(h)} targets above bit {k}: more than the "
                             f"{m - k} free local bits")
    perq: dict[int, list] = {}
    for i, op in enumerate(ops):
        for q in op.targets:
            perq.setdefault(q, []).append(i)
    ptr = {q: 0 for q in perq}
    placed = [False] * len(ops)
    order = list(range(len(ops)))
    passes = []
    left = len(ops)
    while left:
        hset: set = set()
        taken = []
        progress = True
        while progress:
            progress = False
            best = None
            for i in order:
                if placed[i] or any(perq[q][ptr[q]] != i for q in ops[i].targets):
                    continue
                grow = len(high[i] - hset)
                if len(hset) + grow > m - k:
                    continue
                if grow == 0:  # absorb at once, in program order
                    best = i
                    break
                if best is None:
                    best = i
            if best is not None:
                placed[best] = True
                taken.append(best)
                hset |= high[best]
                for q in ops[best].targets:
                    ptr[q] += 1
                left -= 1
                progress = True
        if not taken:
            raise RuntimeError("host-tier planning made no progress")
        bits = sorted(hset)
        fill = [b for b in range(k, n) if b not in hset][: m - k - len(bits)]
        passes.append((list(range(k)) + sorted(bits + fill), taken))
        order = [i for i in order if not placed[i]]
    return passes
