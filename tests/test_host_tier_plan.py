"""Host-memory tier planning on the CPU: every gate lands in exactly one
pass, inside that pass's local bits, in an order that keeps each qubit's gate
sequence; replaying the passes slice by slice with the CPU oracle (gather the
slice by the pass's local bits, apply the gates rewritten onto device qubits,
scatter back) reproduces the whole circuit."""
import numpy as np
import pytest

from oracle import oracle as O


def _gather_index(n, local_bits, v):
    """Host indices of slice v, in device order (device bit j = host bit
    local_bits[j])."""
    m = len(local_bits)
    gbits = [b for b in range(n) if b not in local_bits]
    base = 0
    for i, b in enumerate(gbits):
        if v >> i & 1:
            base |= 1 << b
    idx = np.zeros(1 << m, dtype=np.int64)
    for j, b in enumerate(local_bits):
        idx |= ((np.arange(1 << m) >> j) & 1) << b
    return base | idx


@pytest.mark.parametrize("n,m,k", [(10, 7, 3), (11, 8, 5), (9, 9, 4)])
def test_passes_replayed_with_the_oracle(n, m, k):
    from snuqs_amd.circuits import random_gate_list
    from snuqs_amd.host_tier import _to_device
    from snuqs_amd.transpile import plan_host_passes

    rng = np.random.default_rng(n * 100 + m)
    ops = random_gate_list(n, 80, rng)
    passes = plan_host_passes(ops, n, m, k)
    # every gate once; each qubit's gates in program order
    seen = sorted(i for _, idx in passes for i in idx)
    assert seen == list(range(len(ops)))
    flat = [i for _, idx in passes for i in idx]
    for q in range(n):
        mine = [i for i in flat if q in ops[i].targets]
        assert mine == sorted(mine)
    ref = O.run_circuit(n, [(op.matrix(), op.targets) for op in ops])
    st = O.basis_state(n)
    for local_bits, idx in passes:
        assert len(local_bits) == m and local_bits[:k] == list(range(k))
        assert len(set(local_bits)) == m
        where = {b: j for j, b in enumerate(local_bits)}
        for i in idx:
            assert all((n - 1 - q) in where for q in ops[i].targets)
        dev = [_to_device(ops[i], where, n, m) for i in idx]
        for v in range(1 << (n - m)):
            gi = _gather_index(n, local_bits, v)
            sl = st[gi].copy()
            for op in dev:
                O.apply(sl, np.array(op.matrix()), list(op.targets), m)
            st[gi] = sl
    assert np.abs(st - ref).max() < 1e-12
    if m < n:
        assert len(passes) > 1


def test_plan_rejects_gates_wider_than_the_free_bits():
    from snuqs_amd._C.operation import gate_operations as g
    from snuqs_amd.transpile import plan_host_passes

    ops = [g.CCNot([0, 1, 2])]
    with pytest.raises(ValueError):
        plan_host_passes(ops, 10, 8, 7)  # one free bit, the gate needs three
    with pytest.raises(ValueError):
        plan_host_passes(ops, 10, 11, 3)
