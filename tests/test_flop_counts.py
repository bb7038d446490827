"""The roofline's algorithmic operation counts (bench.py) are the oracle's:
the restatement of the reference's geometry, Hamiltonian flow, DP5 attempt
and per-step invariants instantiated with an operation-counting scalar."""
import ctypes as C

import numpy as np

import oracle

import bench


def _counts():
    o = np.zeros(6, np.int64)
    oracle.lib().oracle_flop_counts(o.ctypes.data_as(C.c_void_p))
    return o


def test_counts_are_pinned():
    per_rhs, att_acc, att_rej, loop, acc_step, rej_att = (int(v) for v in _counts())
    assert att_acc > 0 and att_rej > 0
    assert per_rhs == bench.ALG_FLOPS_PER_RHS
    assert acc_step == bench.ALG_FLOPS_PER_ACCEPTED_STEP == 6 * per_rhs + att_acc + loop
    assert rej_att == bench.ALG_FLOPS_PER_REJECTED_ATTEMPT == 6 * per_rhs + att_rej


def test_accepted_step_adds_dense_output():
    """An accepted attempt differs from a rejected one by the dense-output
    coefficients (8 components x (c2..c5): 1 + 2 + 3 + 12 operations) and the
    controller's memory update."""
    per_rhs, att_acc, att_rej, *_ = (int(v) for v in _counts())
    assert att_acc - att_rej == 8 * 18 + 0  # max(err, 1e-4) is free
