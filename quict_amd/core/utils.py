"""Small dense-matrix helpers used by the front end and by tests.

``matrix_product_to_circuit`` has the semantics of
QuICT/core/utils/utils.py:8-63 (expand a gate matrix acting on ``gate_args``
into the full ``max_q``-qubit space; qubit 0 is the most significant bit,
gate argument 0 is the most significant bit of the gate-matrix index), but is
computed by index arithmetic over numpy arrays rather than the reference's
element-by-element double loop.
"""
from typing import Union

import numpy as np


def matrix_product_to_circuit(gate_matrix: np.ndarray, gate_args: Union[int, list], max_q: int,
                              min_q: int = 0) -> np.ndarray:
    if isinstance(gate_args, (int, np.integer)):
        gate_args = [int(gate_args)]
    nq = max_q - min_q
    k = len(gate_args)
    assert gate_matrix.shape == (1 << k, 1 << k)
    for a in gate_args:
        assert min_q <= a < max_q
    n = 1 << nq
    idx = np.arange(n, dtype=np.int64)
    # sub-index of each full index inside the gate's space
    sub = np.zeros(n, dtype=np.int64)
    gate_mask = 0
    for t, a in enumerate(gate_args):
        bit = nq - 1 - (a - min_q)
        sub |= ((idx >> bit) & 1) << (k - 1 - t)
        gate_mask |= 1 << bit
    rest = idx & ~gate_mask
    same = rest[:, None] == rest[None, :]
    out = np.where(same, gate_matrix[sub[:, None], sub[None, :]], 0).astype(gate_matrix.dtype)
    return out


def unique_id_generator() -> str:
    import uuid
    return uuid.uuid4().hex[:8]
