"""CPU oracle for the circuit-simulation path — TEST INFRASTRUCTURE ONLY.

This package restates the reference's CPU simulator algorithms (numpy, plus a
C restatement under ``oracle/csrc`` used as the timed CPU baseline).  Only the
test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it, and only as the checker; the product path
(``quict_amd.simulation`` / ``quict_amd.gradient``) never calls it and fails
loudly when the HIP library is missing.

How each function is pinned (so that a dropped term, a wrong sign or a
transposed operand fails a check):
  statevector.apply_gate / run   golden 10-qubit vector shipped with the reference
                                 (unit_test/simulation/data/state_vector.npy) at 1e-12,
                                 and brute-force kron expansion on <= 8 qubits
  statevector.measure / sample   closed-form P0, GHZ support, chi-square vs |a|^2,
                                 hand-worked seeded cases
  density_matrix.*               reconstructed golden DM outer(sv, sv*), literal dense
                                 U rho U^dag algorithm on <= 6 qubits, closed-form
                                 depolarizing / amplitude-damping channels, CPTP checks
  unitary.*                      brute-force product of expanded gate matrices on random
                                 1-7 qubit circuits, golden vector U|0>, kron, unitarity,
                                 circuit * inverse = identity
  gradients.*                    parameter-shift (exact for Pauli rotations), closed form
                                 <Z> = cos(theta) for Ry(theta)|0>, finite differences
  random circuits (core)         parity unpinned against the reference generator byte for
                                 byte (it cannot be imported here); the draw order is
                                 restated from QuICT/core/circuit/circuit.py:442-488
"""
