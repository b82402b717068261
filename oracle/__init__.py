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
                                 the reference CPU simulator's own outputs on seeded
                                 8-13-qubit circuits incl. every gate type the golden
                                 circuit lacks (tests/golden/sv_cpu.npz), and brute-force
                                 kron expansion on <= 8 qubits
  statevector.measure / sample   the reference's seeded counts (tests/golden/sv_sample.json),
                                 closed-form P0, GHZ support, chi-square vs |a|^2; the
                                 pyramid rule equals the sequential collapse on random states
  density_matrix.*               the reference DensityMatrixSimulator's outputs with noise
                                 in accumulated and trajectory modes, readout errors and
                                 seeded samples (tests/golden/dm_cpu.npz), reconstructed
                                 golden DM outer(sv, sv*), closed-form depolarizing /
                                 amplitude-damping channels, CPTP checks
  unitary.*                      brute-force product of expanded gate matrices on random
                                 1-7 qubit circuits, golden vector U|0>, kron, unitarity,
                                 circuit * inverse = identity
  gradients.*                    parameter-shift (exact for Pauli rotations), closed form
                                 <Z> = cos(theta) for Ry(theta)|0>, finite differences
  random circuits (core)         the reference generator's own gate lists for the seeded
                                 workloads, compared draw for draw
                                 (tests/golden/gate_lists.json, circuit.py:442-488)
The reference-generated fixtures come from tests/golden/make_reference_fixtures.py,
which imports the reference in the build container only.
"""
