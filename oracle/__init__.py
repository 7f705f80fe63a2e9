"""CPU oracle for lettuce_amd's parity tests.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the lettuce_amd package.
``lettuce_torch`` restates lettuce's torch time step on the CPU.
"""
