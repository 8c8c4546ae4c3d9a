"""CPU oracle of the phase-linking hot path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline and accuracy legs, as the checker. The product never imports it.
"""
