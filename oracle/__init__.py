"""CPU oracle for the SYCLDB operator hot path.

TEST INFRASTRUCTURE ONLY: the checker the HIP product is compared against.
Imported by tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline
leg; never by the product package.
"""
