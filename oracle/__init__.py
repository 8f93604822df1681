"""FP64 CPU oracle of the ONSAS.jl nonlinear static path.

TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py.  The product (onsas.jl_amd/) never
imports, links or runs anything here.
"""
