"""CPU oracle (test infrastructure only): a restatement of KronFreeCircuits.jl's
hot-path algorithm used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  Product code never imports it."""
