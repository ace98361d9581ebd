"""CPU oracle (test infrastructure only): FP64 NumPy restatement of jsolver's
multigrid path.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, as the checker; the product never imports it."""
