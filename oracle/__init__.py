"""The FP64 CPU oracle of the depth-averaged MPM step: test infrastructure
only (imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, as the checker; never by the product)."""
