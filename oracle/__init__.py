"""CPU restatement of the SB solver loop -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package; it is the checker of the HIP path and never part of it."""
