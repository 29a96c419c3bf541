"""Oracle: CPU restatement of the reference hot path.  Test infrastructure only;
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
