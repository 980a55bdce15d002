"""CPU oracle package (test infrastructure only: imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
product)."""
