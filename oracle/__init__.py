"""Test infrastructure only: the CPU oracle (serial restatement of the
reference solver) and the reference build recipe.  Imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the
product package."""
