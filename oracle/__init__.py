"""CPU checker of the matched filter -- test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the apogee_amd package.
"""
