"""CPU oracle (test infrastructure only): a C restatement of the reference
algorithms, loaded by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py.  Never imported by the svfsi_amd package."""
