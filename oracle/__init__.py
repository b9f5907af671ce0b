"""ORACLE — test infrastructure only.

The CPU restatement of the reference renderer (liboracle.so) and, where the reference checkout is
present, the reference's own CPU build (_ref/libphoton_ref.so). Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg import this package; the product (photon_amd) never does.
"""
