"""ORACLE — test infrastructure only.  CPU restatement of the reference's
H·v / Lanczos path (liboracle.so) and the reference's own operator core
compiled from its sources (_ref/libref.so).  Never imported by the product
path: only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it.
"""
