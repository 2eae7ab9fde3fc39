"""CPU oracle: an FP64 numpy restatement of the SaclayMocks hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the checker
(or the timed CPU baseline); the product path in saclaymocks_amd never calls
it.  Every function cites the reference file:line it restates and the test
that pins it (see each module's header).
"""
