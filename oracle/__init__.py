"""CPU oracle for mocat_amd: test infrastructure, never product code.

Plain NumPy (FP64, or FP32 where a step must mirror single-precision rounding) restatements of
the reference algorithms on the accelerated path. Each function cites the reference file:line it
follows. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only as the checker or the timed CPU baseline; the mocat_amd package never does.

Pinning (what checks each part of the oracle independently of its own formulas) is listed in
each module's header. A function without such a check says "parity unpinned".
"""
