"""FP64 CPU oracle: a serial restatement of the reference's update formulae.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may load it, and only as the checker or the
timed CPU baseline -- never as a compute path of the product.
"""
from .binding import load, Oracle  # noqa: F401
