"""CPU oracle for the MI355X tracer -- TEST INFRASTRUCTURE ONLY.

A C++ restatement of the reference tracer's still path (see oracle.h for the
per-part pin status). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may load it, and only as the checker or the CPU baseline;
the product (pfx_amd / libpfxhip.so) never routes through it.
"""
from .oracle import Oracle, build, lib  # noqa: F401
