"""MI355X (gfx950) implementation of the SCFT.c iteration loop.

The product is the C ABI library libscft.so (csrc/) and the C driver
scft_run (csrc/scft_run.c).  This Python package only holds the ctypes
binding (abi.py), the workload definitions (configs.py) and the batched
serving front end (serve.py) used by the tests and bench.py.
"""
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
if PKG_DIR not in sys.path:
    sys.path.insert(0, PKG_DIR)
