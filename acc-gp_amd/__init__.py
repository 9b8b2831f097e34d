"""MI355X multilevel k-way graph partitioner (host-side Python mirror).

The product is libgpart_mi355x.so (C ABI in include/gpart_mi355x.h) plus the
gpart_partition_standard / gpart_verify_csr tools under bin/.  This module
binds the C ABI with ctypes; see gpart.py.
"""
from .gpart import *  # noqa: F401,F403
