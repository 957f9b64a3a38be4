"""MI355X-native UPPE/GNLSE propagation backend for solprop's native seam.

The hot path is the HIP library ``libsolprop_mi355x.so`` (gfx950), reached through the
C ABI of ``include/solprop_mi355x.h``; :mod:`solprop_amd.native` is the host layer that
mirrors the reference's native adapter, and :mod:`solprop_amd.problem` the problem
record and the workload builders.
"""

__version__ = "0.1.0"
