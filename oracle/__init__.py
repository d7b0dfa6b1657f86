"""CPU oracle for the batoid_amd trace hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, as the checker and as the timed CPU baseline.  It binds:

* ``liboracle.so``: the C restatement of the reference (oracle.c);
* ``_ref/libbatoid_ref.so`` (when built): the reference's own C++ core
  compiled from its sources, driven through ref_harness.cpp.

Both take the descriptor records of include/batoid_amd.h with HOST
pointers; ``HostRays`` and ``host_records`` provide those.
"""
from .host import (HostRays, host_records, HostPlan, load, load_ref, ref_available,
                   ORACLE_PATH, REF_PATH)
