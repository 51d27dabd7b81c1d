"""MI355X-native k-clique densest subgraph (k-CDS) solver.

Modules:
  kcds  -- ctypes front end of libkcds.so (HIP kernels behind include/kcds.h)
  gen   -- seeded synthetic workloads (libkcdsgen.so)
"""
