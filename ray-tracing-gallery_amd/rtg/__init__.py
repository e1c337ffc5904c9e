"""MI355X trace step of expenses/ray-tracing-gallery: Python host side.

abi     ctypes mirrors of the C ABI structs (include/rtg.h)
lib     binding of librtg.so (the HIP library; no fallback)
device  device-resident scenes: BLAS/TLAS builds, refits and traces
scenes  seeded synthetic workloads
"""
