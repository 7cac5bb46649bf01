"""fvk: MI355X (gfx950) implementation of the spray-simulation SIMPLE hot path.

The product is libfvk.so (HIP kernels + C ABI, include/fvk.h); programs_amd.fvk
is the Python host mirror of the reference's solver interface over that ABI.
"""
