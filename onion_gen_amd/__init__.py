"""onion-gen candidate-key search on MI355X (gfx950).

The product is the HIP library ``libonion_gen_hip.so`` behind the C ABI of
``include/onion_gen_hip.h``; this package is its Python binding.
"""
__all__ = ["gpu", "filters", "key", "workload"]
