"""MI355X-native SYCLDB physical-operator hot path.

``sycldb_amd`` holds the HIP kernels and C++ engine (csrc/, built into
libsycldb_hip.so) and their Python front ends:

* ``ops``    -- the reference's operator free functions on device tensors
* ``engine`` -- resident tables and plan execution (TransientTable model)
* ``plans``  -- hand-built plans of the SSB templates
* ``schema`` -- table layout and the seeded data specification
"""
__all__ = ["schema", "plans"]
