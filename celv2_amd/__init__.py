"""celv2_amd: CelV2's decision engine, MI355X-native (HIP for gfx950 behind a C ABI).

The product is libcelv2hip.so (celv2_amd/lib/) with its C++ engine `celv2::HipEngine`
(include/celv2/hip_engine.h) and C ABI (include/celv2hip.h); `celv2_amd.engine` drives it from Python.
"""
