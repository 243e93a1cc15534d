"""MI355X (gfx950) implementation of fastwarp's GPU registration hot path.

`fastwarp_amd.backend.HipBackend` mirrors the reference's `Backend`/`LinearEngine` stage interface
over the C ABI in include/fastwarp_hip.h; `fastwarp_amd.register` runs whole registrations.
"""
