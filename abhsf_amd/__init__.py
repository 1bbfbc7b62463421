"""abhsf_amd -- MI355X-native ABHSF SpMV (adaptive-blocking hierarchical storage).

The compute path is libabhsf.so (hand-written gfx950 HIP kernels behind the C
ABI of include/abhsf_c.h); this package is the host-side mirror of the
reference's matrix-class interface plus the benchmark/tuning harness.
"""
from .matrix import AbhsfBatch, AbhsfMatrix, DeltaMatrix, matrix_properties, matrix_type_t, matrix_symmetry_t  # noqa: F401
from .capi import AbhsfError  # noqa: F401

__all__ = ["AbhsfBatch", "AbhsfMatrix", "DeltaMatrix", "matrix_properties", "matrix_type_t", "matrix_symmetry_t", "AbhsfError"]
