"""KillingFusion's per-frame volumetric hot path on AMD MI355X (gfx950).

The compute runs in libkf.so (HIP kernels behind the C ABI of include/kf_api.h); this
package holds its Python mirror (`api`), the build script (`build`) and the synthetic
workload generators (`scenes`).
"""
from .api import (Context, DisplacementField, IndexedMesh, KillingFusion, KillingParams, SDF, SimpleMesh,  # noqa: F401
                  batch_autotune, batch_deform, compute_displacement_field, dims_from_bounds, energy_backward,
                  energy_forward, energy_gradient_ref, field_from_rigid, host_array, rigid_align)
