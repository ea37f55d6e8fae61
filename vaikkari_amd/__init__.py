"""vaikkari_amd -- MI355X (gfx950) Gross-Pitaevskii stepper on the BCC dual mesh.

The compute lives in libvaikkari_gpe.so (HIP kernels + C ABI,
include/vaikkari_gpe.h); this package is the host-side mirror of the
reference's solver interface over that ABI.
"""
from .solver import (  # noqa: F401
    Solver,
    TrapRZ,
    VortexState,
    as_complex,
    bcc_positions,
    bcc_stencil,
    bcc_stencil_derived,
    centered_origin,
    copy_bandwidth,
    field_norm,
    gaussian,
    integrate_in_time,
    integrateInTime,
    integrateInTimeWithOutput,
    leapfrog_coefs,
    node_measure,
    node_positions,
    nodes_of,
    phase_shift,
    reference_dims,
    reference_setup,
    reference_state_recipe,
    stream_bandwidth,
    thomas_fermi,
    trap_potential,
    variants,
    vortex_line,
    vortex_ring,
    write_bmp,
    write_volume_map,
)
from .primal import (  # noqa: F401
    PrimalSolver,
    primal_centered_origin,
    primal_coefficients,
    primal_positions,
    primal_tiles,
    primal_trap_potential,
    site_measure,
    sites_of,
)
