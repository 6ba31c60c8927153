"""MI355X (gfx950) drop-in for HaissinskiJulia's multi-turn longitudinal
tracking loop: a C-ABI library of hand-written HIP kernels
(``libhaissinski_mi355x.so``, declared in ``include/haissinski_mi355x.h``)
and this host-side mirror of the reference's API."""
from .api import (  # noqa: F401
    ELECTRON_CHARGE,
    MASS_ELECTRON,
    SPEED_LIGHT,
    Coordinates,
    DeviceBunch,
    EvolveResult,
    MI355X,
    Particles,
    SimulationBuffers,
    SimulationParameters,
    TrackingOptions,
    apply_wakefield_inplace,
    batch_evolve,
    bin_grid,
    calc_rf_factor,
    calculate_wake_function,
    compute_global_std,
    create_simulation_buffers,
    delta,
    draw_normals,
    drift,
    generate_particles,
    histogram,
    interpolate_kick,
    longitudinal_evolve,
    next_power_of_two,
    quantum_excitation,
    rf_kick,
    synchrotron_radiation,
    wake_potential,
    z_to_ϕ,
    ϕ_to_z,
)
from ._lib import HaissinskiError  # noqa: F401
