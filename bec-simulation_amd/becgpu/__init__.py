"""becgpu: MI355X (gfx950) split-step engine for spinor BECs, a drop-in GPU path
for the hot loop of SpinorBEC.jl.  The compute runs in libbecgpu.so (hand-written
HIP kernels behind the C ABI in include/becgpu.h); this package is the host
driver that keeps the reference's public entry points."""
from .physics import (AdaptiveDtParams, AtomSpecies, Eu151, GridConfig, HarmonicTrap,  # noqa: F401
                      InteractionParams, LossParams, RamanCoupling, GravityPotential,
                      CompositePotential, TimeDependentZeeman, zeeman_at,
                      interaction_params_from_constraint, Na23, NoPotential, Rb87, SimParams,
                      SpinSystem,
                      ZeemanParams, add_noise, cell_volume, classify_phase, init_psi,
                      init_psi_thomas_fermi, make_grid, spin_matrices)
from .api import (SimulationResult, Workspace, energy_terms, fft_gradient,  # noqa: F401
                  fft_partial_derivative, find_ground_state, find_ground_state_multistart,
                  magnetization, make_workspace, run_simulation, run_simulation_adaptive,
                  run_simulation_yoshida, scan_continuation, split_step, strang_core,
                  total_energy, total_norm, yoshida_core)
