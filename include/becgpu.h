/* becgpu -- C ABI of the MI355X (gfx950) split-step engine for spinor BECs.
 *
 * Drop-in boundary for the hot path of SpinorBEC.jl (anko9801/BEC-simulation).
 * The reference is pure Julia with no FFI; each entry point below replaces the
 * Julia function cited next to it, and a host driver (the `becgpu` Python
 * package here, or a Julia `ccall` module) keeps the reference's public entry
 * points (make_workspace, split_step!, run_simulation!, find_ground_state,
 * find_ground_state_multistart, scan_continuation, run_simulation_yoshida!,
 * total_energy) on top of it.
 *
 * Conventions
 *  - psi is exchanged as interleaved complex double (re, im), in the reference's
 *    memory order: column-major (x, y, z, component) -- x fastest, component
 *    slowest; component c <-> m = F - c (types.jl:239-244).
 *  - A workspace holds B >= 1 independent instances of the same shape (grid,
 *    spin F, precision); each instance has its own physical parameters, trap
 *    and psi.  Operations act on the active instances (bec_set_active).
 *  - Every call returns 0 on success or a negative BEC_E* code; the message of
 *    the last failure on the calling thread is available from bec_last_error().
 *  - A workspace is not thread-safe; distinct workspaces may be driven from
 *    distinct host threads (each owns its HIP stream).
 */
#ifndef BECGPU_H
#define BECGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define BEC_OK 0
#define BEC_EINVAL (-1)   /* invalid argument / configuration (ArgumentError)    */
#define BEC_EDEVICE (-2)  /* HIP runtime error                                   */
#define BEC_ENOMEM (-3)   /* device allocation failed                            */
#define BEC_EUNSUP (-4)   /* configuration valid in the reference, not built here */

#define BEC_FP64 0
#define BEC_FP32 1 /* FP32 psi/FFT/pointwise; FP64 reductions and phase arguments */

/* Energy terms returned per instance by bec_energy (energy.jl:6-54). */
enum {
  BEC_E_KIN = 0, BEC_E_TRAP, BEC_E_ZEEMAN, BEC_E_C0, BEC_E_C1, BEC_E_DDI, BEC_E_LHY,
  BEC_E_TENSOR, BEC_E_NORM, BEC_E_MZ, BEC_E_RAMAN, BEC_E_TOTAL, BEC_E_COUNT
};

typedef struct bec_ws bec_ws;

/* Shape and fixed options of a workspace (Workspace, types.jl:387-407;
 * GridConfig types.jl:1-23; SimParams types.jl:217-237). */
typedef struct bec_config {
  int ndim;            /* 1, 2 or 3 */
  int n[3];            /* grid points per axis (x, y, z); unused axes 1.  Any positive
                          even side (types.jl:12): 2^k in [8, 512], 3*2^k in [24, 192],
                          5*2^k in [40, 160] and 7*2^k in [56, 224] have compiled radix
                          plans, every other side runs line passes of run-time length.
                          BEC_EUNSUP when one x row of all 2F+1 components (with the DDI
                          lines) or one y/z line does not fit a block's LDS, or an
                          instance has 2^31 complex values or more */
  double L[3];         /* box size per axis */
  int F;               /* spin (2F+1 components), 0 <= F <= 6 (SpinSystem, types.jl:36-46) */
  int n_instances;     /* batch size B */
  int precision;       /* BEC_FP64 or BEC_FP32 */
  int imaginary_time;  /* SimParams.imaginary_time */
  int normalize_every; /* SimParams.normalize_every (ITP renormalisation cadence) */
  int enable_ddi;      /* make_workspace(enable_ddi=...) */
  int ddi_secular;     /* make_ddi_params(secular=...) */
  int ddi_padding;     /* make_workspace(ddi_padding=...): 2x zero-padded DDI, sides <= 256 */
  int tensor;          /* general-F tensor step replaces spin mixing (TensorInteractionCache
                        * active, initialization.jl:164-185; any F >= 1, also a cache built
                        * directly, _make_tensor_cache_from_channels); couplings via
                        * bec_set_tensor */
  /* Cell-centred coordinates and fftfreq wavenumbers of each axis, concatenated
   * x then y then z (grid.jl:1-23). */
  const double* x_tab;
  const double* k_tab;
} bec_config;

/* Per-instance physics (InteractionParams, ZeemanParams, HarmonicTrap, dt). */
typedef struct bec_inst_params {
  double dt;
  double c0, c1, c_lhy; /* interactions.jl / types.jl:336-346 */
  double c2;            /* nematic coupling get_cn(ip, 2) = c_extra[1] (types.jl:157-162); set
                         * 0 when the tensor step carries the contact channels */
  double c_dd;          /* DDIParams.C_dd */
  double p, q;          /* ZeemanParams (zeeman.jl:15-19) */
  double omega[3];      /* HarmonicTrap omega per axis (potentials.jl:5-15) */
} bec_inst_params;

/* Workspace lifetime -- replaces make_workspace (initialization.jl:86-191). */
int bec_ws_create(const bec_config* cfg, bec_ws** out);
void bec_ws_destroy(bec_ws* ws);
int bec_set_params(bec_ws* ws, int inst, const bec_inst_params* p);
/* Explicit potential array (evaluate_potential, potentials.jl:1-32), n sites, or
 * NULL to return to the harmonic trap of the instance parameters. */
int bec_set_potential(bec_ws* ws, int inst, const double* V);
/* Channel couplings g_S for S = 0, 2, ..., 2F (n = F+1) of the general-F tensor step
 * (TensorInteractionCache.g_values, tensor_interaction.jl:106-119; the caller forms
 * g_S = _c0c1_to_gS + _c_extra_to_delta_gS and zeroes c0, c1 as make_workspace does). */
int bec_set_tensor(bec_ws* ws, int inst, const double* gS, int n);
/* Loss rates gamma_c + L3 per component (n = 2F+1) of apply_loss_step! (losses.jl:25-82;
 * the caller forms gamma_c = _dipolar_relaxation_rates(F, gamma_dr)), or rates = NULL for
 * no loss (LossParams absent).  Real-time split_step and leapfrog segments then apply it. */
int bec_set_loss(bec_ws* ws, int inst, const double* rates, int n);
/* RamanCoupling of one instance (types.jl:179-183): raman = {Omega_R, delta, k_eff[ndim]},
 * or NULL for none.  With it, every half potential step applies apply_raman_step!
 * (raman.jl:8-31) with dt_half/2 before and after the DDI step (split_step.jl:77-107), and
 * bec_energy reports _raman_energy (energy.jl:123-135) in BEC_E_RAMAN. */
int bec_set_raman(bec_ws* ws, int inst, const double* raman);
/* apply_loss_step! with step dt (losses.jl:25-49), as run_simulation_yoshida! applies it
 * after each accepted step (yoshida.jl:64-66). */
int bec_loss_step(bec_ws* ws, double dt);
/* psi upload/download (SimState.psi). */
int bec_set_psi(bec_ws* ws, int inst, const double* psi);
int bec_get_psi(bec_ws* ws, int inst, double* psi);
/* Restrict subsequent operations to the listed instances. */
int bec_set_active(bec_ws* ws, const int* insts, int n);

/* split_step! x n_steps (split_step.jl:10-35): V(dt/2) K(dt) V(dt/2), ITP
 * renormalisation every normalize_every steps.  t/step bookkeeping is host-side. */
int bec_split_step(bec_ws* ws, int n_steps);
/* The same n_steps as one captured HIP graph, replayed on later calls with the same step
 * count, active list and set of active operators (no counterpart in the reference: it
 * removes the per-kernel launch cost of small grids).  Parameters (dt, couplings, trap,
 * loss rates) may change between replays: the kernels read them from device memory.
 * Results are bitwise those of bec_split_step.  With profiling on it runs uncaptured. */
int bec_split_step_graph(bec_ws* ws, int n_steps);
/* One leapfrog segment of _run_simulation_leapfrog! (simulation.jl:49-94):
 * V(dt/2) [K(dt) V(dt)]^(n-1) K(dt) V(dt/2). */
int bec_leapfrog_segment(bec_ws* ws, int n_steps);
/* _half_potential_step! (split_step.jl:48-133) with explicit dt_half. */
int bec_half_potential(bec_ws* ws, double dt_half, int imaginary_time);
/* apply_kinetic_step_batched! with the phase for dt (propagators.jl:130-145). */
int bec_kinetic(bec_ws* ws, double dt, int imaginary_time);
/* _strang_core! / _yoshida_core! (split_step.jl:146-183), real time. */
int bec_strang_core(bec_ws* ws, double dt);
int bec_yoshida_core(bec_ws* ws, double dt);
/* The same cores started from psi in slot src_slot (1..3; 0: the cores above): slot 0
 * receives the result and src_slot keeps the starting state, as run_simulation_yoshida!
 * needs for psi_old (src/yoshida.jl:36-44) without a copy of psi: the first launch of the
 * core reads src_slot (with every instance active; otherwise the slot is copied first). */
int bec_strang_core_from(bec_ws* ws, double dt, int src_slot);
int bec_yoshida_core_from(bec_ws* ws, double dt, int src_slot);
/* _normalize_psi! (split_step.jl:185-195). */
int bec_normalize(bec_ws* ws);
/* _normalize_psi_constrained! (ground_state.jl:279-320) of every active instance, with
 * target <F_z> target_mz[inst] (array of B values). */
int bec_normalize_constrained(bec_ws* ws, const double* target_mz);

/* total_energy terms (energy.jl:6-121) + norm and Mz (observables.jl:1-33),
 * evaluated in FP64.  FP32-bulk imaginary-time workspaces (normalize_every > 0) report
 * every term except the norm for psi / sqrt(norm), the state that their FP64
 * renormalisation describes (its FP32 copy carries a rounding of the norm of up to 1e-7);
 * out[B][BEC_E_COUNT] for every instance (inactive rows untouched). */
int bec_energy(bec_ws* ws, double* out);
/* Copies between psi slots: slot 0 is psi, slots 1..3 are device-resident
 * copies (psi_prev / psi_old / psi_strang of the drivers). */
int bec_copy_slot(bec_ws* ws, int dst_slot, int src_slot);
/* Exchanges the contents of two slots for EVERY instance of the workspace (a pointer swap:
 * no data moves), where the drivers would copy psi out of a slot they overwrite next
 * (the psi_strang copy of run_simulation_yoshida!, yoshida.jl:36-44).  Callers with inactive
 * instances whose slots must stay in place use bec_copy_slot instead. */
int bec_swap_slots(bec_ws* ws, int slot_a, int slot_b);
/* mode 0: out[B][2] = {sum |a-b|^2, sum |b|^2}  (_wavefunction_l2_change, adaptive.jl:42-50)
 * mode 1: out[B][2] = {max |a-b|, max |a|}      (dpsi test, ground_state.jl:148-149) */
int bec_diff(bec_ws* ws, int slot_a, int slot_b, int mode, double* out);
/* _compute_spin_density! (observables.jl:61-88) of instance inst, in FP64:
 * out[3][N] = Fx, Fy, Fz per site in psi's spatial memory order, with
 * Fz = sum_m m |psi_m|^2 and Fx + i Fy = sum_m f+(m) conj(psi_{m+1}) psi_m. */
int bec_spin_density(bec_ws* ws, int inst, double* out);
/* One check interval of find_ground_state's loop (ground_state.jl:132-162) with a single
 * host synchronisation: n_steps ITP split steps, then for every active instance
 * out[B][BEC_CHUNK_COUNT] = {total energy, dpsi = max|psi - psi_prev| / max|psi| against
 * slot 1, norm, max |Phi| of the last DDI convolution (0 without DDI; the overflow guard's
 * input, ground_state.jl:4-29), status (0, or 1 when the energy or norm is not finite)},
 * and slot 1 <- psi for the next interval.  The caller seeds slot 1 with the starting psi
 * (bec_copy_slot(ws, 1, 0)) and keeps E_prev; inactive rows are untouched. */
enum { BEC_CHUNK_E = 0, BEC_CHUNK_DPSI, BEC_CHUNK_NORM, BEC_CHUNK_PHIMAX, BEC_CHUNK_STATUS,
       BEC_CHUNK_COUNT };
int bec_itp_chunk(bec_ws* ws, int n_steps, double* out);
/* ITP guard (ground_state.jl:4-29): out[B][2] = {sum |psi|^2 (NaN if any NaN),
 * max |Phi| of the last DDI convolution (0 without DDI)}. */
int bec_itp_guard(bec_ws* ws, double* out);

/* Per-launch HIP-event timing on the workspace stream, by kernel class:
 *  BEC_K_XPASS      x-axis pass (FFT along x + fused per-site operators)
 *  BEC_K_FWD        y/z forward FFT pass over psi
 *  BEC_K_KIN        last-axis pass: forward FFT, kinetic phase, inverse FFT
 *  BEC_K_INV        y/z inverse FFT pass over psi
 *  BEC_K_KE         last-axis forward FFT + kinetic-energy partial sums
 *  BEC_K_DDI_FFT    y forward/inverse FFT of the DDI spectra
 *  BEC_K_DDI_CONV   last-axis DDI pass: forward FFT, C_dd Q_ab(k) contraction, inverse FFT
 *  BEC_K_REDUCE     partial-sum reduction
 *  BEC_K_SITES      lane-per-site pass (13-component tensor step and spin rotations)
 *  BEC_K_ENERGY     a whole bec_energy evaluation (all of its launches)
 *  BEC_K_XSITE      fused x pass + lane-per-site pass (F = 2 and 6: x FFTs, DDI x lines and
 *                   the heavy per-site operators in one launch)
 * No counterpart in the reference (which uses TimerOutputs, SpinorBEC.jl:12). */
enum {
  BEC_K_XPASS = 0, BEC_K_FWD, BEC_K_KIN, BEC_K_INV, BEC_K_KE, BEC_K_DDI_FFT, BEC_K_DDI_CONV,
  BEC_K_REDUCE, BEC_K_SITES, BEC_K_ENERGY, BEC_K_XSITE, BEC_K_COUNT
};
int bec_profile_enable(bec_ws* ws, int on);
/* ms[BEC_K_COUNT] total milliseconds and count[BEC_K_COUNT] launches since enable;
 * resets the record. */
int bec_profile_read(bec_ws* ws, double* ms, long long* count);
/* Work executed by the lane-per-site pass since bec_profile_enable (or the last call):
 * out[0] tensor matvecs (Taylor terms of the tensor step), out[1] Euler spin rotations,
 * out[2] flops per tensor matvec (the channel-folded form on FP64 workspaces), out[3] flops
 * per rotation, out[4 + k] algorithmic HBM
 * bytes issued by kernel class k (BEC_K_*; 0 for BEC_K_REDUCE and BEC_K_ENERGY), so
 * out[4 + k] / time of class k is its achieved algorithmic bandwidth;
 * out[4 + BEC_K_COUNT] FFT flops of the fused x/site launches (5 N log2 N per complex
 * line of N points); out[5 + BEC_K_COUNT] the matvecs among out[0] that were a site's first
 * Taylor term in the symmetric form (v = psi: one product per pair), and
 * out[6 + BEC_K_COUNT] the flops of such a matvec, so the tensor flops are
 * out[0] out[2] - out[5 + K] (out[2] - out[6 + K]).  out holds 7 + BEC_K_COUNT doubles.
 * Resets the counters.  No counterpart in the reference. */
int bec_profile_work(bec_ws* ws, double* out);

/* Spectral derivatives of a real field f[N] (memory order as psi without the component
 * axis) on the grid that cfg describes (ndim, n, L, x_tab, k_tab; every other field is
 * ignored), in FP64: out = Re IFFT(i k_dim FFT(f)).
 * bec_fft_derivative replaces _fft_partial_derivative (src/fft_utils.jl:4-19); dim is
 * 0-based (the reference's dim - 1).  bec_fft_gradient replaces _fft_gradient
 * (src/fft_utils.jl:25-44) and writes out[ndim][N]. */
int bec_fft_derivative(const bec_config* cfg, const double* f, int dim, double* out);
int bec_fft_gradient(const bec_config* cfg, const double* f, double* out);
/* bec_fft_gradient without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, result and errors as
 * bec_fft_gradient, computed by the same kernels; every device buffer it needs lives in one
 * fixed array per GPU (about 21 MB, held until the process ends) that keeps the last call's
 * contents.  Grids above 262144 sites (64^3) are rejected with BEC_EINVAL.  Calls on one GPU
 * share that array: this library serialises them with a process-wide lock, and work that a
 * caller issues on the array by other means would race with them. */
int bec_fft_gradient_noalloc(const bec_config* cfg, const double* f, double* out);

int bec_synchronize(bec_ws* ws);
/* Kernel schedule switch (no counterpart in the reference): on (the default, or the
 * environment variable BEC_XSITE=0 for off), spins F = 2 and 6 run every half potential
 * step with its neighbouring x FFTs and DDI x lines in the fused x/site kernel; off, in
 * separate x passes and lane-per-site passes.  Both give the same psi.  Returns the
 * setting in force (the fused kernel needs 2-D/3-D grids whose sides are all powers of two,
 * with one x-pass tile of 16..256 sites; other shapes always run separately). */
int bec_set_xsite(bec_ws* ws, int on);
/* Tile autotuner (no counterpart in the reference, whose FFTW plans are tuned by
 * FFTW_MEASURE, src/grid.jl:37-53): times the x-pass tile candidates on this workspace's
 * data and keeps the fastest for the shape (precision, n_x, D, grid rows) for the rest of
 * the process; later workspaces of the shape reuse the decision without timing.  psi is
 * not modified.  rows_out (optional): the x rows per block now in use. */
int bec_autotune(bec_ws* ws, int* rows_out);
/* Non-blocking completion test of the work queued on the workspace's stream: 0 when all
 * of it has finished, 1 while some is still running, a negative code on a device error.
 * No counterpart in the reference (its calls are synchronous); a request loop polls it
 * to serve each workspace as soon as its own work ends. */
int bec_query(bec_ws* ws);
/* Transform a line of n points runs through (no counterpart in the reference, whose FFTW
 * plans any length, src/grid.jl:37-53): 0 a built radix plan (2^k 8..512, 3 2^k 24..192,
 * 5 2^k 40..160, 7 2^k 56..224), 1 runtime mixed-radix passes, 2 Bluestein's algorithm
 * (lengths with a large prime factor, from about 214 points), BEC_EINVAL for n < 2. */
int bec_line_plan(int n);
const char* bec_last_error(void);
/* Bytes of device memory held by a workspace of this configuration (admission). */
long long bec_workspace_bytes(const bec_config* cfg);

/* Grouped split steps.  No counterpart in the reference, whose batched drivers
 * (find_ground_state_multistart, scan_continuation: ground_state.jl:239-271, 328-378) loop
 * over solves one at a time.  A group holds workspaces of different grid shapes with the
 * same precision, spin, dimension and time kind (n >= 1, each listed once; the group does
 * not own them and must be destroyed before them).  bec_group_split_step advances every
 * member's active instances by n_steps split steps (bec_split_step) with each position of
 * their common launch sequence issued as ONE kernel launch over all members (the x pass:
 * one launch per four members, whose descriptors travel as kernel arguments; use_graph:
 * the whole sequence replayed as one captured HIP graph, re-captured when a member's
 * active set, tiles or operator set change).  Results are bitwise those of bec_split_step
 * on each member.  Members whose shapes or passes have no grouped kernel (x sides other
 * than 32, 64, 128, 256; more than 5 components; dipolar, tensor or lane-per-site passes;
 * unequal x-pass block sizes) are advanced one workspace at a time instead, with the same
 * results.  Ordering: the group's launches follow the work queued on every member's
 * stream, and later work on a member's stream follows them. */
typedef struct bec_group bec_group;
int bec_group_create(bec_ws* const* ws, int n, bec_group** out);
int bec_group_split_step(bec_group* g, int n_steps, int use_graph);
/* 1 when the last bec_group_split_step issued grouped launches, 0 when it advanced the
 * members one at a time (or had nothing to do). */
int bec_group_grouped(const bec_group* g);
void bec_group_destroy(bec_group* g);

#ifdef __cplusplus
}
#endif
#endif /* BECGPU_H */
