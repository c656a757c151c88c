// Types shared by the HIP kernels and the host-side workspace code.
#pragma once
#include <stdint.h>

namespace bec {

constexpr int kMaxD = 13;   // 2F+1 for F <= 6
constexpr int kMaxOps = 12;

// Per-instance physical parameters (one per independent solve in a batch).
// Device-resident; every block reads the entry of the instance it works on.
struct InstParams {
  double dt;         // time step of this instance
  double c0, c1, c_lhy;
  double cdd;        // dipolar strength C_dd
  double zee[kMaxD]; // Zeeman energies E_c = -p m + q m^2, m = F - c
  double zee_shift;  // min_c E_c (subtracted in imaginary time)
  double gS[7];      // channel couplings g_S, S = 0, 2, ..., 2F (general-F tensor step)
  int active;
  int pad_;
};

enum OpKind : int {
  OP_DIAG = 1,      // _diagonal_step_svec! (dt*frac)
  OP_SPINMIX = 2,   // apply_spin_mixing_step! (dt*frac)
  OP_DDI_ROT = 3,   // Euler rotation by the DDI potential (dt*frac)
  OP_SCALE = 4,     // multiply by 1/sqrt(norm2) of the instance (ITP renormalisation)
  OP_TENSOR = 5,    // apply_tensor_interaction_step! (dt*frac)
};

struct Op {
  int kind;
  int pad_;
  double frac;
};

// Program executed by the x-axis pass kernel on a tile of whole x-rows holding
// all 2F+1 components (so every per-site operation sees the full spinor).
enum XStage : int {
  XS_INV_FFT = 1 << 0,     // inverse x FFT on load (end of a kinetic step)
  XS_LOAD_PHI = 1 << 1,    // read DDI Phi_hat half rows, C2R along x
  XS_OPS_A = 1 << 2,       // per-site op list A
  XS_KIN_1D = 1 << 3,      // 1-D grid: forward FFT, kinetic phase, inverse FFT in LDS
  XS_OPS_B = 1 << 4,       // per-site op list B (after the 1-D kinetic step)
  XS_WRITE_FHAT = 1 << 5,  // spin density -> R2C along x -> F_hat half rows
  XS_NORM = 1 << 6,        // per-block partial of sum |psi|^2 (after ops)
  XS_FWD_FFT = 1 << 7,     // forward x FFT before the store
  XS_NO_STORE = 1 << 8,    // do not write psi back
  XS_KE_1D = 1 << 10,      // 1-D kinetic energy partial sum k^2 |psi_hat|^2 after forward FFT
  XS_STORE_PHI = 1 << 12,  // after XS_LOAD_PHI: write Phi(x) (3 real fields) for the site kernel
  XS_NO_LOAD = 1 << 13,    // psi is not read (with XS_NO_STORE: a pass over the DDI spectra only)
};

struct XProgram {
  int stages;
  int nA, nB;
  int itp;            // imaginary time
  double dt_override; // > 0: use instead of the per-instance dt
  double kin_frac;    // 1-D kinetic step dt fraction
  Op A[kMaxOps];
  Op B[kMaxOps];
};

// Program of the lane-per-site kernel (heavy per-site operators of large spins).
struct SProgram {
  int n;
  int itp;
  double dt_override;
  Op ops[kMaxOps];
};

// Reduction slots written per block into the partials buffer.
enum RedSlot : int {
  RS_NORM = 0,
  RS_KIN = 1,
  RS_TRAP = 2,
  RS_ZEE = 3,
  RS_C0 = 4,
  RS_C1 = 5,
  RS_LHY = 6,
  RS_DDI = 7,
  RS_MZ = 8,
  RS_PHIMAX = 9,
  RS_TENSOR = 10,
  RS_COUNT = 11,
};

}  // namespace bec

// Enumeration of the Clebsch-Gordan table used by the tensor step: for every
// M = -2F..2F, every unordered pair m1 <= m2 with m1 + m2 = M, every even
// S >= |M| up to 2F: <F m1; F m2 | S M>.  Shared by host (table build) and
// device (compile-time offsets).
namespace tens {
__host__ __device__ constexpr int m1lo(int F, int M) { return M - F > -F ? M - F : -F; }
__host__ __device__ constexpr int m1hi(int M) { return M >= 0 ? M / 2 : -((-M + 1) / 2); }
__host__ __device__ constexpr int npairs(int F, int M) { return m1hi(M) - m1lo(F, M) + 1; }
__host__ __device__ constexpr int s0(int M) { return ((M < 0 ? -M : M) + 1) / 2 * 2; }
__host__ __device__ constexpr int ns(int F, int M) { return (2 * F - s0(M)) / 2 + 1; }
__host__ __device__ constexpr int base(int F, int M) {
  int b = 0;
  for (int q = -2 * F; q < M; ++q) b += npairs(F, q) * ns(F, q);
  return b;
}
__host__ __device__ constexpr int size(int F) { return base(F, 2 * F + 1); }
__host__ __device__ constexpr int off(int F, int M, int m1, int S) {
  return base(F, M) + (m1 - m1lo(F, M)) * ns(F, M) + (S - s0(M)) / 2;
}
}  // namespace tens
