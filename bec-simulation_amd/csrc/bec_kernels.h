// Host-visible launch interface of the split-step kernels (implemented in
// bec_kernels.hip).  Plain pointers and sizes; T is double (FP64 path) or float
// (FP32-bulk path, FP64 reductions and phase arguments).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bec_types.h"

namespace bec {

// Geometry shared by all instances of a workspace (instances differ only in
// their InstParams, trap tables and psi).
struct Geom {
  int ndim;
  int n[3];            // valid sizes, 1 for absent axes
  int D;               // 2F+1
  int nrows;           // n[1]*n[2]
  long long nsites;    // n[0]*n[1]*n[2]
  long long inst_stride;  // D*nsites (complex elements)
  double inv_ntot;     // 1/nsites
  double dV;
  int nmax;            // max(n)
  // DDI spectral grid: sn = 2n per axis with the zero-padded convolution (ddi_padded.jl),
  // else sn = n.  Spin densities fill the first n entries of each axis, Phi is read back
  // from the same entries.
  int sn[3];
  int smax;            // max(sn)
  int spad;            // 1: zero-padded DDI convolution
  int nxh, nxh_pitch;  // rFFT half-row length sn[0]/2+1 and row pitch of the DDI spectra
  long long fhat_field_stride;  // nxh_pitch*sn[1]*sn[2]
  long long fhat_inst_stride;   // 3*fhat_field_stride
  int ddi_secular;
  int xrows;           // x-pass rows per block chosen by the tile autotuner (0: size heuristic)
};

constexpr int kWorkSlots = 256;
constexpr int kWorkStride = 8;

// Line lengths the FFTs are built for: 2^k, 3 * 2^k, 5 * 2^k and 7 * 2^k from 8 to 512
// (every one a multiple of 8).  Twiddle table i belongs to kFftSizes[i].
constexpr int kFftSizes[] = {8,  16,  24,  32,  40,  48,  56,  64,  80,  96,
                             112, 128, 160, 192, 224, 256, 320, 384, 512};
constexpr int kNumFftSizes = 19;
constexpr int fft_size_id(int n) {
  for (int i = 0; i < kNumFftSizes; ++i)
    if (kFftSizes[i] == n) return i;
  return -1;
}
// Values per thread of a line FFT of a built length (the radix plan's E): 12 for 3 * 2^k,
// 20 for 5 * 2^k, 28 for 7 * 2^k, else 8.
constexpr int line_e(int n) { return n % 3 == 0 ? 12 : n % 5 == 0 ? 20 : n % 7 == 0 ? 28 : 8; }

// Runtime-length twiddle tables (bec_generic.hip): one per distinct side and zero-padded
// spectral length of a grid.
constexpr int kGenTabs = 6;

struct DevTables {
  const double* ktab;     // [3][nmax] fftfreq k values per axis
  const double* xtab;     // [3][nmax] cell-centred coordinates per axis
  const double* sktab;    // [3][smax] fftfreq k values of the DDI spectral grid
  const double* axtab;    // [B][3][nmax] omega_d^2 x_i^2 (trap), per instance
  const double* Varr;     // [B][nsites] explicit potential or null
  const InstParams* params;  // [B]
  const int* active;      // active instance list
  const void* tw[kNumFftSizes];  // per line length: exp(-2 pi i m/N), FP64 for every T
  const void* gtw[kGenTabs];     // runtime-length tables exp(-2 pi i m/N), m < gtw_n[i]
  int gtw_n[kGenTabs];           // their lengths (0: unused)
  // Bluestein tables of the same lengths (gbl_m[i] = 0: none): chirp c[n] = exp(i pi n^2/N)
  // (n < N), the transform of the wrapped chirp over M points divided by M, and
  // exp(-2 pi i m/M) (m < M), M = gbl_m[i] a power of two >= 2N - 1
  const void* gbl[kGenTabs];
  int gbl_m[kGenTabs];
  const double* red;      // [B][RS_COUNT] reduced scalars (norm for OP_SCALE)
  double* partials;       // [RS_COUNT][B][max_blocks]
  // profiling only (else null): kWorkSlots counter pairs, one per 64-byte line, so that the
  // blocks of a launch spread their atomics over many addresses; slot s holds
  // [s * kWorkStride] tensor matvecs and [s * kWorkStride + 1] spin rotations
  unsigned long long* work;
  // [B][tens::kKfoldMax] channel-folded tensor kernels (FP64 tensor step), or null
  const double* kfold;
  int max_blocks;
  int maxB;
};

enum StridedMode : int {
  SM_FWD = 0,
  SM_INV = 1,
  SM_KIN = 2,   // forward, kinetic phase/N, inverse
  SM_DDI = 3,   // 3 fields: forward, C_dd Q(k)/N contraction, inverse
  SM_KE = 4,    // forward then sum k^2 |v|^2 (no store)
  SM_DERIV = 5, // forward, i k_c / N on component c (c < ndim), inverse (FP64 only)
};

// x-axis pass: tiles of whole x rows, all components; executes `prog`.
template <typename T>
hipError_t launch_xpass(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                        const void* src, void* dst, void* fhat, int n_active, void* phi = nullptr);

// Fused x pass + lane-per-site pass (bec_xsite.hip): stages XS_INV_FFT, XS_LOAD_PHI,
// XS_OPS_A (any operator), XS_WRITE_FHAT, XS_NORM, XS_FWD_FFT from src into psi (the same
// buffer, or another slot of the same layout), with the x-pass tiling (xpass_grid blocks,
// so its partials line up with the x pass's).
template <typename T>
hipError_t launch_xsite(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                        const void* src, void* psi, void* fhat, int n_active);
// Whether the fused kernel covers this geometry (F = 2 or 6, 2-D/3-D, one site per thread).
bool xsite_ok(const Geom& g);
// Lane-per-site pass: one thread per grid site applies `prog` to its spinor in VGPRs
// (phi: Phi(x) fields [inst][3][nsites] written by an XS_STORE_PHI x pass, or null).
template <typename T>
hipError_t launch_sites(hipStream_t st, const Geom& g, const DevTables& tb, const SProgram& prog,
                        void* psi, const void* phi, int n_active);

// Pointwise energy partials of psi (phi: Phi(x) of an XS_STORE_PHI pass, or null).
template <typename T>
hipError_t launch_energy(hipStream_t st, const Geom& g, const DevTables& tb, const void* psi,
                         const void* phi, int with_tensor, int n_active, int* nblocks_out);

// y or z axis pass (axis = 1 or 2) over psi (fields = D) or over the DDI
// spectra (fields = 3, DDI mode).  SM_KIN uses dt = (dt_override > 0 ? dt_override :
// instance dt) * kin_frac.  nblocks_out (optional): partials per instance (SM_KE).
template <typename T>
hipError_t launch_strided(hipStream_t st, const Geom& g, const DevTables& tb, int axis, int mode,
                          void* data, int n_active, double kin_frac, double dt_override, int itp,
                          int ddi_data, int* nblocks_out);

// 1-D grids: Q_ab product of the DDI half spectra (the step the y/z spectra passes carry
// on 2-D and 3-D grids).
template <typename T>
hipError_t launch_ddi_line(hipStream_t st, const Geom& g, const DevTables& tb, void* fhat,
                           int n_active);

// Tile shapes (shared by the launchers and the workspace sizing).
constexpr int strided_tb_c(int N, int fields, int csize) {
  return (fields * N * 16 * csize <= 65536) ? 16 : (fields * N * 8 * csize <= 65536) ? 8 : 4;
}
int strided_tb(int N, int fields, int csize);
int xpass_rows(int D, int NX, int nrows);
// rows per x-pass block for this geometry (the tuned value when set)
int xpass_rows_g(const Geom& g);
// whether `rows` rows per block is a valid x-pass tile for this geometry
bool xpass_rows_ok(const Geom& g, int rows);
// Blocks per instance of an x-pass launch (= number of per-block partial sums it writes).
int xpass_grid(const Geom& g);

// Line passes of runtime length (bec_generic.hip) for grid sides without a built FFT:
// the x pass (same stages as launch_xpass) and the y/z passes (same modes as
// launch_strided, over `other_lines` lines of the other transverse axis).  launch_xpass
// and launch_strided hand over to them by themselves.
bool xpass_generic(const Geom& g);     // the x pass of this geometry runs the runtime kernel
int xgen_rows(const Geom& g);          // its rows per block
int sgen_tb(int N, int fields, int csize);  // y/z tile columns of a runtime length (0: too long)
// Bluestein length M of a runtime line length N (0: N keeps its mixed-radix Stockham passes)
int gen_blue_m(int N);
// every pass of this geometry fits one block's LDS (csz: bytes per complex element)
bool generic_fits(const Geom& g, int csz, int ddi);
template <typename T>
hipError_t launch_xpass_gen(hipStream_t st, const Geom& g, const DevTables& tb,
                            const XProgram& prog, const void* src, void* dst, void* fhat,
                            int n_active, void* phi);
template <typename T>
hipError_t launch_strided_gen(hipStream_t st, const Geom& g, const DevTables& tb, int axis,
                              int mode, void* data, int n_active, double kin_frac,
                              double dt_override, int itp, int ddi_data, int other_lines);

// Grouped launches (bec_group.hip): one launch runs a pass of the same schedule position
// for several workspaces (different grid shapes), each block picking its workspace from a
// device array of descriptors by its linear index (block0 = first block of the descriptor).
struct XGroupDesc {
  Geom g;
  DevTables tb;
  XProgram prog;
  const void* src;
  void* dst;
  void* fhat;
  void* phi;
  int rows;    // x rows per tile
  int gx;      // tiles per instance (blocks = gx * active instances)
  int block0;
  int pad_;
};
struct SGroupDesc {
  Geom g;
  DevTables tb;
  void* data;
  double kin_frac, dt_override;
  int axis, itp;
  int gx, gy;  // x chunks * lines of the other axis, component groups (blocks = gx*gy*active)
  int block0;
  int pad_;
};
struct RGroupDesc {
  DevTables tb;
  double* red;
  int nblocks;
  uint32_t slot_mask, max_mask;
  int gx;      // slots reduced (blocks = gx * active instances)
  int block0;
  int pad_;
};
struct CGroupDesc {  // ITP renormalisation psi *= 1/sqrt(norm dV)
  Geom g;
  DevTables tb;
  void* psi;
  int gx;
  int block0;
};
// Shapes the grouped kernels are built for: x sides 32, 64, 128, 256 with 1, 3 or 5
// components (x pass); lines of 32 ... 256 points (y/z pass, modes SM_FWD, SM_INV, SM_KIN).
bool xgroup_ok(const Geom& g);
bool sgroup_ok(const Geom& g, int axis, int mode, int ddi_data);
constexpr int kSGroupThreads = 256;  // y/z tiles of 2048 elements: 2048 / N columns
// x pass: descs is a HOST array of nd <= kXGroupMax descriptors (passed as the kernel
// argument), block0 counted from the first; the other grouped launches read theirs from
// device memory.
constexpr int kXGroupMax = 4;
template <typename T>
hipError_t launch_xpass_group(hipStream_t st, int D, const XGroupDesc* descs, int nd, int blocks,
                              int threads, size_t lds);
template <typename T>
hipError_t launch_strided_group(hipStream_t st, int mode, const SGroupDesc* descs, int nd,
                                int blocks);
hipError_t launch_reduce_group(hipStream_t st, const RGroupDesc* descs, int nd, int blocks);
template <typename T>
hipError_t launch_scale_group(hipStream_t st, const CGroupDesc* descs, int nd, int blocks);
// LDS bytes of an x-pass launch of this geometry and program (as launch_xpass sizes it)
size_t xpass_lds_g(const Geom& g, const XProgram& prog, size_t csz);
// threads of an x-pass block of this geometry
int xpass_threads_g(const Geom& g);

// Spin density (Fx, Fy, Fz) of one instance's psi into out[3][nsites], FP64 (bec_aux.hip).
template <typename T>
hipError_t launch_spin_density(hipStream_t st, const void* psi, double* out, long long nsites,
                               int D);

// dst[i] = (double)src[i], i < n (FP32 psi widened for FP64 energy evaluation)
hipError_t launch_widen(hipStream_t st, const void* src, void* dst, long long n);

// Sum (or max) the per-block partials of `slots` into red[inst][slot].
hipError_t launch_reduce(hipStream_t st, const DevTables& tb, double* red, int n_active,
                         int nblocks, uint32_t slot_mask, uint32_t max_mask);

// psi *= 1/sqrt(red[inst][RS_NORM] * dV)
template <typename T>
hipError_t launch_scale_norm(hipStream_t st, const Geom& g, const DevTables& tb, void* psi,
                             int n_active);

// Magnetisation-constrained normalisation of every active instance (target[inst] = Mz):
// per-component norms into part[B][D][nblk], factors into fac[B][kMaxD], then the scaling.
template <typename T>
hipError_t launch_constrained_norm(hipStream_t st, const Geom& g, const DevTables& tb, void* psi,
                                   int n_active, double* part, int nblk, const double* target,
                                   double* fac);

// dst <- src for every active instance (bytes_per_inst a multiple of 16)
hipError_t launch_copy_slot(hipStream_t st, const DevTables& tb, const void* src, void* dst,
                            long long bytes_per_inst, int n_active);

// partials: RS_NORM <- sum|a-b|^2 (or max|a-b| when want_max), RS_KIN <- sum|b|^2 (or max|a|)
template <typename T>
hipError_t launch_diff(hipStream_t st, const Geom& g, const DevTables& tb, const void* a,
                       const void* b, int n_active, int want_max, int nblocks);

}  // namespace bec
