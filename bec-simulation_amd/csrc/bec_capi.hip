// C ABI of the split-step engine: workspace ownership, parameter upload and the
// schedule compiler that turns the reference's operator sequence
// (split_step.jl / simulation.jl / yoshida.jl) into fused HBM passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/becgpu.h"
#include "bec_aux.h"
#include "bec_kernels.h"

using namespace bec;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return fail(BEC_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));    \
  } while (0)

// One launch of a split-step schedule, as recorded for a grouped launch (bec_group_*).
enum RecKind { R_X, R_STRIDED, R_REDUCE, R_SCALE, R_OTHER };
struct Rec {
  int kind;
  XProgram xp;          // R_X
  const void* src;
  void* dst;
  int axis, mode, itp, ddi;  // R_STRIDED
  void* data;
  double kin_frac, dt_ovr;
  int nblocks;          // R_REDUCE
  uint32_t slots, maxmask;
};

struct bec_ws {
  // set while a grouped launch records this workspace's schedule: the launch helpers append
  // to it instead of launching
  std::vector<Rec>* rec = nullptr;
  Geom g{};
  int F = 0, D = 0, B = 0, prec = 0, itp = 0, normalize_every = 1, ddi = 0, tensor = 0;
  size_t csz = 16;
  std::vector<InstParams> hp;
  std::vector<bec_inst_params> user;
  std::vector<int> active;
  std::vector<double> xtab, ktab;  // host copies, concatenated per axis
  std::vector<int> has_varr;
  void* slot[4] = {nullptr, nullptr, nullptr, nullptr};
  // slot the next launch of a core reads psi from (bec_strang_core_from and
  // bec_yoshida_core_from; -1: slot 0 in place)
  int src_once = -1;
  void* scratch = nullptr;
  void* fhat = nullptr;
  void* phi = nullptr;  // Phi(x) fields for the lane-per-site DDI rotation (F=6)
  // FP64 evaluation copies for FP32-bulk workspaces (energies are always FP64)
  void* psi64 = nullptr;
  void* scratch64 = nullptr;
  void* fhat64 = nullptr;
  void* phi64 = nullptr;
  double* d_ktab = nullptr;
  double* d_xtab = nullptr;   // [3][nmax] cell-centred coordinates (Raman phase k.r)
  double* d_sktab = nullptr;  // k of the zero-padded DDI grid (null: the grid's own k)
  double* d_axtab = nullptr;
  double* d_varr = nullptr;
  double* d_red = nullptr;
  double* d_partials = nullptr;
  InstParams* d_params = nullptr;
  int* d_active = nullptr;
  unsigned long long* d_work = nullptr;  // profiling counters (tensor matvecs, spin rotations)
  double* d_sdens = nullptr;  // bec_spin_density output staging, [3][nsites] (first use)
  double* d_kfold = nullptr;  // [B][tens::kKfoldMax] channel-folded tensor kernels (tensor ws)
  std::vector<double> h_kfold;
  double* d_cpart = nullptr;   // constrained normalisation: per-component partials, targets,
  double* d_ctarget = nullptr; // factors (allocated on first use)
  double* d_cfac = nullptr;
  void* d_tw[kNumFftSizes] = {};
  void* d_gtw[kGenTabs] = {};  // runtime-length twiddle tables (sides without a built FFT)
  int gtw_n[kGenTabs] = {};
  void* d_gbl[kGenTabs] = {};  // Bluestein tables of the same lengths (or null)
  int gbl_m[kGenTabs] = {};
  double* h_red = nullptr;   // pinned
  double* h_red2 = nullptr;  // pinned, second staging buffer of bec_itp_chunk (first use)
  int max_blocks = 0;
  int n_xblocks = 0;
  bool params_dirty = true;
  hipStream_t stream = nullptr;
  DevTables tb{};
  // per-launch HIP-event profiling (bec_profile_*)
  bool prof_on = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> prof_cls;
  double prof_bytes[BEC_K_COUNT] = {};  // algorithmic HBM bytes issued per kernel class
  double prof_fft_flops = 0.0;          // FFT flops issued by fused x/site launches
  // heavy per-site operators of F = 2 and 6 run in the fused x/site kernel (bec_set_xsite)
  bool xsite = true;
  // captured split-step runs (bec_split_step_graph), keyed by step count, active list and
  // the schedule flags; dropped whenever a captured kernel argument could change
  struct Graph {
    int n;
    std::vector<int> active;
    int flags;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
};

static void drop_graphs(bec_ws* ws) {
  for (auto& g : ws->graphs) (void)hipGraphExecDestroy(g.exec);
  ws->graphs.clear();
}

// Kernel classes reported by bec_profile_read (order = BEC_K_* in becgpu.h).
// `bytes`: algorithmic HBM bytes of the launch (each array the pass must touch read
// and/or written once), accumulated per class for bec_profile_work.
static int prof_begin(bec_ws* ws, int cls, double bytes = 0.0) {
  if (!ws->prof_on) return -1;
  ws->prof_bytes[cls] += bytes;
  if (ws->ev_used + 2 > ws->ev_pool.size()) {
    for (int i = 0; i < 256; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      ws->ev_pool.push_back(e);
    }
  }
  const int idx = (int)ws->ev_used;
  ws->ev_used += 2;
  ws->prof_cls.push_back(cls);
  (void)hipEventRecord(ws->ev_pool[idx], ws->stream);
  return idx;
}
static void prof_end(bec_ws* ws, int idx) {
  if (idx >= 0) (void)hipEventRecord(ws->ev_pool[idx + 1], ws->stream);
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static constexpr size_t kWorkBytes = sizeof(unsigned long long) * kWorkSlots * kWorkStride;

// Grid sides with built x-line FFTs: 2^k (8..512), 3 * 2^k (24..192), 5 * 2^k (40..160),
// 7 * 2^k (56..224).  (Zero-padded DDI lines are twice the side and must be built too.)
static bool side_supported(int n) {
  if (n == 320 || n == 384) return false;
  return fft_size_id(n) >= 0;
}

static void refresh_tables(bec_ws* ws) {
  drop_graphs(ws);  // captured launches hold the previous tables by value
  DevTables& t = ws->tb;
  t.ktab = ws->d_ktab;
  t.xtab = ws->d_xtab;
  t.sktab = ws->d_sktab ? ws->d_sktab : ws->d_ktab;
  t.axtab = ws->d_axtab;
  t.Varr = ws->d_varr && std::any_of(ws->has_varr.begin(), ws->has_varr.end(), [](int v) { return v; })
               ? ws->d_varr
               : nullptr;
  t.params = ws->d_params;
  t.active = ws->d_active;
  for (int i = 0; i < kNumFftSizes; ++i) t.tw[i] = ws->d_tw[i];
  for (int i = 0; i < kGenTabs; ++i) {
    t.gtw[i] = ws->d_gtw[i];
    t.gtw_n[i] = ws->gtw_n[i];
    t.gbl[i] = ws->d_gbl[i];
    t.gbl_m[i] = ws->gbl_m[i];
  }
  t.red = ws->d_red;
  t.partials = ws->d_partials;
  t.max_blocks = ws->max_blocks;
  t.maxB = ws->B;
  t.work = ws->prof_on ? ws->d_work : nullptr;
  t.kfold = ws->d_kfold;
}

static int sync_params(bec_ws* ws) {
  if (!ws->params_dirty) return BEC_OK;
  HIPCHK(hipMemcpyAsync(ws->d_params, ws->hp.data(), sizeof(InstParams) * ws->B,
                        hipMemcpyHostToDevice, ws->stream));
  if (ws->d_kfold)
    HIPCHK(hipMemcpyAsync(ws->d_kfold, ws->h_kfold.data(), ws->h_kfold.size() * 8,
                          hipMemcpyHostToDevice, ws->stream));
  HIPCHK(hipStreamSynchronize(ws->stream));
  ws->params_dirty = false;
  return BEC_OK;
}

static int n_active(const bec_ws* ws) { return (int)ws->active.size(); }

static double psi_bytes(const bec_ws* ws) {
  return (double)ws->D * (double)ws->g.nsites * (double)ws->csz * n_active(ws);
}
static double half_field_bytes(const bec_ws* ws) {  // 3 rFFT half-grid fields
  return 3.0 * ws->g.nxh * (double)(ws->g.nsites / ws->g.n[0]) * (double)ws->csz * n_active(ws);
}
// DDI spectra pass along `axis`: lines of the other transverse axis it visits, entries of
// each line it reads and writes (zero padding: the first n entries only on one side)
static double strided_ddi_bytes(const bec_ws* ws, int axis, int mode) {
  const Geom& g = ws->g;
  const int N = g.sn[axis];
  const int n_in = (g.spad && mode != SM_INV) ? g.n[axis] : N;
  const int n_out = (g.spad && mode != SM_FWD) ? g.n[axis] : N;
  const int n_other = axis == 1 ? g.n[2] : g.sn[1];
  return 3.0 * g.nxh * (double)n_other * (n_in + n_out) * (double)ws->csz * n_active(ws);
}
static double phi_bytes(const bec_ws* ws) {  // 3 real fields
  return 3.0 * (double)ws->g.nsites * (ws->csz / 2) * n_active(ws);
}

#define CHK(expr)                 \
  do {                            \
    int _r = (expr);              \
    if (_r != BEC_OK) return _r;  \
  } while (0)

// A core that starts from another slot (src_once): x passes and fused passes that load
// and store psi read that slot directly and write slot 0; before any other first launch
// the slot is copied into slot 0.
static int flush_src(bec_ws* ws) {
  if (ws->src_once < 0) return BEC_OK;
  const int s = ws->src_once;
  ws->src_once = -1;
  const size_t b = ws->csz * (size_t)ws->g.inst_stride;
  if (b % 16 == 0) {
    HIPCHK(launch_copy_slot(ws->stream, ws->tb, ws->slot[s], ws->slot[0], (long long)b,
                            n_active(ws)));
    return BEC_OK;
  }
  for (int inst : ws->active)
    HIPCHK(hipMemcpyAsync((char*)ws->slot[0] + b * inst, (char*)ws->slot[s] + b * inst, b,
                          hipMemcpyDeviceToDevice, ws->stream));
  return BEC_OK;
}

static int do_x(bec_ws* ws, const XProgram& p, const void* src, void* dst) {
  if (ws->src_once >= 0 && !ws->rec) {
    if (src == ws->slot[0] && dst == ws->slot[0] && !(p.stages & (XS_NO_LOAD | XS_NO_STORE))) {
      src = ws->slot[ws->src_once];
      ws->src_once = -1;
    } else {
      CHK(flush_src(ws));
    }
  }
  if (ws->rec) {
    Rec r{};
    r.kind = R_X;
    r.xp = p;
    r.src = src;
    r.dst = dst;
    ws->rec->push_back(r);
    return BEC_OK;
  }
  double b = 0.0;
  if (ws->prof_on) {
    const double pb = psi_bytes(ws);
    if (!(p.stages & XS_NO_LOAD)) b += pb;
    if (!(p.stages & XS_NO_STORE)) b += pb;
    if (p.stages & (XS_WRITE_FHAT)) b += half_field_bytes(ws);
    if (p.stages & (XS_LOAD_PHI)) b += half_field_bytes(ws);
    if (p.stages & (XS_STORE_PHI)) b += phi_bytes(ws);
  }
  const int pe = prof_begin(ws, 0, b);
  hipError_t e = ws->prec == BEC_FP32
                     ? launch_xpass<float>(ws->stream, ws->g, ws->tb, p, src, dst, ws->fhat,
                                           n_active(ws), ws->phi)
                     : launch_xpass<double>(ws->stream, ws->g, ws->tb, p, src, dst, ws->fhat,
                                            n_active(ws), ws->phi);
  prof_end(ws, pe);
  HIPCHK(e);
  return BEC_OK;
}
static int do_sites(bec_ws* ws, const SProgram& p) {
  if (ws->rec) {
    ws->rec->push_back(Rec{R_OTHER});
    return BEC_OK;
  }
  CHK(flush_src(ws));
  double b = 0.0;
  if (ws->prof_on) {
    b = 2.0 * psi_bytes(ws);
    for (int i = 0; i < p.n; ++i)
      if (p.ops[i].kind == OP_DDI_ROT) {
        b += phi_bytes(ws);
        break;
      }
  }
  const int pe = prof_begin(ws, 8, b);
  hipError_t e = ws->prec == BEC_FP32
                     ? launch_sites<float>(ws->stream, ws->g, ws->tb, p, ws->slot[0], ws->phi,
                                           n_active(ws))
                     : launch_sites<double>(ws->stream, ws->g, ws->tb, p, ws->slot[0], ws->phi,
                                            n_active(ws));
  prof_end(ws, pe);
  HIPCHK(e);
  return BEC_OK;
}
// Fused x pass + lane-per-site pass on psi in place (slot 0).
static int do_xs(bec_ws* ws, const XProgram& p) {
  if (ws->rec) {
    ws->rec->push_back(Rec{R_OTHER});
    return BEC_OK;
  }
  double b = 0.0;
  if (ws->prof_on) {
    b = 2.0 * psi_bytes(ws);
    if (p.stages & XS_WRITE_FHAT) b += half_field_bytes(ws);
    if (p.stages & XS_LOAD_PHI) b += half_field_bytes(ws);
    // complex FFTs of n_x points: D lines per grid row for each tile FFT, two
    // auxiliary lines per row (of the spectral x length) for each DDI stage
    const Geom& g = ws->g;
    auto fl = [](double n) { return 5.0 * n * std::log2(n); };
    const double rows = (double)g.nrows * n_active(ws);
    const int ntile = ((p.stages & XS_INV_FFT) ? 1 : 0) + ((p.stages & XS_FWD_FFT) ? 1 : 0);
    const int naux = ((p.stages & XS_LOAD_PHI) ? 1 : 0) + ((p.stages & XS_WRITE_FHAT) ? 1 : 0);
    ws->prof_fft_flops += rows * (ntile * g.D * fl(g.n[0]) + naux * 2.0 * fl(g.sn[0]));
  }
  const void* src = ws->src_once >= 0 ? ws->slot[ws->src_once] : ws->slot[0];
  ws->src_once = -1;
  const int pe = prof_begin(ws, BEC_K_XSITE, b);
  hipError_t e = ws->prec == BEC_FP32
                     ? launch_xsite<float>(ws->stream, ws->g, ws->tb, p, src, ws->slot[0],
                                           ws->fhat, n_active(ws))
                     : launch_xsite<double>(ws->stream, ws->g, ws->tb, p, src, ws->slot[0],
                                            ws->fhat, n_active(ws));
  prof_end(ws, pe);
  HIPCHK(e);
  return BEC_OK;
}
static int do_strided(bec_ws* ws, int axis, int mode, void* data, double kin_frac, double dt_ovr,
                      int itp, int ddi_data, int* nb = nullptr) {
  if (ws->rec) {
    Rec r{};
    r.kind = R_STRIDED;
    r.axis = axis;
    r.mode = mode;
    r.data = data;
    r.kin_frac = kin_frac;
    r.dt_ovr = dt_ovr;
    r.itp = itp;
    r.ddi = ddi_data;
    ws->rec->push_back(r);
    return BEC_OK;
  }
  CHK(flush_src(ws));
  hipError_t e;
  const int cls = ddi_data ? 5 + (mode == SM_DDI ? 1 : 0)
                           : 1 + (mode == SM_KIN ? 1 : mode == SM_INV ? 2 : mode == SM_KE ? 3 : 0);
  double b = 0.0;
  if (ws->prof_on)
    b = ddi_data ? strided_ddi_bytes(ws, axis, mode) : (mode == SM_KE ? 1.0 : 2.0) * psi_bytes(ws);
  const int pe = prof_begin(ws, cls, b);
  if (ws->prec == BEC_FP32)
    e = launch_strided<float>(ws->stream, ws->g, ws->tb, axis, mode, data, n_active(ws), kin_frac,
                              dt_ovr, itp, ddi_data, nb);
  else
    e = launch_strided<double>(ws->stream, ws->g, ws->tb, axis, mode, data, n_active(ws), kin_frac,
                               dt_ovr, itp, ddi_data, nb);
  prof_end(ws, pe);
  HIPCHK(e);
  return BEC_OK;
}
// 1-D DDI: the Q_ab product of the three half spectra (read and written once)
static int do_ddi_line(bec_ws* ws) {
  if (ws->rec) {
    ws->rec->push_back(Rec{R_OTHER});
    return BEC_OK;
  }
  CHK(flush_src(ws));
  const int pe = prof_begin(ws, 6, ws->prof_on ? 2.0 * half_field_bytes(ws) : 0.0);
  const hipError_t e =
      ws->prec == BEC_FP32
          ? launch_ddi_line<float>(ws->stream, ws->g, ws->tb, ws->fhat, n_active(ws))
          : launch_ddi_line<double>(ws->stream, ws->g, ws->tb, ws->fhat, n_active(ws));
  prof_end(ws, pe);
  HIPCHK(e);
  return BEC_OK;
}
static int do_reduce(bec_ws* ws, int nblocks, uint32_t slots, uint32_t maxmask = 0) {
  if (ws->rec) {
    Rec r{};
    r.kind = R_REDUCE;
    r.nblocks = nblocks;
    r.slots = slots;
    r.maxmask = maxmask;
    ws->rec->push_back(r);
    return BEC_OK;
  }
  CHK(flush_src(ws));
  const int pe = prof_begin(ws, 7);
  HIPCHK(launch_reduce(ws->stream, ws->tb, ws->d_red, n_active(ws), nblocks, slots, maxmask));
  prof_end(ws, pe);
  return BEC_OK;
}


// ---------------------------------------------------------------------------
// Schedule compiler
// ---------------------------------------------------------------------------
enum ItemKind { IT_V, IT_K, IT_NORM, IT_LOSS };
struct Item {
  ItemKind kind;
  double frac;  // V: dt_half = frac*dt ; K: kinetic dt = frac*dt
};

struct Sched {
  bec_ws* ws;
  int itp;
  double dt_ovr;
  XProgram p{};
  SProgram sp{};  // open lane-per-site program (13-component operators)
  bool scale_pending = false;
  // fuse: heavy operators join the x program, which then runs in the fused x/site kernel
  // (heavy: the open x program holds one)
  bool fuse = false;
  bool heavy = false;

  void reset() {
    std::memset(&p, 0, sizeof(p));
    p.itp = itp;
    p.dt_override = dt_ovr;
    heavy = false;
  }
  void reset_s() {
    std::memset(&sp, 0, sizeof(sp));
    sp.itp = itp;
    sp.dt_override = dt_ovr;
  }
  bool empty() const { return p.stages == 0 && p.nA == 0 && p.nB == 0; }
  // Euler spin rotations (DDI, spin mixing for F >= 2) and the tensor step run in the
  // lane-per-site kernel; the x pass keeps FFTs and light pointwise work
  bool site_op(int kind) const {
    return kind == OP_TENSOR || kind == OP_DDI_ROT || kind == OP_RAMAN ||
           (kind == OP_SPINMIX && ws->D >= 5) ||
           (kind == OP_NEMATIC && ws->D >= 13);
  }
  int flush_s() {
    if (sp.n == 0) return BEC_OK;
    int r = do_sites(ws, sp);
    reset_s();
    return r;
  }
  int flush() {
    CHK(flush_s());
    if (empty()) return BEC_OK;
    if (p.nA) p.stages |= XS_OPS_A;
    if (p.nB) p.stages |= XS_OPS_B;
    int r;
    if (heavy) {
      r = do_xs(ws, p);
    } else {
      // passes that only read psi (norm, spin density) do not write it back
      const int ro = XS_NORM | XS_WRITE_FHAT;
      if (p.nA == 0 && p.nB == 0 && (p.stages & ~ro) == 0) p.stages |= XS_NO_STORE;
      r = do_x(ws, p, ws->slot[0], ws->slot[0]);
    }
    reset();
    return r;
  }
  // In real time two adjacent spin-1 spin-mixing steps are one step of the summed time:
  // exp(-i c1 dt F.S) rotates the spinor about its own spin vector F, which the rotation
  // leaves unchanged.  The two halves of _half_potential_step!'s spin mixing are adjacent
  // when no DDI, nematic or Raman step sits between them, and then run as one Rodrigues
  // rotation.  (Only these: they are adjacent within one potential step, so a run's
  // launches do not depend on how its steps are split into calls.  In imaginary time the
  // factor changes F and both halves run as written; so do the Euler rotations of larger
  // spins, whose composed rounding drifts from the reference's two halves by more than the
  // FP64 parity bar over long adaptive runs.)
  bool merge(Op* ops, int n, int kind, double frac) const {
    if (itp || ws->D != 3 || n == 0 || kind != OP_SPINMIX || ops[n - 1].kind != kind)
      return false;
    ops[n - 1].frac += frac;
    return true;
  }
  int add_op(int kind, double frac) {
    if (scale_pending) {
      scale_pending = false;
      CHK(add_op(OP_SCALE, 0.0));
    }
    if (fuse) {  // one op list; the x stages after it close the program
      if (p.stages & (XS_FWD_FFT | XS_NORM | XS_WRITE_FHAT)) CHK(flush());
      if (merge(p.A, p.nA, kind, frac)) return BEC_OK;
      if (p.nA >= kMaxOps) CHK(flush());
      p.A[p.nA].kind = kind;
      p.A[p.nA].frac = frac;
      ++p.nA;
      if (site_op(kind)) heavy = true;
      return BEC_OK;
    }
    if (sp.n > 0 || site_op(kind)) {
      if (!empty()) {  // the x pass before the site kernel must store psi first
        if (p.nA) p.stages |= XS_OPS_A;
        if (p.nB) p.stages |= XS_OPS_B;
        CHK(do_x(ws, p, ws->slot[0], ws->slot[0]));
        reset();
      }
      if (merge(sp.ops, sp.n, kind, frac)) return BEC_OK;
      if (sp.n >= kMaxOps) CHK(flush_s());
      sp.ops[sp.n].kind = kind;
      sp.ops[sp.n].frac = frac;
      ++sp.n;
      return BEC_OK;
    }
    if (p.stages & (XS_FWD_FFT | XS_NORM | XS_WRITE_FHAT)) CHK(flush());
    {
      const bool toB = (p.stages & XS_KIN_1D) != 0;
      if (merge(toB ? p.B : p.A, toB ? p.nB : p.nA, kind, frac)) return BEC_OK;
    }
    if (((p.stages & XS_KIN_1D) ? p.nB : p.nA) >= kMaxOps) CHK(flush());
    const bool toB = (p.stages & XS_KIN_1D) != 0;
    Op* ops = toB ? p.B : p.A;
    int& n = toB ? p.nB : p.nA;
    ops[n].kind = kind;
    ops[n].frac = frac;
    ++n;
    return BEC_OK;
  }
  int ddi_conv() {
    // F_hat half rows were written by the launch just flushed
    const int nd = ws->g.ndim;
    if (nd == 3) {
      CHK(do_strided(ws, 1, SM_FWD, ws->fhat, 0, 0, itp, 1));
      CHK(do_strided(ws, 2, SM_DDI, ws->fhat, 0, 0, itp, 1));
      CHK(do_strided(ws, 1, SM_INV, ws->fhat, 0, 0, itp, 1));
    } else if (nd == 1) {
      CHK(do_ddi_line(ws));
    } else {
      CHK(do_strided(ws, 1, SM_DDI, ws->fhat, 0, 0, itp, 1));
    }
    return BEC_OK;
  }
  // _half_potential_step! (split_step.jl:48-133) for dt_half = f*dt
  // spin mixing is the identity when |c1| < 1e-30 (spin_mixing.jl:9): with every active
  // instance in that case the operator (and, for large spins, its lane-per-site pass) is
  // left out of the schedule
  bool mixing_active() const {
    if (ws->tensor) return true;
    for (int inst : ws->active)
      if (std::fabs(ws->hp[inst].c1) >= 1e-30) return true;
    return false;
  }
  // the Raman step runs when a RamanCoupling is present (split_step.jl:77-83,101-107)
  bool raman_active() const {
    for (int inst : ws->active)
      if (ws->hp[inst].has_raman) return true;
    return false;
  }
  // likewise the nematic step with |c2| < 1e-30 (nematic.jl:43); never with the tensor step
  bool nematic_active() const {
    if (ws->tensor) return false;
    for (int inst : ws->active)
      if (std::fabs(ws->hp[inst].c2) >= 1e-30) return true;
    return false;
  }
  int V(double f) {
    const int mix = ws->tensor ? OP_TENSOR : OP_SPINMIX;
    const bool mixing = mixing_active();
    const bool nematic = nematic_active();
    CHK(add_op(OP_DIAG, f / 2));
    if (mixing) CHK(add_op(mix, f / 2));
    if (nematic) CHK(add_op(OP_NEMATIC, f / 2));
    const bool raman = raman_active();
    if (raman) CHK(add_op(OP_RAMAN, f / 2));
    if (ws->ddi) {
      CHK(flush_s());
      if (p.stages & (XS_FWD_FFT | XS_NORM | XS_WRITE_FHAT)) CHK(flush());
      p.stages |= XS_WRITE_FHAT;
      CHK(flush());
      CHK(ddi_conv());
      if (fuse) {
        // C2R of Phi in LDS of the launch that rotates (the op list is empty here, so the
        // rotation cannot be split from it)
        p.stages |= XS_LOAD_PHI;
      } else {
        // C2R of Phi into its real-space buffer; the rotation runs per site
        p.stages |= XS_LOAD_PHI | XS_STORE_PHI | XS_NO_LOAD | XS_NO_STORE;
        CHK(flush());
      }
      CHK(add_op(OP_DDI_ROT, f));
    }
    if (raman) CHK(add_op(OP_RAMAN, f / 2));
    if (nematic) CHK(add_op(OP_NEMATIC, f / 2));
    if (mixing) CHK(add_op(mix, f / 2));
    CHK(add_op(OP_DIAG, f / 2));
    return BEC_OK;
  }
  int K(double f) {
    if (scale_pending) {  // renormalise before the (linear) kinetic step
      scale_pending = false;
      CHK(add_op(OP_SCALE, 0.0));
    }
    CHK(flush_s());
    if (ws->g.ndim == 1) {
      if (p.stages & (XS_KIN_1D | XS_FWD_FFT | XS_NORM | XS_WRITE_FHAT)) CHK(flush());
      p.stages |= XS_KIN_1D;
      p.kin_frac = f;
      return BEC_OK;
    }
    p.stages |= XS_FWD_FFT;
    CHK(flush());
    if (ws->g.ndim == 3) {
      CHK(do_strided(ws, 1, SM_FWD, ws->slot[0], 0, 0, itp, 0));
      CHK(do_strided(ws, 2, SM_KIN, ws->slot[0], f, dt_ovr, itp, 0));
      CHK(do_strided(ws, 1, SM_INV, ws->slot[0], 0, 0, itp, 0));
    } else {
      CHK(do_strided(ws, 1, SM_KIN, ws->slot[0], f, dt_ovr, itp, 0));
    }
    p.stages |= XS_INV_FFT;
    return BEC_OK;
  }
  // apply_loss_step! (losses.jl:25-49) for dt = f*dt; fuses into the next x pass
  int LOSS(double f) { return add_op(OP_LOSS, f); }
  int NORM() {
    if (scale_pending) {
      scale_pending = false;
      CHK(add_op(OP_SCALE, 0.0));
    }
    CHK(flush_s());
    p.stages |= XS_NORM;
    CHK(flush());
    CHK(do_reduce(ws, xpass_grid(ws->g), 1u << RS_NORM));
    scale_pending = true;
    return BEC_OK;
  }
  int finish() {
    CHK(flush());
    if (scale_pending) {
      scale_pending = false;
      if (ws->rec) {
        ws->rec->push_back(Rec{R_SCALE});
        return BEC_OK;
      }
      CHK(flush_src(ws));
      hipError_t e = ws->prec == BEC_FP32
                         ? launch_scale_norm<float>(ws->stream, ws->g, ws->tb, ws->slot[0], n_active(ws))
                         : launch_scale_norm<double>(ws->stream, ws->g, ws->tb, ws->slot[0], n_active(ws));
      HIPCHK(e);
    }
    return BEC_OK;
  }
};

static int run_items(bec_ws* ws, const std::vector<Item>& items, int itp, double dt_ovr) {
  if (n_active(ws) == 0) return BEC_OK;
  CHK(sync_params(ws));
  Sched s;
  s.ws = ws;
  s.itp = itp;
  s.dt_ovr = dt_ovr;
  s.fuse = ws->xsite && xsite_ok(ws->g);
  s.reset();
  s.reset_s();
  for (const Item& it : items) {
    if (it.kind == IT_V) CHK(s.V(it.frac));
    else if (it.kind == IT_K) CHK(s.K(it.frac));
    else if (it.kind == IT_LOSS) CHK(s.LOSS(it.frac));
    else CHK(s.NORM());
  }
  return s.finish();
}

// Floating-point operations of one tensor matvec plus its Taylor update, as written in
// tensor_matvec / op_tensor (complex*complex = 6, complex fma = 8, complex*real fma = 4;
// zero Clebsch-Gordan entries are skipped at compile time).
template <int F>
static double tensor_matvec_flops() {
  constexpr tens::CGTab<F> cg{};
  double f = 0.0;
  for (int M = -2 * F; M <= 2 * F; ++M) {
    const int ns = tens::ns(F, M);
    for (int m1 = tens::m1lo(F, M); m1 <= tens::m1hi(M); ++m1) {
      const bool off = m1 != M - m1;
      int nz = 0;
      for (int k = 0; k < ns; ++k) nz += cg.v[tens::off(F, M, m1, tens::s0(M)) + k] != 0.0;
      f += off ? 6 + 8 : 6;  // pair product(s) psi * v
      f += 8.0 * nz;         // B_SM and C accumulations
      f += off ? 16 : 8;     // w update(s) conj(psi) * C
    }
    f += 2.0 * ns;  // B_SM *= g_S
  }
  return f + 10.0 * (2 * F + 1);  // v = -i f w, acc += v, |v|_1 and |acc|_1
}
// Flops a site's first matvec (v = psi) saves in either form: per pair with m1 != m2 the
// second complex product and the sum become one doubling (6 + 8 -> 6 + 2).
template <int F>
static double first_term_saving() {
  double f = 0.0;
  for (int M = -2 * F; M <= 2 * F; ++M)
    for (int i = 0; i < tens::npairs(F, M); ++i)
      if (tens::m1lo(F, M) + i != M - (tens::m1lo(F, M) + i)) f += 6.0;
  return f;
}

// The same for the channel-folded matvec of the FP64 path (tensor_matvec_folded): pair
// products, K_M P (complex*real: 2 flop for the first entry of a row, 4 after), w updates.
template <int F>
static double tensor_matvec_folded_flops() {
  double f = 0.0;
  for (int M = -2 * F; M <= 2 * F; ++M) {
    const int np = tens::npairs(F, M);
    for (int i = 0; i < np; ++i) {
      const int m1 = tens::m1lo(F, M) + i;
      const bool off = m1 != M - m1;
      f += off ? 6 + 8 : 6;            // P_i
      f += 2.0 + 4.0 * (np - 1);       // C_i = sum_j K_ij P_j
      f += off ? 16 : 8;               // w update(s) conj(psi) * C
    }
  }
  return f + 10.0 * (2 * F + 1);  // v = -i f w, acc += v, |v|_1 and |acc|_1
}

// One Euler rotation (op_euler): 4 products of the real matrix d^F(pi/2) (or its
// transpose) with a complex spinor, as dhalf_apply forms them (the 2F sums and differences
// v_c +- v_{D-1-c}, 2 flop each, then 4 flop per nonzero entry of each row's half: c <= F
// for an even row, c < F for an odd one, all D columns for a row of neither parity), and
// 5 diagonal phase sweeps (F - 1 powers and 2F component products, 6 flop each).
template <int F, bool TRANS>
static double dhalf_flops() {
  constexpr int D = 2 * F + 1;
  constexpr tens::DHalf<F> d{};
  constexpr tens::DHalfParity<F, TRANS> par{};
  double f = 4.0 * F;
  for (int i = 0; i < D; ++i) {
    const int ncol = par.s[i] == 1 ? F + 1 : par.s[i] == -1 ? F : D;
    for (int j = 0; j < ncol; ++j) f += (TRANS ? d.v[j * D + i] : d.v[i * D + j]) != 0.0 ? 4.0 : 0.0;
  }
  return f;
}
template <int F>
static double euler_flops() {
  return 2.0 * (dhalf_flops<F, false>() + dhalf_flops<F, true>()) + 30.0 * (F > 0 ? 3 * F - 1 : 0);
}

// K_M[i][j] = sum_S CG(m1_i, M - m1_i | S M) g_S CG(m1_j, M - m1_j | S M) for every M, in
// tens::kfold_off order (the kernel of tensor_matvec_folded), from the compile-time table.
template <int F>
static void fold_channels_f(const double* gS, double* out) {
  constexpr tens::CGTab<F> cg{};
  for (int M = -2 * F; M <= 2 * F; ++M) {
    const int lo = tens::m1lo(F, M), np = tens::npairs(F, M), s0 = tens::s0(M);
    const int ns = tens::ns(F, M);
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) {
        double k = 0.0;
        for (int q = 0; q < ns; ++q)
          k += cg.v[tens::off(F, M, lo + i, s0) + q] * gS[s0 / 2 + q] *
               cg.v[tens::off(F, M, lo + j, s0) + q];
        out[tens::kfold_off(F, M) + i * np + j] = k;
      }
  }
}
static void fold_channels(int F, const double* gS, double* out) {
  switch (F) {
    case 1: fold_channels_f<1>(gS, out); break;
    case 2: fold_channels_f<2>(gS, out); break;
    case 3: fold_channels_f<3>(gS, out); break;
    case 4: fold_channels_f<4>(gS, out); break;
    case 5: fold_channels_f<5>(gS, out); break;
    case 6: fold_channels_f<6>(gS, out); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* bec_last_error(void) { return g_err.c_str(); }

// Geometry of a workspace with D components on the grid c describes (validated by the caller).
static void fill_geom(const bec_config* c, int D, Geom& g) {
  g = Geom{};
  g.ndim = c->ndim;
  for (int d = 0; d < 3; ++d) g.n[d] = d < c->ndim ? c->n[d] : 1;
  g.D = D;
  g.nrows = g.n[1] * g.n[2];
  g.nsites = (long long)g.n[0] * g.n[1] * g.n[2];
  g.inst_stride = g.nsites * D;
  g.inv_ntot = 1.0 / (double)g.nsites;
  double dV = 1.0;
  for (int d = 0; d < c->ndim; ++d) dV *= c->L[d] / c->n[d];
  g.dV = dV;
  g.nmax = std::max(g.n[0], std::max(g.n[1], g.n[2]));
  g.spad = c->enable_ddi && c->ddi_padding && c->F > 0 ? 1 : 0;
  for (int d = 0; d < 3; ++d) g.sn[d] = d < c->ndim ? g.n[d] * (g.spad ? 2 : 1) : 1;
  g.smax = std::max(g.sn[0], std::max(g.sn[1], g.sn[2]));
  g.nxh = g.sn[0] / 2 + 1;
  g.nxh_pitch = (g.nxh + 15) / 16 * 16;
  g.fhat_field_stride = (long long)g.nxh_pitch * g.sn[1] * g.sn[2];
  g.fhat_inst_stride = 3 * g.fhat_field_stride;
  g.ddi_secular = c->ddi_secular;
}


// The same grid with the unpadded DDI spectra (energy evaluation).
static Geom unpadded_geom(const Geom& g0) {
  Geom g = g0;
  g.spad = 0;
  for (int d = 0; d < 3; ++d) g.sn[d] = g.n[d];
  g.smax = g.nmax;
  g.nxh = g.n[0] / 2 + 1;
  g.nxh_pitch = (g.nxh + 15) / 16 * 16;
  g.fhat_field_stride = (long long)g.nxh_pitch * g.n[1] * g.n[2];
  g.fhat_inst_stride = 3 * g.fhat_field_stride;
  return g;
}

static long long ws_bytes(const bec_config* c, int* max_blocks_out, int* nxb_out) {
  const int D = 2 * c->F + 1;
  long long ns = 1;
  for (int d = 0; d < 3; ++d) ns *= (d < c->ndim ? c->n[d] : 1);
  const long long csz = c->precision == BEC_FP32 ? 8 : 16;
  const int B = c->n_instances;
  const int nx = c->n[0];
  const int ny = c->ndim >= 2 ? c->n[1] : 1, nz = c->ndim >= 3 ? c->n[2] : 1;
  const int nrows = ny * nz;
  Geom g;
  fill_geom(c, D, g);
  const Geom gu = unpadded_geom(g);
  const int nxb = xpass_grid(g);
  // the tile autotuner may pick down to a quarter of the default rows per block
  int maxb = std::max(std::min(nrows, 4 * nxb), 1024);
  maxb = std::max(maxb, xpass_grid(gu));  // energy evaluation on the unpadded spectra
  // strided SM_KE partial count (the energy runs in FP64 for both precisions)
  for (int axis = 1; axis < c->ndim; ++axis) {
    const int N = axis == 1 ? ny : nz;
    const int other = axis == 1 ? nz : ny;
    for (int cs : {(int)csz, 16}) {
      const int tbw = std::max(1, strided_tb(N, 1, cs));
      maxb = std::max(maxb, ((nx + tbw - 1) / tbw) * other * D);
    }
  }
  if (max_blocks_out) *max_blocks_out = maxb;
  if (nxb_out) *nxb_out = nxb;
  long long bytes = 2LL * B * D * ns * csz;  // psi + scratch
  if (c->enable_ddi) {
    const int nxh = nx / 2 + 1;
    const int pitch = (nxh + 15) / 16 * 16;
    const int pf = c->ddi_padding ? 2 : 1;  // padded spectra: (2nx/2+1) x 2ny x 2nz
    const int ppitch = (pf * nx / 2 + 1 + 15) / 16 * 16;
    bytes += 3LL * B * ppitch * (pf * ny) * (c->ndim >= 3 ? pf * nz : nz) * csz;
    bytes += 3LL * B * ns * (csz / 2);
    if (c->precision == BEC_FP32) bytes += 3LL * B * pitch * ny * nz * 16;
  }
  if (c->precision == BEC_FP32) bytes += 2LL * B * D * ns * 16;  // FP64 energy copies
  bytes += (long long)RS_COUNT * B * maxb * 8 + (long long)B * RS_COUNT * 8;
  bytes += (long long)B * 3 * std::max(nx, std::max(ny, nz)) * 8;
  return bytes;
}

long long bec_workspace_bytes(const bec_config* cfg) {
  if (!cfg || cfg->ndim < 1 || cfg->ndim > 3 || cfg->F < 0 || cfg->F > 6) return -1;
  for (int d = 0; d < cfg->ndim; ++d)
    if (cfg->n[d] <= 0) return -1;
  return ws_bytes(cfg, nullptr, nullptr);
}

// Bluestein tables of a runtime line length N with M-point convolutions (layout as the
// device reads them: chirp, chirp transform / M, M-point twiddles; interleaved re, im):
// c[n] = exp(i pi n^2 / N) with n^2 reduced mod 2N (exact phase), the wrapped chirp
// (c[m] at m and M - m, zero between) transformed by a radix-2 FFT in long double.
static std::vector<double> bluestein_tables(int N, int M) {
  const long double pi = 3.14159265358979323846264338327950288L;
  std::vector<long double> cr(N), ci(N);
  for (int n = 0; n < N; ++n) {
    const long long q = ((long long)n * n) % (2LL * N);
    cr[n] = cosl(pi * q / N);
    ci[n] = sinl(pi * q / N);
  }
  std::vector<long double> wr(M, 0.0L), wi(M, 0.0L);
  for (int n = 0; n < N; ++n) {
    wr[n] = cr[n];
    wi[n] = ci[n];
    if (n > 0) {
      wr[M - n] = cr[n];
      wi[M - n] = ci[n];
    }
  }
  // in-place iterative radix-2 FFT (forward, exp(-2 pi i jk / M))
  for (int i = 1, j = 0; i < M; ++i) {
    int bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(wr[i], wr[j]);
      std::swap(wi[i], wi[j]);
    }
  }
  for (int len = 2; len <= M; len <<= 1) {
    for (int i = 0; i < M; i += len) {
      for (int k = 0; k < len / 2; ++k) {
        const long double ang = -2.0L * pi * k / len;
        const long double tr = cosl(ang), ti = sinl(ang);
        const long double xr = wr[i + k + len / 2] * tr - wi[i + k + len / 2] * ti;
        const long double xi = wr[i + k + len / 2] * ti + wi[i + k + len / 2] * tr;
        wr[i + k + len / 2] = wr[i + k] - xr;
        wi[i + k + len / 2] = wi[i + k] - xi;
        wr[i + k] += xr;
        wi[i + k] += xi;
      }
    }
  }
  std::vector<double> out(2 * ((size_t)N + 2 * (size_t)M));
  size_t o = 0;
  for (int n = 0; n < N; ++n) {
    out[o++] = (double)cr[n];
    out[o++] = (double)ci[n];
  }
  for (int k = 0; k < M; ++k) {
    out[o++] = (double)(wr[k] / M);
    out[o++] = (double)(wi[k] / M);
  }
  for (int m = 0; m < M; ++m) {
    const long double ang = -2.0L * pi * m / M;
    out[o++] = (double)cosl(ang);
    out[o++] = (double)sinl(ang);
  }
  return out;
}

int bec_ws_create(const bec_config* c, bec_ws** out) {
  if (!c || !out) return fail(BEC_EINVAL, "null argument");
  *out = nullptr;
  if (c->ndim < 1 || c->ndim > 3) return fail(BEC_EINVAL, "ndim must be 1, 2 or 3");
  for (int d = 0; d < c->ndim; ++d) {
    if (c->n[d] <= 0 || (c->n[d] % 2) != 0)
      return fail(BEC_EINVAL, "n_points must be positive even integers");
    if (!(c->L[d] > 0)) return fail(BEC_EINVAL, "box_size must be positive");
  }
  if (c->F < 0) return fail(BEC_EINVAL, "F must be non-negative");  // types.jl:43
  if (c->F > 6) return fail(BEC_EUNSUP, "F above 6 (more than 13 components) is not built");
  if (c->n_instances < 1) return fail(BEC_EINVAL, "n_instances must be >= 1");
  if (c->precision != BEC_FP64 && c->precision != BEC_FP32)
    return fail(BEC_EINVAL, "precision must be BEC_FP64 or BEC_FP32");
  if (c->ddi_padding && !c->enable_ddi) return fail(BEC_EINVAL, "ddi_padding needs enable_ddi");
  if (!c->x_tab || !c->k_tab) return fail(BEC_EINVAL, "missing tables");
  {
    // site indices are 32-bit within an instance; every line pass holds whole lines (all
    // 2F+1 components of a row for the x pass) in one block's LDS
    long long ns = 1;
    for (int d = 0; d < c->ndim; ++d) ns *= c->n[d];
    if (ns * (2 * c->F + 1) >= (1LL << 31))
      return fail(BEC_EUNSUP, "more than 2^31 complex values per instance");
    Geom g0;
    fill_geom(c, 2 * c->F + 1, g0);
    if (!generic_fits(g0, c->precision == BEC_FP32 ? 8 : 16, c->enable_ddi && c->F > 0))
      return fail(BEC_EUNSUP, "grid side too long: one x row of all 2F+1 components (with the "
                              "dipolar lines) or one y/z line does not fit a block's LDS");
  }

  bec_ws* ws = new bec_ws();
  ws->F = c->F;
  ws->D = 2 * c->F + 1;
  ws->B = c->n_instances;
  ws->prec = c->precision;
  ws->itp = c->imaginary_time;
  ws->normalize_every = c->normalize_every;
  // The scalar gas (F = 0) has no spin density, so its dipolar potential and the rotation
  // by it are exactly zero and the identity: the convolution is left out (ddi.jl:108-201)
  ws->ddi = c->enable_ddi && c->F > 0;
  ws->tensor = c->tensor;
  if (ws->tensor && ws->F < 1) {
    delete ws;
    return fail(BEC_EINVAL, "the tensor step needs F >= 1");
  }
  ws->csz = ws->prec == BEC_FP32 ? 8 : 16;
  Geom& g = ws->g;
  fill_geom(c, ws->D, g);
  ws_bytes(c, &ws->max_blocks, &ws->n_xblocks);
  {
    const char* e = getenv("BEC_XSITE");
    ws->xsite = !(e && atoi(e) == 0);
  }

  ws->hp.assign(ws->B, InstParams{});
  ws->user.assign(ws->B, bec_inst_params{});
  ws->has_varr.assign(ws->B, 0);
  ws->active.resize(ws->B);
  for (int i = 0; i < ws->B; ++i) ws->active[i] = i;
  int off = 0;
  for (int d = 0; d < c->ndim; ++d) {
    ws->xtab.insert(ws->xtab.end(), c->x_tab + off, c->x_tab + off + c->n[d]);
    ws->ktab.insert(ws->ktab.end(), c->k_tab + off, c->k_tab + off + c->n[d]);
    off += c->n[d];
  }

  auto cleanup = [&](int code, const std::string& m) {
    bec_ws_destroy(ws);
    return fail(code, m);
  };
  if (hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(BEC_EDEVICE, "hipStreamCreate failed");
  const size_t psi_bytes = ws->csz * (size_t)g.inst_stride * ws->B;
  if (hipMalloc(&ws->slot[0], psi_bytes) != hipSuccess) return cleanup(BEC_ENOMEM, "psi alloc");
  if (hipMalloc(&ws->scratch, psi_bytes) != hipSuccess) return cleanup(BEC_ENOMEM, "scratch alloc");
  if (hipMemset(ws->slot[0], 0, psi_bytes) != hipSuccess) return cleanup(BEC_EDEVICE, "memset");
  if (ws->ddi) {
    const size_t fb = ws->csz * (size_t)g.fhat_inst_stride * ws->B;
    if (hipMalloc(&ws->fhat, fb) != hipSuccess) return cleanup(BEC_ENOMEM, "fhat alloc");
    if (hipMemset(ws->fhat, 0, fb) != hipSuccess) return cleanup(BEC_EDEVICE, "memset");
    const size_t pb = (ws->csz / 2) * 3 * (size_t)g.nsites * ws->B;
    if (hipMalloc(&ws->phi, pb) != hipSuccess) return cleanup(BEC_ENOMEM, "phi alloc");
  }
  // k table [3][nmax]
  std::vector<double> kt(3 * g.nmax, 0.0);
  off = 0;
  for (int d = 0; d < c->ndim; ++d) {
    for (int i = 0; i < c->n[d]; ++i) kt[d * g.nmax + i] = c->k_tab[off + i];
    off += c->n[d];
  }
  if (hipMalloc(&ws->d_ktab, kt.size() * 8) != hipSuccess) return cleanup(BEC_ENOMEM, "ktab");
  if (hipMemcpy(ws->d_ktab, kt.data(), kt.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(BEC_EDEVICE, "ktab copy");
  {  // x table [3][nmax]
    std::vector<double> xt(3 * g.nmax, 0.0);
    int xo = 0;
    for (int d = 0; d < c->ndim; ++d) {
      for (int i = 0; i < c->n[d]; ++i) xt[d * g.nmax + i] = c->x_tab[xo + i];
      xo += c->n[d];
    }
    if (hipMalloc(&ws->d_xtab, xt.size() * 8) != hipSuccess) return cleanup(BEC_ENOMEM, "xtab");
    if (hipMemcpy(ws->d_xtab, xt.data(), xt.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(BEC_EDEVICE, "xtab copy");
  }
  if (g.spad) {
    // k of the padded grid (same dx, twice the points): fftfreq(2n) * 2 pi / (2n dx)
    std::vector<double> sk(3 * g.smax, 0.0);
    for (int d = 0; d < c->ndim; ++d) {
      const int ns = g.sn[d];
      const double dk = 2.0 * M_PI / (ns * (c->L[d] / c->n[d]));
      for (int m = 0; m < ns; ++m) sk[d * g.smax + m] = (m < ns / 2 ? m : m - ns) * dk;
    }
    if (hipMalloc(&ws->d_sktab, sk.size() * 8) != hipSuccess) return cleanup(BEC_ENOMEM, "sktab");
    if (hipMemcpy(ws->d_sktab, sk.data(), sk.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(BEC_EDEVICE, "sktab copy");
  }
  const size_t axb = (size_t)ws->B * 3 * g.nmax * 8;
  if (hipMalloc(&ws->d_axtab, axb) != hipSuccess) return cleanup(BEC_ENOMEM, "axtab");
  if (hipMemset(ws->d_axtab, 0, axb) != hipSuccess) return cleanup(BEC_EDEVICE, "memset");
  if (hipMalloc(&ws->d_params, sizeof(InstParams) * ws->B) != hipSuccess)
    return cleanup(BEC_ENOMEM, "params");
  if (ws->tensor) {
    ws->h_kfold.assign((size_t)ws->B * tens::kKfoldMax, 0.0);
    if (hipMalloc(&ws->d_kfold, ws->h_kfold.size() * 8) != hipSuccess)
      return cleanup(BEC_ENOMEM, "tensor kernels");
    if (hipMemset(ws->d_kfold, 0, ws->h_kfold.size() * 8) != hipSuccess)
      return cleanup(BEC_EDEVICE, "memset");
  }
  if (hipMalloc(&ws->d_active, sizeof(int) * ws->B) != hipSuccess) return cleanup(BEC_ENOMEM, "active");
  if (hipMemcpy(ws->d_active, ws->active.data(), sizeof(int) * ws->B, hipMemcpyHostToDevice) !=
      hipSuccess)
    return cleanup(BEC_EDEVICE, "active copy");
  const size_t redb = (size_t)ws->B * RS_COUNT * 8;
  if (hipMalloc(&ws->d_red, redb) != hipSuccess) return cleanup(BEC_ENOMEM, "red");
  if (hipMemset(ws->d_red, 0, redb) != hipSuccess) return cleanup(BEC_EDEVICE, "memset");
  const size_t pb = (size_t)RS_COUNT * ws->B * ws->max_blocks * 8;
  if (hipMalloc(&ws->d_partials, pb) != hipSuccess) return cleanup(BEC_ENOMEM, "partials");
  if (hipMalloc(&ws->d_work, kWorkBytes) != hipSuccess)
    return cleanup(BEC_ENOMEM, "work counters");
  if (hipMemset(ws->d_work, 0, kWorkBytes) != hipSuccess) return cleanup(BEC_EDEVICE, "memset");
  if (hipMemset(ws->d_partials, 0, pb) != hipSuccess) return cleanup(BEC_EDEVICE, "memset");
  if (hipHostMalloc(&ws->h_red, redb) != hipSuccess) return cleanup(BEC_ENOMEM, "pinned red");
  // twiddles exp(-2 pi i m / N), computed in long double and stored in FP64 for both
  // precisions (FP32-bulk butterflies multiply by them in FP64 with one rounding)
  for (int l = 0; l < kNumFftSizes; ++l) {
    const int N = kFftSizes[l];
    std::vector<double> buf(2 * N);
    for (int m = 0; m < N; ++m) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * m / N;
      buf[2 * m] = (double)cosl(ang);
      buf[2 * m + 1] = (double)sinl(ang);
    }
    if (hipMalloc(&ws->d_tw[l], buf.size() * 8) != hipSuccess) return cleanup(BEC_ENOMEM, "twiddles");
    if (hipMemcpy(ws->d_tw[l], buf.data(), buf.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(BEC_EDEVICE, "twiddle copy");
  }
  {
    // runtime-length tables: every side and zero-padded spectral length of the grid
    std::vector<int> lens;
    for (int d = 0; d < c->ndim; ++d) {
      lens.push_back(g.n[d]);
      if (g.spad) lens.push_back(g.sn[d]);
    }
    std::sort(lens.begin(), lens.end());
    lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
    for (size_t l = 0; l < lens.size() && l < (size_t)kGenTabs; ++l) {
      const int N = lens[l];
      std::vector<double> buf(2 * N);
      for (int m = 0; m < N; ++m) {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * m / N;
        buf[2 * m] = (double)cosl(ang);
        buf[2 * m + 1] = (double)sinl(ang);
      }
      if (hipMalloc(&ws->d_gtw[l], buf.size() * 8) != hipSuccess)
        return cleanup(BEC_ENOMEM, "twiddles");
      if (hipMemcpy(ws->d_gtw[l], buf.data(), buf.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(BEC_EDEVICE, "twiddle copy");
      ws->gtw_n[l] = N;
      const int M = gen_blue_m(N);
      if (M > 0) {
        const std::vector<double> bt = bluestein_tables(N, M);
        if (hipMalloc(&ws->d_gbl[l], bt.size() * 8) != hipSuccess)
          return cleanup(BEC_ENOMEM, "Bluestein tables");
        if (hipMemcpy(ws->d_gbl[l], bt.data(), bt.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
          return cleanup(BEC_EDEVICE, "Bluestein table copy");
        ws->gbl_m[l] = M;
      }
    }
  }
  refresh_tables(ws);
  *out = ws;
  return BEC_OK;
}

void bec_ws_destroy(bec_ws* ws) {
  if (!ws) return;
  if (ws->stream) (void)hipStreamSynchronize(ws->stream);
  for (int i = 0; i < 4; ++i)
    if (ws->slot[i]) (void)hipFree(ws->slot[i]);
  if (ws->scratch) (void)hipFree(ws->scratch);
  if (ws->fhat) (void)hipFree(ws->fhat);
  if (ws->phi) (void)hipFree(ws->phi);
  for (void* q : {ws->psi64, ws->scratch64, ws->fhat64, ws->phi64})
    if (q) (void)hipFree(q);
  for (double* p : {ws->d_ktab, ws->d_xtab, ws->d_sktab, ws->d_axtab, ws->d_varr, ws->d_red,
                    ws->d_partials})
    if (p) (void)hipFree(p);
  if (ws->d_params) (void)hipFree(ws->d_params);
  if (ws->d_active) (void)hipFree(ws->d_active);
  if (ws->d_work) (void)hipFree(ws->d_work);
  for (double* q : {ws->d_cpart, ws->d_ctarget, ws->d_cfac, ws->d_sdens, ws->d_kfold})
    if (q) (void)hipFree(q);
  for (int i = 0; i < kNumFftSizes; ++i)
    if (ws->d_tw[i]) (void)hipFree(ws->d_tw[i]);
  for (int i = 0; i < kGenTabs; ++i) {
    if (ws->d_gtw[i]) (void)hipFree(ws->d_gtw[i]);
    if (ws->d_gbl[i]) (void)hipFree(ws->d_gbl[i]);
  }
  if (ws->h_red) (void)hipHostFree(ws->h_red);
  if (ws->h_red2) (void)hipHostFree(ws->h_red2);
  for (hipEvent_t e : ws->ev_pool) (void)hipEventDestroy(e);
  drop_graphs(ws);
  if (ws->stream) (void)hipStreamDestroy(ws->stream);
  delete ws;
}

int bec_set_params(bec_ws* ws, int inst, const bec_inst_params* p) {
  if (!ws || !p) return fail(BEC_EINVAL, "null argument");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  if (!(p->dt > 0)) return fail(BEC_EINVAL, "dt must be positive");
  ws->user[inst] = *p;
  InstParams& q = ws->hp[inst];
  const InstParams keep = q;  // tensor couplings, loss rates, Raman have their own setters
  q = InstParams{};
  std::memcpy(q.gS, keep.gS, sizeof(q.gS));
  std::memcpy(q.loss, keep.loss, sizeof(q.loss));
  q.has_loss = keep.has_loss;
  std::memcpy(q.raman, keep.raman, sizeof(q.raman));
  q.has_raman = keep.has_raman;
  q.dt = p->dt;
  q.c0 = p->c0;
  q.c1 = p->c1;
  q.c_lhy = p->c_lhy;
  q.c2 = ws->tensor ? 0.0 : p->c2;
  q.cdd = p->c_dd;
  const int F = ws->F;
  double zmin = 0.0;
  for (int c = 0; c < ws->D; ++c) {
    const double m = (double)(F - c);
    q.zee[c] = -p->p * m + p->q * (m * m);
    zmin = c == 0 ? q.zee[c] : std::min(zmin, q.zee[c]);
  }
  q.zee_shift = zmin;
  q.active = 1;
  ws->params_dirty = true;
  // trap tables omega_d^2 x_i^2
  const Geom& g = ws->g;
  std::vector<double> ax(3 * g.nmax, 0.0);
  int off = 0;
  for (int d = 0; d < g.ndim; ++d) {
    const double w2 = p->omega[d] * p->omega[d];
    for (int i = 0; i < g.n[d]; ++i) {
      const double x = ws->xtab[off + i];
      ax[d * g.nmax + i] = w2 * (x * x);
    }
    off += g.n[d];
  }
  HIPCHK(hipMemcpyAsync(ws->d_axtab + (size_t)inst * 3 * g.nmax, ax.data(), ax.size() * 8,
                        hipMemcpyHostToDevice, ws->stream));
  HIPCHK(hipStreamSynchronize(ws->stream));
  return BEC_OK;
}

int bec_set_tensor(bec_ws* ws, int inst, const double* gS, int n) {
  if (!ws || !gS) return fail(BEC_EINVAL, "null argument");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  if (!ws->tensor) return fail(BEC_EINVAL, "workspace was created without the tensor step");
  if (n != ws->F + 1) return fail(BEC_EINVAL, "expected F+1 channel couplings (S = 0, 2, ..., 2F)");
  for (int k = 0; k < 7; ++k) ws->hp[inst].gS[k] = k < n ? gS[k] : 0.0;
  fold_channels(ws->F, gS, ws->h_kfold.data() + (size_t)inst * tens::kKfoldMax);
  ws->params_dirty = true;
  return BEC_OK;
}

int bec_set_potential(bec_ws* ws, int inst, const double* V) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  const Geom& g = ws->g;
  if (!V) {
    ws->has_varr[inst] = 0;
    if (std::any_of(ws->has_varr.begin(), ws->has_varr.end(), [](int v) { return v; }))
      return fail(BEC_EUNSUP, "mixing explicit and harmonic potentials across instances");
    refresh_tables(ws);
    return BEC_OK;
  }
  if (!ws->d_varr) {
    if (hipMalloc(&ws->d_varr, (size_t)ws->B * g.nsites * 8) != hipSuccess)
      return fail(BEC_ENOMEM, "potential alloc");
  }
  HIPCHK(hipMemcpy(ws->d_varr + (size_t)inst * g.nsites, V, g.nsites * 8, hipMemcpyHostToDevice));
  ws->has_varr[inst] = 1;
  refresh_tables(ws);
  return BEC_OK;
}

int bec_set_psi(bec_ws* ws, int inst, const double* psi) {
  if (!ws || !psi) return fail(BEC_EINVAL, "null argument");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  const size_t n = (size_t)ws->g.inst_stride;
  char* dst = (char*)ws->slot[0] + ws->csz * n * inst;
  if (ws->prec == BEC_FP64) {
    HIPCHK(hipMemcpy(dst, psi, n * 16, hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(2 * n);
    for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (float)psi[i];
    HIPCHK(hipMemcpy(dst, tmp.data(), n * 8, hipMemcpyHostToDevice));
  }
  return BEC_OK;
}

int bec_get_psi(bec_ws* ws, int inst, double* psi) {
  if (!ws || !psi) return fail(BEC_EINVAL, "null argument");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  HIPCHK(hipStreamSynchronize(ws->stream));
  const size_t n = (size_t)ws->g.inst_stride;
  const char* src = (const char*)ws->slot[0] + ws->csz * n * inst;
  if (ws->prec == BEC_FP64) {
    HIPCHK(hipMemcpy(psi, src, n * 16, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(2 * n);
    HIPCHK(hipMemcpy(tmp.data(), src, n * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < 2 * n; ++i) psi[i] = tmp[i];
  }
  return BEC_OK;
}

int bec_set_active(bec_ws* ws, const int* insts, int n) {
  if (!ws || (n > 0 && !insts)) return fail(BEC_EINVAL, "null argument");
  if (n < 0 || n > ws->B) return fail(BEC_EINVAL, "bad active count");
  std::vector<int> a(insts, insts + n);
  for (int v : a)
    if (v < 0 || v >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  HIPCHK(hipStreamSynchronize(ws->stream));
  if (n) HIPCHK(hipMemcpy(ws->d_active, a.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  ws->active = a;
  return BEC_OK;
}

static bool any_loss(const bec_ws* ws) {
  for (int inst : ws->active)
    if (ws->hp[inst].has_loss) return true;
  return false;
}

// 1-D grids whose split step is a single x pass: one block per instance holds the whole
// line, so n steps can run in one launch that repeats the step with the line resident in
// LDS (XProgram::nrep).  The step's program is taken from the schedule compiler itself (one
// step recorded, not launched): eligible when it is exactly one x pass of the stages
// ops A / 1-D kinetic / ops B / norm partial (+ its reduction and the closing
// renormalisation in imaginary time), i.e. no lane-per-site operators (F >= 2 spin mixing,
// tensor, Raman), no DDI and a built line length.  Within the launch the op sequence of every
// step is the one the separate launches run, and each renormalisation uses the same norm
// (the repetition loop of bec_xpass_body.h), so the result is bitwise that of the per-step
// schedule.
static bool repeat_1d_program(bec_ws* ws, XProgram* out) {
  // power-of-two lines only: their repeat kernels compile without register spills, and each
  // is tested bit for bit against the per-step launches; a 112-point line (radix 7) gave
  // wrong repeated steps while its per-step kernel was right, and the cause is not known
  const int nx = ws->g.n[0];
  if (ws->rec || ws->g.ndim != 1 || ws->D > 3 || ws->ddi || (nx & (nx - 1)) != 0 || nx < 8 ||
      nx > 512)
    return false;
  const int itp = ws->itp;
  if (itp && ws->normalize_every != 1) return false;
  std::vector<Item> items = {{IT_V, 0.5}, {IT_K, 1.0}, {IT_V, 0.5}};
  if (!itp && any_loss(ws)) items.push_back({IT_LOSS, 1.0});
  if (itp) items.push_back({IT_NORM, 0.0});
  std::vector<Rec> rec;
  ws->rec = &rec;
  const int r = run_items(ws, items, itp, 0.0);
  ws->rec = nullptr;
  if (r != BEC_OK || rec.empty() || rec[0].kind != R_X) return false;
  const XProgram& p = rec[0].xp;
  const int allowed = XS_OPS_A | XS_KIN_1D | XS_OPS_B | XS_NORM;
  if ((p.stages & ~allowed) != 0 || !(p.stages & XS_KIN_1D)) return false;
  if (rec[0].src != ws->slot[0] || rec[0].dst != ws->slot[0]) return false;
  if (itp) {
    if (rec.size() != 3 || rec[1].kind != R_REDUCE || rec[2].kind != R_SCALE ||
        !(p.stages & XS_NORM))
      return false;
  } else if (rec.size() != 1 || (p.stages & XS_NORM)) {
    return false;
  }
  for (int i = 0; i < p.nA; ++i)
    if (p.A[i].kind == OP_SCALE) return false;
  for (int i = 0; i < p.nB; ++i)
    if (p.B[i].kind == OP_SCALE) return false;
  *out = p;
  return true;
}

int bec_split_step(bec_ws* ws, int n_steps) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (n_steps < 0) return fail(BEC_EINVAL, "n_steps < 0");
  if (n_steps >= 2 && n_active(ws) > 0) {
    XProgram p;
    if (repeat_1d_program(ws, &p)) {
      p.nrep = n_steps;
      CHK(do_x(ws, p, ws->slot[0], ws->slot[0]));
      if (p.stages & XS_NORM) {
        CHK(do_reduce(ws, xpass_grid(ws->g), 1u << RS_NORM));
        HIPCHK(ws->prec == BEC_FP32
                   ? launch_scale_norm<float>(ws->stream, ws->g, ws->tb, ws->slot[0], n_active(ws))
                   : launch_scale_norm<double>(ws->stream, ws->g, ws->tb, ws->slot[0], n_active(ws)));
      }
      return BEC_OK;
    }
  }
  const bool loss = !ws->itp && any_loss(ws);  // split_step.jl:21-23: real time only
  std::vector<Item> items;
  items.reserve(5 * (size_t)n_steps);
  for (int s = 1; s <= n_steps; ++s) {
    items.push_back({IT_V, 0.5});
    items.push_back({IT_K, 1.0});
    items.push_back({IT_V, 0.5});
    if (loss) items.push_back({IT_LOSS, 1.0});
    if (ws->itp && ws->normalize_every > 0 && (s % ws->normalize_every) == 0)
      items.push_back({IT_NORM, 0.0});
  }
  return run_items(ws, items, ws->itp, 0.0);
}

// Which operators the schedule compiler includes for the active instances (the captured
// launch sequence depends on these and on nothing else in the parameters: every value is
// read from device memory by the kernels).
static int sched_flags(const bec_ws* ws) {
  int f = 0;
  for (int inst : ws->active) {
    if (std::fabs(ws->hp[inst].c1) >= 1e-30) f |= 1;
    if (std::fabs(ws->hp[inst].c2) >= 1e-30) f |= 2;
    if (ws->hp[inst].has_loss) f |= 4;
    if (ws->hp[inst].has_raman) f |= 8;
  }
  return f;
}

int bec_split_step_graph(bec_ws* ws, int n_steps) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (n_steps < 0) return fail(BEC_EINVAL, "n_steps < 0");
  if (n_steps == 0 || n_active(ws) == 0) return BEC_OK;
  if (ws->prof_on) return bec_split_step(ws, n_steps);  // per-launch events are not captured
  CHK(sync_params(ws));  // parameters reach the device before (never inside) the capture
  const int flags = sched_flags(ws);
  bec_ws::Graph* hit = nullptr;
  for (auto& g : ws->graphs)
    if (g.n == n_steps && g.flags == flags && g.active == ws->active) hit = &g;
  if (!hit) {
    constexpr size_t kMaxGraphs = 32;
    if (ws->graphs.size() >= kMaxGraphs) {
      (void)hipGraphExecDestroy(ws->graphs.front().exec);
      ws->graphs.erase(ws->graphs.begin());
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal));
    const int rc = bec_split_step(ws, n_steps);
    const hipError_t ec = hipStreamEndCapture(ws->stream, &graph);
    if (rc != BEC_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIPCHK(ec);
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(ei);
    ws->graphs.push_back({n_steps, ws->active, flags, exec});
    hit = &ws->graphs.back();
  }
  HIPCHK(hipGraphLaunch(hit->exec, ws->stream));
  return BEC_OK;
}

int bec_leapfrog_segment(bec_ws* ws, int n_steps) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (n_steps < 1) return fail(BEC_EINVAL, "n_steps < 1");
  const bool loss = any_loss(ws);
  std::vector<Item> items;
  items.push_back({IT_V, 0.5});
  for (int s = 1; s <= n_steps; ++s) {
    items.push_back({IT_K, 1.0});
    items.push_back({IT_V, s == n_steps ? 0.5 : 1.0});
    if (loss) items.push_back({IT_LOSS, 1.0});  // simulation.jl:74-76
  }
  return run_items(ws, items, 0, 0.0);
}

int bec_normalize_constrained(bec_ws* ws, const double* target_mz) {
  if (!ws || !target_mz) return fail(BEC_EINVAL, "null argument");
  if (n_active(ws) == 0) return BEC_OK;
  constexpr int kBlk = 256;
  if (!ws->d_cpart) {
    if (hipMalloc(&ws->d_cpart, sizeof(double) * ws->B * ws->D * kBlk) != hipSuccess ||
        hipMalloc(&ws->d_ctarget, sizeof(double) * ws->B) != hipSuccess ||
        hipMalloc(&ws->d_cfac, sizeof(double) * ws->B * kMaxD) != hipSuccess)
      return fail(BEC_ENOMEM, "constrained normalisation buffers");
  }
  CHK(sync_params(ws));
  HIPCHK(hipMemcpyAsync(ws->d_ctarget, target_mz, sizeof(double) * ws->B, hipMemcpyHostToDevice,
                        ws->stream));
  hipError_t e = ws->prec == BEC_FP32
                     ? launch_constrained_norm<float>(ws->stream, ws->g, ws->tb, ws->slot[0],
                                                      n_active(ws), ws->d_cpart, kBlk,
                                                      ws->d_ctarget, ws->d_cfac)
                     : launch_constrained_norm<double>(ws->stream, ws->g, ws->tb, ws->slot[0],
                                                       n_active(ws), ws->d_cpart, kBlk,
                                                       ws->d_ctarget, ws->d_cfac);
  HIPCHK(e);
  // the host array must stay valid until the copy has been consumed
  HIPCHK(hipStreamSynchronize(ws->stream));
  return BEC_OK;
}

int bec_loss_step(bec_ws* ws, double dt) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (!any_loss(ws)) return BEC_OK;
  return run_items(ws, {{IT_LOSS, 1.0}}, 0, dt);
}

int bec_set_loss(bec_ws* ws, int inst, const double* rates, int n) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  InstParams& q = ws->hp[inst];
  if (!rates) {
    q.has_loss = 0;
    for (double& r : q.loss) r = 0.0;
  } else {
    if (n != ws->D) return fail(BEC_EINVAL, "loss rates: expected 2F+1 values");
    for (int c = 0; c < ws->D; ++c) {
      if (!(rates[c] >= 0.0)) return fail(BEC_EINVAL, "loss rates must be >= 0");
      q.loss[c] = rates[c];
    }
    q.has_loss = 1;
  }
  ws->params_dirty = true;
  return BEC_OK;
}

int bec_set_raman(bec_ws* ws, int inst, const double* raman) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  InstParams& q = ws->hp[inst];
  if (!raman) {
    q.has_raman = 0;
    for (double& r : q.raman) r = 0.0;
  } else {
    for (int i = 0; i < 2 + ws->g.ndim; ++i)
      if (!std::isfinite(raman[i])) return fail(BEC_EINVAL, "Raman parameters must be finite");
    for (int i = 0; i < 5; ++i) q.raman[i] = i < 2 + ws->g.ndim ? raman[i] : 0.0;
    q.has_raman = 1;
  }
  ws->params_dirty = true;
  return BEC_OK;
}

int bec_half_potential(bec_ws* ws, double dt_half, int imaginary_time) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  return run_items(ws, {{IT_V, 1.0}}, imaginary_time, dt_half);
}

int bec_kinetic(bec_ws* ws, double dt, int imaginary_time) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  return run_items(ws, {{IT_K, 1.0}}, imaginary_time, dt);
}

int bec_strang_core(bec_ws* ws, double dt) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  return run_items(ws, {{IT_V, 0.5}, {IT_K, 1.0}, {IT_V, 0.5}}, 0, dt);
}

int bec_yoshida_core(bec_ws* ws, double dt) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  const double w1 = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0));
  const double w0 = 1.0 - 2.0 * w1;
  const double wm = (w1 + w0) / 2;
  return run_items(ws,
                   {{IT_V, w1 / 2}, {IT_K, w1}, {IT_V, wm}, {IT_K, w0}, {IT_V, wm}, {IT_K, w1},
                    {IT_V, w1 / 2}},
                   0, dt);
}

// The cores from another slot: slot src holds the state, slot 0 receives the result, src
// is left unchanged.  Only with every instance active (slot 0's inactive instances would
// keep stale data) and outside recording; otherwise slot src is copied first.
static int ensure_slot(bec_ws* ws, int s);
static int core_from(bec_ws* ws, int src, int (*core)(bec_ws*, double), double dt) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (src < 0 || src > 3) return fail(BEC_EINVAL, "slot out of range");
  if (src == 0) return core(ws, dt);
  CHK(ensure_slot(ws, src));
  ws->src_once = src;
  if (ws->rec || n_active(ws) != ws->B) CHK(flush_src(ws));
  const int r = core(ws, dt);
  if (r != BEC_OK) {
    ws->src_once = -1;
    return r;
  }
  return flush_src(ws);  // a core that launched nothing (no active instance)
}

int bec_strang_core_from(bec_ws* ws, double dt, int src_slot) {
  return core_from(ws, src_slot, bec_strang_core, dt);
}

int bec_yoshida_core_from(bec_ws* ws, double dt, int src_slot) {
  return core_from(ws, src_slot, bec_yoshida_core, dt);
}

int bec_normalize(bec_ws* ws) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  return run_items(ws, {{IT_NORM, 0.0}}, ws->itp, 0.0);
}

int bec_synchronize(bec_ws* ws) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  HIPCHK(hipStreamSynchronize(ws->stream));
  return BEC_OK;
}

int bec_set_xsite(bec_ws* ws, int on) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (ws->xsite != (on != 0)) drop_graphs(ws);  // captured launches follow the old schedule
  ws->xsite = on != 0;
  return ws->xsite && xsite_ok(ws->g) ? 1 : 0;
}

// fft_utils.jl:4-19,25-44: d_d f = Re IFFT(i k_d FFT(f)) of a real field.  A private FP64
// spin-1 workspace of the grid holds f in component d for every d < ndim, so one schedule
// (forward transforms, i k_d / N on component d, inverse transforms) yields the gradient.
static int fft_gradient_impl(const bec_config* grid, const double* f, int only, double* out) {
  if (!grid || !f || !out) return fail(BEC_EINVAL, "null argument");
  bec_config c = *grid;
  c.F = 1;
  c.n_instances = 1;
  c.precision = BEC_FP64;
  c.imaginary_time = 0;
  c.normalize_every = 0;
  c.enable_ddi = 0;
  c.ddi_secular = 0;
  c.ddi_padding = 0;
  c.tensor = 0;
  bec_ws* ws = nullptr;
  CHK(bec_ws_create(&c, &ws));
  struct Destroy {
    bec_ws* w;
    ~Destroy() { bec_ws_destroy(w); }
  } destroy{ws};
  const int nd = ws->g.ndim;
  if (only >= nd) return fail(BEC_EINVAL, "dim must be in [0, ndim)");
  const size_t N = (size_t)ws->g.nsites;
  std::vector<double> h(2 * 3 * N, 0.0);
  for (int d = 0; d < nd; ++d)
    for (size_t i = 0; i < N; ++i) h[2 * (d * N + i)] = f[i];
  CHK(bec_set_psi(ws, 0, h.data()));
  CHK(sync_params(ws));
  XProgram p{};
  if (nd == 1) {
    p.stages = XS_KIN_1D;
    p.deriv = 1;
    CHK(do_x(ws, p, ws->slot[0], ws->slot[0]));
  } else {
    p.stages = XS_FWD_FFT;
    CHK(do_x(ws, p, ws->slot[0], ws->slot[0]));
    if (nd == 3) {
      CHK(do_strided(ws, 1, SM_FWD, ws->slot[0], 0.0, 0.0, 0, 0));
      CHK(do_strided(ws, 2, SM_DERIV, ws->slot[0], 0.0, 0.0, 0, 0));
      CHK(do_strided(ws, 1, SM_INV, ws->slot[0], 0.0, 0.0, 0, 0));
    } else {
      CHK(do_strided(ws, 1, SM_DERIV, ws->slot[0], 0.0, 0.0, 0, 0));
    }
    XProgram q{};
    q.stages = XS_INV_FFT;
    CHK(do_x(ws, q, ws->slot[0], ws->slot[0]));
  }
  CHK(bec_get_psi(ws, 0, h.data()));
  for (int d = 0; d < nd; ++d) {
    if (only >= 0 && d != only) continue;
    double* o = out + (only >= 0 ? 0 : d * N);
    for (size_t i = 0; i < N; ++i) o[i] = h[2 * (d * N + i)];
  }
  return BEC_OK;
}

int bec_fft_derivative(const bec_config* grid, const double* f, int dim, double* out) {
  if (dim < 0) return fail(BEC_EINVAL, "dim must be in [0, ndim)");
  return fft_gradient_impl(grid, f, dim, out);
}

int bec_fft_gradient(const bec_config* grid, const double* f, double* out) {
  return fft_gradient_impl(grid, f, -1, out);
}

// ---------------------------------------------------------------------------
// bec_fft_gradient_noalloc: bec_fft_gradient for callers that cannot allocate device memory.
// Every device buffer of the private spin-1 workspace (psi, the staging of f and of the
// result, axis tables, twiddles, parameter block, active list, reduction partials) lives in
// one file-scope array per GPU.  Host code never passes the array's name to the device: the
// address comes from hipGetSymbolAddress, once per device, on that device's first call.
// ---------------------------------------------------------------------------
constexpr long long kNaMaxSites = 1LL << 18;  // 64^3, 512^2, ...
constexpr size_t kNaAlign = 256;
constexpr size_t na_up(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr int kNaMaxSide = 512;
constexpr int kNaMaxBlocks = 1024;
constexpr size_t kNaPsi = 0;
constexpr size_t kNaF = kNaPsi + na_up(3 * kNaMaxSites * 16);
constexpr size_t kNaOut = kNaF + na_up(kNaMaxSites * 8);
constexpr size_t kNaK = kNaOut + na_up(3 * kNaMaxSites * 8);
constexpr size_t kNaX = kNaK + na_up(3 * kNaMaxSide * 8);
constexpr size_t kNaAx = kNaX + na_up(3 * kNaMaxSide * 8);
constexpr size_t kNaParams = kNaAx + na_up(3 * kNaMaxSide * 8);
constexpr size_t kNaActive = kNaParams + na_up(sizeof(InstParams));
constexpr size_t kNaRed = kNaActive + na_up(sizeof(int));
constexpr size_t kNaPart = kNaRed + na_up(RS_COUNT * 8);
constexpr size_t kNaTw = kNaPart + na_up((size_t)RS_COUNT * kNaMaxBlocks * 8);
constexpr size_t na_tw_bytes() {
  size_t b = 0;
  for (int i = 0; i < kNumFftSizes; ++i) b += na_up(2 * (size_t)kFftSizes[i] * 8);
  return b;
}
constexpr size_t kNaBytes = kNaTw + na_tw_bytes();  // about 21.1 MB

static __device__ __align__(256) unsigned char noalloc_scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static char* g_na_ptr[kNaMaxDevices] = {};  // device address of noalloc_scratch, per device

// Device address of this device's array, with the twiddle tables (which never change)
// written on the first call.  Caller holds g_na_mu.
static int noalloc_setup(int dev, char** base) {
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    char* b = static_cast<char*>(p);
    size_t off = kNaTw;
    double buf[2 * kNaMaxSide];
    for (int l = 0; l < kNumFftSizes; ++l) {
      const int N = kFftSizes[l];
      for (int m = 0; m < N; ++m) {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * m / N;
        buf[2 * m] = (double)cosl(ang);
        buf[2 * m + 1] = (double)sinl(ang);
      }
      HIPCHK(hipMemcpy(b + off, buf, 2 * (size_t)N * 8, hipMemcpyHostToDevice));
      off += na_up(2 * (size_t)N * 8);
    }
    g_na_ptr[dev] = b;
  }
  *base = g_na_ptr[dev];
  return BEC_OK;
}

int bec_fft_gradient_noalloc(const bec_config* grid, const double* f, double* out) {
  if (!grid || !f || !out) return fail(BEC_EINVAL, "null argument");
  const bec_config& c = *grid;
  if (c.ndim < 1 || c.ndim > 3) return fail(BEC_EINVAL, "ndim must be 1, 2 or 3");
  long long N = 1;
  for (int d = 0; d < c.ndim; ++d) {
    if (c.n[d] <= 0 || (c.n[d] % 2) != 0)
      return fail(BEC_EINVAL, "n_points must be positive even integers");
    if (!(c.L[d] > 0)) return fail(BEC_EINVAL, "box_size must be positive");
    if (!side_supported(c.n[d]))
      return fail(BEC_EUNSUP, "grid side without a built FFT line on this build");
    N *= c.n[d];
  }
  if (N > kNaMaxSites)
    return fail(BEC_EINVAL, "bec_fft_gradient_noalloc: grid of " + std::to_string(N) +
                                " sites exceeds the fixed buffer of " +
                                std::to_string(kNaMaxSites) +
                                " sites (use bec_fft_gradient for larger grids)");
  if (!c.x_tab || !c.k_tab) return fail(BEC_EINVAL, "missing tables");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) return fail(BEC_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_na_mu);
  char* b = nullptr;
  CHK(noalloc_setup(dev, &b));
  const hipStream_t st = nullptr;
  Geom g;
  fill_geom(&c, 3, g);
  const int nd = g.ndim;
  // axis tables [3][nmax]: only the first n[d] entries of row d are read
  int off = 0;
  for (int d = 0; d < nd; ++d) {
    HIPCHK(hipMemcpy(b + kNaK + (size_t)d * g.nmax * 8, c.k_tab + off, (size_t)c.n[d] * 8,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + kNaX + (size_t)d * g.nmax * 8, c.x_tab + off, (size_t)c.n[d] * 8,
                     hipMemcpyHostToDevice));
    off += c.n[d];
  }
  InstParams ip{};
  const int active0 = 0;
  HIPCHK(hipMemcpy(b + kNaParams, &ip, sizeof(ip), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b + kNaActive, &active0, sizeof(int), hipMemcpyHostToDevice));
  DevTables tb{};
  tb.ktab = reinterpret_cast<const double*>(b + kNaK);
  tb.xtab = reinterpret_cast<const double*>(b + kNaX);
  tb.sktab = tb.ktab;
  tb.axtab = reinterpret_cast<const double*>(b + kNaAx);
  tb.Varr = nullptr;
  tb.params = reinterpret_cast<const InstParams*>(b + kNaParams);
  tb.active = reinterpret_cast<const int*>(b + kNaActive);
  size_t toff = kNaTw;
  for (int l = 0; l < kNumFftSizes; ++l) {
    tb.tw[l] = b + toff;
    toff += na_up(2 * (size_t)kFftSizes[l] * 8);
  }
  tb.red = reinterpret_cast<const double*>(b + kNaRed);
  tb.partials = reinterpret_cast<double*>(b + kNaPart);
  tb.max_blocks = kNaMaxBlocks;
  tb.maxB = 1;
  tb.work = nullptr;
  void* psi = b + kNaPsi;
  double* fdev = reinterpret_cast<double*>(b + kNaF);
  double* odev = reinterpret_cast<double*>(b + kNaOut);
  // every component of psi is written by the packing kernel (the array keeps the previous
  // call's contents)
  HIPCHK(hipMemcpy(fdev, f, (size_t)N * 8, hipMemcpyHostToDevice));
  HIPCHK(launch_grad_pack(st, fdev, psi, N, nd));
  // the schedule of bec_fft_gradient: forward transforms, i k_d / N on component d, inverse
  XProgram p{};
  if (nd == 1) {
    p.stages = XS_KIN_1D;
    p.deriv = 1;
    HIPCHK(launch_xpass<double>(st, g, tb, p, psi, psi, nullptr, 1, nullptr));
  } else {
    p.stages = XS_FWD_FFT;
    HIPCHK(launch_xpass<double>(st, g, tb, p, psi, psi, nullptr, 1, nullptr));
    if (nd == 3) {
      HIPCHK(launch_strided<double>(st, g, tb, 1, SM_FWD, psi, 1, 0.0, 0.0, 0, 0, nullptr));
      HIPCHK(launch_strided<double>(st, g, tb, 2, SM_DERIV, psi, 1, 0.0, 0.0, 0, 0, nullptr));
      HIPCHK(launch_strided<double>(st, g, tb, 1, SM_INV, psi, 1, 0.0, 0.0, 0, 0, nullptr));
    } else {
      HIPCHK(launch_strided<double>(st, g, tb, 1, SM_DERIV, psi, 1, 0.0, 0.0, 0, 0, nullptr));
    }
    XProgram q{};
    q.stages = XS_INV_FFT;
    HIPCHK(launch_xpass<double>(st, g, tb, q, psi, psi, nullptr, 1, nullptr));
  }
  HIPCHK(launch_grad_unpack(st, psi, odev, N, nd));
  HIPCHK(hipMemcpy(out, odev, (size_t)nd * N * 8, hipMemcpyDeviceToHost));
  return BEC_OK;
}

// Tile autotuner.  The x pass's rows per block set its LDS tile, its thread count and the
// number of blocks in flight; the best value depends on (precision, n_x, D, number of rows).
// Candidates (powers of two that divide the rows, within 512 threads and the partials
// buffer) are timed on the workspace's own data: an inverse FFT + diagonal step + forward
// FFT pass from psi into the scratch buffer, one warm-up and three timed launches each.
// A candidate replaces the size heuristic only when it is 3 % faster.  Decisions are kept
// per process and shape, so every workspace of a shape uses the same tiles (the block
// partition is also the partition of the block-partial reductions).
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, int> g_tune;

int bec_autotune(bec_ws* ws, int* rows_out) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  if (xpass_generic(ws->g)) {  // the runtime-length x pass keeps its fixed partition
    if (rows_out) *rows_out = xpass_rows_g(ws->g);
    return BEC_OK;
  }
  const auto key = std::make_tuple((int)ws->prec, ws->g.n[0], ws->g.D, ws->g.nrows);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find(key);
  if (it != g_tune.end()) {
    if (ws->g.xrows != it->second) drop_graphs(ws);
    ws->g.xrows = it->second;
    if (rows_out) *rows_out = xpass_rows_g(ws->g);
    return BEC_OK;
  }
  Geom g = ws->g;
  g.xrows = 0;
  const int r0 = xpass_rows_g(g);
  // Shapes that run the fused x/site kernel by default keep its default tile: the timing
  // below is of the x pass, whose best tile is not the fused kernel's (128^3 spin 6 FP64,
  // one box: 624-627 us per fused launch at one row per block, 662-673 us at the two rows
  // that the x-pass timing picked on some runs), and a tile above kXsiteMaxThreads sites
  // would switch the shape to the separate passes for the rest of the process.
  const bool fusable = xsite_ok(g);
  std::vector<int> cand;
  for (int r = 1; r <= g.nrows && !fusable; r *= 2) {
    g.xrows = r;
    if (xpass_rows_ok(g, r) && g.nrows / r <= ws->max_blocks && r * 4 >= r0) cand.push_back(r);
  }
  int best = 0;
  if (cand.size() > 1 && n_active(ws) > 0) {
    CHK(sync_params(ws));
    XProgram p{};
    p.stages = XS_INV_FFT | XS_FWD_FFT | XS_OPS_A;
    p.itp = ws->itp;
    p.nA = 1;
    p.A[0].kind = OP_DIAG;
    p.A[0].frac = 0.5;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double t_best = 0.0, t_def = 0.0;
    int rc = BEC_OK;
    for (int r : cand) {
      ws->g.xrows = r;
      hipError_t e = hipSuccess;
      for (int k = 0; k < 4 && e == hipSuccess; ++k) {
        if (k == 1) e = hipEventRecord(e0, ws->stream);
        if (e == hipSuccess)
          e = ws->prec == BEC_FP32
                  ? launch_xpass<float>(ws->stream, ws->g, ws->tb, p, ws->slot[0], ws->scratch,
                                        ws->fhat, n_active(ws), ws->phi)
                  : launch_xpass<double>(ws->stream, ws->g, ws->tb, p, ws->slot[0], ws->scratch,
                                         ws->fhat, n_active(ws), ws->phi);
      }
      if (e == hipSuccess) e = hipEventRecord(e1, ws->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(BEC_EDEVICE, std::string("autotune: ") + hipGetErrorString(e));
        break;
      }
      if (r == r0) t_def = ms;
      if (best == 0 || ms < t_best) {
        best = r;
        t_best = ms;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ws->g.xrows = 0;
    if (rc != BEC_OK) return rc;
    if (t_def > 0.0 && t_best > 0.97 * t_def) best = 0;
  }
  g_tune[key] = best;
  if (ws->g.xrows != best) drop_graphs(ws);
  ws->g.xrows = best;
  if (rows_out) *rows_out = xpass_rows_g(ws->g);
  return BEC_OK;
}

int bec_line_plan(int n) {
  if (n < 2) return fail(BEC_EINVAL, "line length below 2");
  if (fft_size_id(n) >= 0) return 0;
  return gen_blue_m(n) > 0 ? 2 : 1;
}

int bec_query(bec_ws* ws) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  const hipError_t e = hipStreamQuery(ws->stream);
  if (e == hipErrorNotReady) return 1;
  HIPCHK(e);
  return BEC_OK;
}

static int fetch_red(bec_ws* ws) {
  HIPCHK(hipMemcpyAsync(ws->h_red, ws->d_red, (size_t)ws->B * RS_COUNT * 8, hipMemcpyDeviceToHost,
                        ws->stream));
  HIPCHK(hipStreamSynchronize(ws->stream));
  return BEC_OK;
}

// Energy partial sums of psi (FP64 arithmetic on FP64 data), reduced into d_red.
// The DDI energy always uses the unpadded convolution (energy.jl:111-121), also in a
// workspace whose steps use the zero-padded one: same grid, spectra of the unpadded shape.

static int energy_partials_fp64(bec_ws* ws, const DevTables& tb0, void* psi, void* scratch,
                                void* fhat, void* phi) {
  const Geom g = unpadded_geom(ws->g);
  DevTables tb = tb0;
  tb.sktab = tb.ktab;
  const int na = n_active(ws);
  XProgram p{};
  p.itp = ws->itp;
  auto X = [&](const XProgram& q, const void* src, void* dst) -> int {
    HIPCHK(launch_xpass<double>(ws->stream, g, tb, q, src, dst, fhat, na, phi));
    return BEC_OK;
  };
  auto S = [&](int axis, int mode, void* data, int ddi, int* nb) -> int {
    HIPCHK(launch_strided<double>(ws->stream, g, tb, axis, mode, data, na, 0, 0, 0, ddi, nb));
    return BEC_OK;
  };
  if (ws->ddi) {
    p.stages = XS_WRITE_FHAT | XS_NO_STORE;
    CHK(X(p, psi, psi));
    if (g.ndim == 3) {
      CHK(S(1, SM_FWD, fhat, 1, nullptr));
      CHK(S(2, SM_DDI, fhat, 1, nullptr));
      CHK(S(1, SM_INV, fhat, 1, nullptr));
    } else if (g.ndim == 1) {
      HIPCHK(launch_ddi_line<double>(ws->stream, g, tb, fhat, na));
    } else {
      CHK(S(1, SM_DDI, fhat, 1, nullptr));
    }
    p.stages = XS_LOAD_PHI | XS_STORE_PHI | XS_NO_LOAD | XS_NO_STORE;
    CHK(X(p, psi, psi));
  }
  int nbe = 0;
  HIPCHK(launch_energy<double>(ws->stream, g, tb, psi, ws->ddi ? phi : nullptr, ws->tensor, na,
                               &nbe));
  const uint32_t emask = (1u << RS_NORM) | (1u << RS_TRAP) | (1u << RS_ZEE) | (1u << RS_C0) |
                         (1u << RS_C1) | (1u << RS_LHY) | (1u << RS_DDI) | (1u << RS_MZ) |
                         (1u << RS_TENSOR) | (1u << RS_RAMAN);
  HIPCHK(launch_reduce(ws->stream, tb, ws->d_red, na, nbe, emask, 0));
  // kinetic energy: forward FFT of every component, sum k^2 |psi_hat|^2 (energy.jl:56-65)
  if (g.ndim == 1) {
    p.stages = XS_FWD_FFT | XS_KE_1D | XS_NO_STORE;
    CHK(X(p, psi, psi));
    HIPCHK(launch_reduce(ws->stream, tb, ws->d_red, na, xpass_grid(g), 1u << RS_KIN, 0));
  } else {
    p.stages = XS_FWD_FFT;
    CHK(X(p, psi, scratch));
    int nb = 0;
    if (g.ndim == 3) CHK(S(1, SM_FWD, scratch, 0, nullptr));
    CHK(S(g.ndim - 1, SM_KE, scratch, 0, &nb));
    HIPCHK(launch_reduce(ws->stream, tb, ws->d_red, na, nb, 1u << RS_KIN, 0));
  }
  return BEC_OK;
}

// Energy partials of every active instance, reduced into d_red (no synchronisation).
static int energy_enqueue(bec_ws* ws) {
  CHK(sync_params(ws));
  const Geom& g = ws->g;
  const int pe = prof_begin(ws, 9);
  if (ws->prec == BEC_FP64) {
    CHK(energy_partials_fp64(ws, ws->tb, ws->slot[0], ws->scratch, ws->fhat, ws->phi));
  } else {
    // FP32-bulk: the state is widened to FP64 and every energy term (FFT, DDI
    // convolution, reductions) is evaluated in FP64
    const size_t n = (size_t)g.inst_stride * ws->B;
    if (!ws->psi64) {
      if (hipMalloc(&ws->psi64, n * 16) != hipSuccess) return fail(BEC_ENOMEM, "psi64");
      if (hipMalloc(&ws->scratch64, n * 16) != hipSuccess) return fail(BEC_ENOMEM, "scratch64");
      if (ws->ddi) {
        const size_t fb = 16 * (size_t)unpadded_geom(g).fhat_inst_stride * ws->B;
        if (hipMalloc(&ws->fhat64, fb) != hipSuccess) return fail(BEC_ENOMEM, "fhat64");
        HIPCHK(hipMemset(ws->fhat64, 0, fb));
        if (hipMalloc(&ws->phi64, 24 * (size_t)g.nsites * ws->B) != hipSuccess)
          return fail(BEC_ENOMEM, "phi64");
      }
    }
    HIPCHK(launch_widen(ws->stream, ws->slot[0], ws->psi64, (long long)(2 * n)));
    DevTables tb = ws->tb;
    CHK(energy_partials_fp64(ws, tb, ws->psi64, ws->scratch64, ws->fhat64, ws->phi64));
  }
  prof_end(ws, pe);
  return BEC_OK;
}

// Energy terms (BEC_E_*) of the active instances from reduced sums red[B][RS_COUNT].
static void energy_terms_from(const bec_ws* ws, const double* red, double* out) {
  const Geom& g = ws->g;
  const double dV = g.dV;
  for (int inst : ws->active) {
    const double* r = red + (size_t)inst * RS_COUNT;
    const InstParams& q = ws->hp[inst];
    double* o = out + (size_t)inst * BEC_E_COUNT;
    o[BEC_E_KIN] = 0.5 * (r[RS_KIN] * dV / (double)g.nsites);
    o[BEC_E_TRAP] = r[RS_TRAP] * dV;
    o[BEC_E_ZEEMAN] = r[RS_ZEE] * dV;
    o[BEC_E_C0] = 0.5 * q.c0 * r[RS_C0] * dV;
    o[BEC_E_C1] = 0.5 * q.c1 * r[RS_C1] * dV;
    o[BEC_E_DDI] = ws->ddi ? 0.5 * r[RS_DDI] * dV : 0.0;
    o[BEC_E_LHY] = q.c_lhy != 0.0 ? (2.0 / 5.0) * q.c_lhy * r[RS_LHY] * dV : 0.0;
    o[BEC_E_TENSOR] = r[RS_TENSOR] * dV;  // tensor channels, or the nematic term without them
    o[BEC_E_NORM] = r[RS_NORM] * dV;
    o[BEC_E_MZ] = r[RS_MZ] * dV;
    o[BEC_E_RAMAN] = q.has_raman ? r[RS_RAMAN] * dV : 0.0;
    if (ws->prec == BEC_FP32 && ws->itp && ws->normalize_every > 0 && o[BEC_E_NORM] > 0.0) {
      // FP32-bulk imaginary time: the FP64 renormalisation factor is applied with one
      // rounding per element, and for a smooth state those roundings share a sign, so the
      // stored state's norm is off by up to ~1e-7.  Its energies are those of the state
      // the renormalisation means, psi / sqrt(norm): quadratic terms scale by 1/N,
      // quartic ones by 1/N^2, the LHY term by N^-5/2.
      const double in = 1.0 / o[BEC_E_NORM];
      for (int k : {BEC_E_KIN, BEC_E_TRAP, BEC_E_ZEEMAN, BEC_E_RAMAN, BEC_E_MZ}) o[k] *= in;
      for (int k : {BEC_E_C0, BEC_E_C1, BEC_E_DDI, BEC_E_TENSOR}) o[k] *= in * in;
      o[BEC_E_LHY] *= in * in * std::sqrt(in);
    }
    o[BEC_E_TOTAL] = o[BEC_E_KIN] + o[BEC_E_TRAP] + o[BEC_E_ZEEMAN] + o[BEC_E_C0] + o[BEC_E_C1] +
                     o[BEC_E_DDI] + o[BEC_E_LHY] + o[BEC_E_TENSOR] + o[BEC_E_RAMAN];
  }
}

int bec_energy(bec_ws* ws, double* out) {
  if (!ws || !out) return fail(BEC_EINVAL, "null argument");
  if (n_active(ws) == 0) return BEC_OK;
  CHK(energy_enqueue(ws));
  CHK(fetch_red(ws));
  energy_terms_from(ws, ws->h_red, out);
  return BEC_OK;
}

static int ensure_slot(bec_ws* ws, int s) {
  if (s < 0 || s > 3) return fail(BEC_EINVAL, "slot out of range");
  if (!ws->slot[s]) {
    const size_t b = ws->csz * (size_t)ws->g.inst_stride * ws->B;
    if (hipMalloc(&ws->slot[s], b) != hipSuccess) return fail(BEC_ENOMEM, "slot alloc");
    HIPCHK(hipMemsetAsync(ws->slot[s], 0, b, ws->stream));
  }
  return BEC_OK;
}

int bec_copy_slot(bec_ws* ws, int dst, int src) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  CHK(ensure_slot(ws, dst));
  CHK(ensure_slot(ws, src));
  const size_t b = ws->csz * (size_t)ws->g.inst_stride;
  if (dst == src || n_active(ws) == 0) return BEC_OK;
  if (b % 16 == 0) {
    HIPCHK(launch_copy_slot(ws->stream, ws->tb, ws->slot[src], ws->slot[dst], (long long)b,
                            n_active(ws)));
    return BEC_OK;
  }
  for (int inst : ws->active)
    HIPCHK(hipMemcpyAsync((char*)ws->slot[dst] + b * inst, (char*)ws->slot[src] + b * inst, b,
                          hipMemcpyDeviceToDevice, ws->stream));
  return BEC_OK;
}

int bec_swap_slots(bec_ws* ws, int a, int b) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  CHK(ensure_slot(ws, a));
  CHK(ensure_slot(ws, b));
  if (a == b) return BEC_OK;
  std::swap(ws->slot[a], ws->slot[b]);
  drop_graphs(ws);  // captured launches hold the slot addresses
  return BEC_OK;
}

int bec_diff(bec_ws* ws, int a, int b, int mode, double* out) {
  if (!ws || !out) return fail(BEC_EINVAL, "null argument");
  CHK(ensure_slot(ws, a));
  CHK(ensure_slot(ws, b));
  if (n_active(ws) == 0) return BEC_OK;
  const int nb = 512;
  hipError_t e = ws->prec == BEC_FP32
                     ? launch_diff<float>(ws->stream, ws->g, ws->tb, ws->slot[a], ws->slot[b],
                                          n_active(ws), mode, nb)
                     : launch_diff<double>(ws->stream, ws->g, ws->tb, ws->slot[a], ws->slot[b],
                                           n_active(ws), mode, nb);
  HIPCHK(e);
  const uint32_t m = (1u << RS_NORM) | (1u << RS_KIN);
  CHK(do_reduce(ws, nb, m, mode ? m : 0u));
  CHK(fetch_red(ws));
  for (int inst : ws->active) {
    out[2 * inst] = ws->h_red[(size_t)inst * RS_COUNT + RS_NORM];
    out[2 * inst + 1] = ws->h_red[(size_t)inst * RS_COUNT + RS_KIN];
  }
  return BEC_OK;
}

int bec_spin_density(bec_ws* ws, int inst, double* out) {
  if (!ws || !out) return fail(BEC_EINVAL, "null argument");
  if (inst < 0 || inst >= ws->B) return fail(BEC_EINVAL, "instance out of range");
  const long long N = ws->g.nsites;
  if (!ws->d_sdens && hipMalloc(&ws->d_sdens, 3 * (size_t)N * 8) != hipSuccess)
    return fail(BEC_ENOMEM, "spin density buffer");
  const void* psi = (const char*)ws->slot[0] + ws->csz * (size_t)ws->g.inst_stride * inst;
  HIPCHK(ws->prec == BEC_FP32 ? launch_spin_density<float>(ws->stream, psi, ws->d_sdens, N, ws->D)
                              : launch_spin_density<double>(ws->stream, psi, ws->d_sdens, N, ws->D));
  HIPCHK(hipMemcpyAsync(out, ws->d_sdens, 3 * (size_t)N * 8, hipMemcpyDeviceToHost, ws->stream));
  HIPCHK(hipStreamSynchronize(ws->stream));
  return BEC_OK;
}

int bec_itp_chunk(bec_ws* ws, int n_steps, double* out) {
  if (!ws || !out) return fail(BEC_EINVAL, "null argument");
  if (!ws->itp) return fail(BEC_EINVAL, "bec_itp_chunk needs an imaginary-time workspace");
  if (n_steps < 0) return fail(BEC_EINVAL, "n_steps < 0");
  if (n_active(ws) == 0) return BEC_OK;
  CHK(ensure_slot(ws, 1));
  const size_t redb = (size_t)ws->B * RS_COUNT * 8;
  if (!ws->h_red2 && hipHostMalloc(&ws->h_red2, redb) != hipSuccess)
    return fail(BEC_ENOMEM, "pinned reduction buffer");
  CHK(bec_split_step(ws, n_steps));
  // max |Phi| of the chunk's last convolution, before the energy's own convolution
  // rewrites the per-block partials (the energy does not reduce that slot)
  if (ws->ddi) CHK(do_reduce(ws, xpass_grid(ws->g), 1u << RS_PHIMAX, 1u << RS_PHIMAX));
  CHK(energy_enqueue(ws));
  HIPCHK(hipMemcpyAsync(ws->h_red, ws->d_red, redb, hipMemcpyDeviceToHost, ws->stream));
  // max |psi - psi_prev| and max |psi| (ground_state.jl:148-149), then psi_prev <- psi
  const int nb = 512;
  HIPCHK(ws->prec == BEC_FP32
             ? launch_diff<float>(ws->stream, ws->g, ws->tb, ws->slot[0], ws->slot[1],
                                  n_active(ws), 1, nb)
             : launch_diff<double>(ws->stream, ws->g, ws->tb, ws->slot[0], ws->slot[1],
                                   n_active(ws), 1, nb));
  const uint32_t m = (1u << RS_NORM) | (1u << RS_KIN);
  CHK(do_reduce(ws, nb, m, m));
  HIPCHK(hipMemcpyAsync(ws->h_red2, ws->d_red, redb, hipMemcpyDeviceToHost, ws->stream));
  CHK(bec_copy_slot(ws, 1, 0));
  HIPCHK(hipStreamSynchronize(ws->stream));
  std::vector<double> e((size_t)ws->B * BEC_E_COUNT, 0.0);
  energy_terms_from(ws, ws->h_red, e.data());
  for (int inst : ws->active) {
    const double* r2 = ws->h_red2 + (size_t)inst * RS_COUNT;
    double* o = out + (size_t)inst * BEC_CHUNK_COUNT;
    const double E = e[(size_t)inst * BEC_E_COUNT + BEC_E_TOTAL];
    const double nrm = e[(size_t)inst * BEC_E_COUNT + BEC_E_NORM];
    o[BEC_CHUNK_E] = E;
    o[BEC_CHUNK_DPSI] = r2[RS_KIN] > 0.0 ? r2[RS_NORM] / r2[RS_KIN] : 0.0;
    o[BEC_CHUNK_NORM] = nrm;
    o[BEC_CHUNK_PHIMAX] = ws->ddi ? ws->h_red[(size_t)inst * RS_COUNT + RS_PHIMAX] : 0.0;
    o[BEC_CHUNK_STATUS] = (std::isfinite(E) && std::isfinite(nrm)) ? 0.0 : 1.0;
  }
  return BEC_OK;
}

int bec_itp_guard(bec_ws* ws, double* out) {
  if (!ws || !out) return fail(BEC_EINVAL, "null argument");
  if (n_active(ws) == 0) return BEC_OK;
  const int nb = 512;
  hipError_t e = ws->prec == BEC_FP32
                     ? launch_diff<float>(ws->stream, ws->g, ws->tb, ws->slot[0], ws->slot[0],
                                          n_active(ws), 0, nb)
                     : launch_diff<double>(ws->stream, ws->g, ws->tb, ws->slot[0], ws->slot[0],
                                           n_active(ws), 0, nb);
  HIPCHK(e);
  CHK(do_reduce(ws, nb, 1u << RS_KIN));
  if (ws->ddi) CHK(do_reduce(ws, xpass_grid(ws->g), 1u << RS_PHIMAX, 1u << RS_PHIMAX));
  CHK(fetch_red(ws));
  for (int inst : ws->active) {
    out[2 * inst] = ws->h_red[(size_t)inst * RS_COUNT + RS_KIN];
    out[2 * inst + 1] = ws->ddi ? ws->h_red[(size_t)inst * RS_COUNT + RS_PHIMAX] : 0.0;
  }
  return BEC_OK;
}

int bec_profile_enable(bec_ws* ws, int on) {
  if (!ws) return fail(BEC_EINVAL, "null workspace");
  HIPCHK(hipStreamSynchronize(ws->stream));
  ws->prof_on = on != 0;
  ws->ev_used = 0;
  ws->prof_cls.clear();
  HIPCHK(hipMemset(ws->d_work, 0, kWorkBytes));
  for (double& b : ws->prof_bytes) b = 0.0;
  refresh_tables(ws);
  return BEC_OK;
}

int bec_profile_work(bec_ws* ws, double* out) {
  if (!ws || !out) return fail(BEC_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(ws->stream));
  std::vector<unsigned long long> h((size_t)kWorkSlots * kWorkStride);
  HIPCHK(hipMemcpy(h.data(), ws->d_work, kWorkBytes, hipMemcpyDeviceToHost));
  unsigned long long nmv = 0, nrot = 0, nsym = 0;
  for (int sl = 0; sl < kWorkSlots; ++sl) {
    nmv += h[(size_t)sl * kWorkStride];
    nrot += h[(size_t)sl * kWorkStride + 1];
    nsym += h[(size_t)sl * kWorkStride + 2];
  }
  out[0] = (double)nmv;
  out[1] = (double)nrot;
  static const double kTensorFlops[7] = {0.0,
                                         tensor_matvec_flops<1>(), tensor_matvec_flops<2>(),
                                         tensor_matvec_flops<3>(), tensor_matvec_flops<4>(),
                                         tensor_matvec_flops<5>(), tensor_matvec_flops<6>()};
  static const double kEulerFlops[7] = {euler_flops<0>(), euler_flops<1>(), euler_flops<2>(),
                                        euler_flops<3>(), euler_flops<4>(), euler_flops<5>(),
                                        euler_flops<6>()};
  static const double kFoldedFlops[7] = {
      0.0, tensor_matvec_folded_flops<1>(), tensor_matvec_folded_flops<2>(),
      tensor_matvec_folded_flops<3>(), tensor_matvec_folded_flops<4>(),
      tensor_matvec_folded_flops<5>(), tensor_matvec_folded_flops<6>()};
  out[2] = !ws->tensor ? 0.0 : ws->prec == BEC_FP64 ? kFoldedFlops[ws->F] : kTensorFlops[ws->F];
  out[3] = kEulerFlops[ws->F];
  for (int k = 0; k < BEC_K_COUNT; ++k) {
    out[4 + k] = ws->prof_bytes[k];
    ws->prof_bytes[k] = 0.0;
  }
  out[4 + BEC_K_COUNT] = ws->prof_fft_flops;
  ws->prof_fft_flops = 0.0;
  static const double kFirstSaving[7] = {0.0, first_term_saving<1>(), first_term_saving<2>(),
                                         first_term_saving<3>(), first_term_saving<4>(),
                                         first_term_saving<5>(), first_term_saving<6>()};
  out[5 + BEC_K_COUNT] = (double)nsym;
  out[6 + BEC_K_COUNT] = !ws->tensor ? 0.0 : out[2] - kFirstSaving[ws->F];
  HIPCHK(hipMemset(ws->d_work, 0, kWorkBytes));
  return BEC_OK;
}

int bec_profile_read(bec_ws* ws, double* ms, long long* count) {
  if (!ws || !ms || !count) return fail(BEC_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(ws->stream));
  for (int k = 0; k < BEC_K_COUNT; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  for (size_t i = 0; i < ws->prof_cls.size(); ++i) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ws->ev_pool[2 * i], ws->ev_pool[2 * i + 1]));
    const int c = ws->prof_cls[i];
    if (c >= 0 && c < BEC_K_COUNT) {
      ms[c] += t;
      count[c] += 1;
    }
  }
  ws->ev_used = 0;
  ws->prof_cls.clear();
  return BEC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Grouped split steps over several workspaces (batched solves of different grid shapes):
// each position of the common launch sequence is one kernel launch for all workspaces.
// ---------------------------------------------------------------------------
struct bec_group {
  std::vector<bec_ws*> ws;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<hipEvent_t> ready;
  void* d_blob = nullptr;  // descriptors of the uncaptured path
  size_t blob_cap = 0;
  struct Graph {
    int n;
    std::vector<unsigned char> blob;
    void* d_blob;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  int grouped = 0;  // whether the last call issued grouped launches
};

namespace {
struct GPos {
  int kind, mode;
  size_t off;
  int nd, blocks, threads;
  size_t lds;
  // x pass: launches of at most kXGroupMax descriptors each (kernel arguments; the copy in
  // the blob only keys the captured graphs)
  std::vector<std::vector<XGroupDesc>> xchunks;
  std::vector<int> xblocks;
};

size_t blob_align(size_t b) { return (b + 255) / 256 * 256; }

template <typename Desc>
size_t blob_put(std::vector<unsigned char>& blob, const std::vector<Desc>& d) {
  const size_t off = blob_align(blob.size());
  blob.resize(off + sizeof(Desc) * d.size());
  std::memcpy(blob.data() + off, d.data(), sizeof(Desc) * d.size());
  return off;
}

// The launch plan of n_steps split steps for every workspace with active instances, or an
// empty plan when a position has no grouped kernel (shape, pass or schedule mismatch).
int group_plan(bec_group* grp, int n_steps, std::vector<bec_ws*>& live, std::vector<GPos>& plan,
               std::vector<unsigned char>& blob) {
  plan.clear();
  blob.clear();
  live.clear();
  for (bec_ws* w : grp->ws)
    if (n_active(w) > 0) live.push_back(w);
  if (live.empty()) return BEC_OK;
  std::vector<std::vector<Rec>> recs(live.size());
  for (size_t i = 0; i < live.size(); ++i) {
    bec_ws* w = live[i];
    if (w->prof_on) return BEC_OK;
    w->rec = &recs[i];
    const int rc = bec_split_step(w, n_steps);
    w->rec = nullptr;
    CHK(rc);
  }
  const size_t npos = recs[0].size();
  for (auto& r : recs)
    if (r.size() != npos) return BEC_OK;
  const bool f32 = live[0]->prec == BEC_FP32;
  const size_t csz = f32 ? 8 : 16;
  std::vector<GPos> out;
  for (size_t p = 0; p < npos; ++p) {
    const Rec& r0 = recs[0][p];
    for (auto& r : recs)
      if (r[p].kind != r0.kind || (r0.kind == R_STRIDED && (r[p].axis != r0.axis ||
                                                             r[p].mode != r0.mode)))
        return BEC_OK;
    GPos pos{};
    pos.kind = r0.kind;
    pos.mode = r0.mode;
    pos.nd = (int)live.size();
    int blocks = 0;
    if (r0.kind == R_X) {
      std::vector<XGroupDesc> d(live.size());
      for (size_t i = 0; i < live.size(); ++i) {
        bec_ws* w = live[i];
        if (!xgroup_ok(w->g)) return BEC_OK;
        const int thr = xpass_threads_g(w->g);
        if (i == 0) pos.threads = thr;
        else if (thr != pos.threads) return BEC_OK;
        XGroupDesc& x = d[i];
        std::memset(&x, 0, sizeof(x));
        x.g = w->g;
        x.tb = w->tb;
        x.prog = recs[i][p].xp;
        x.src = recs[i][p].src;
        x.dst = recs[i][p].dst;
        x.fhat = w->fhat;
        x.phi = w->phi;
        x.rows = xpass_rows_g(w->g);
        x.gx = xpass_grid(w->g);
        if (i % kXGroupMax == 0) {
          pos.xchunks.emplace_back();
          pos.xblocks.push_back(0);
        }
        x.block0 = pos.xblocks.back();
        pos.xblocks.back() += x.gx * n_active(w);
        blocks += x.gx * n_active(w);
        pos.xchunks.back().push_back(x);
        pos.lds = std::max(pos.lds, xpass_lds_g(w->g, x.prog, csz));
      }
      pos.off = blob_put(blob, d);
    } else if (r0.kind == R_STRIDED) {
      std::vector<SGroupDesc> d(live.size());
      for (size_t i = 0; i < live.size(); ++i) {
        bec_ws* w = live[i];
        const Rec& r = recs[i][p];
        if (!sgroup_ok(w->g, r.axis, r.mode, r.ddi)) return BEC_OK;
        SGroupDesc& x = d[i];
        std::memset(&x, 0, sizeof(x));
        x.g = w->g;
        x.tb = w->tb;
        x.data = r.data;
        x.kin_frac = r.kin_frac;
        x.dt_override = r.dt_ovr;
        x.axis = r.axis;
        x.itp = r.itp;
        const int TB = 2048 / w->g.n[r.axis];
        const int n_other = r.axis == 1 ? w->g.n[2] : w->g.n[1];
        x.gx = ((w->g.n[0] + TB - 1) / TB) * n_other;
        x.gy = w->g.D;
        x.block0 = blocks;
        blocks += x.gx * x.gy * n_active(w);
      }
      pos.threads = kSGroupThreads;
      pos.off = blob_put(blob, d);
    } else if (r0.kind == R_REDUCE) {
      std::vector<RGroupDesc> d(live.size());
      for (size_t i = 0; i < live.size(); ++i) {
        bec_ws* w = live[i];
        const Rec& r = recs[i][p];
        RGroupDesc& x = d[i];
        std::memset(&x, 0, sizeof(x));
        x.tb = w->tb;
        x.red = w->d_red;
        x.nblocks = r.nblocks;
        x.slot_mask = r.slots;
        x.max_mask = r.maxmask;
        x.gx = __builtin_popcount(r.slots);
        x.block0 = blocks;
        blocks += x.gx * n_active(w);
      }
      pos.off = blob_put(blob, d);
    } else if (r0.kind == R_SCALE) {
      std::vector<CGroupDesc> d(live.size());
      for (size_t i = 0; i < live.size(); ++i) {
        bec_ws* w = live[i];
        CGroupDesc& x = d[i];
        std::memset(&x, 0, sizeof(x));
        x.g = w->g;
        x.tb = w->tb;
        x.psi = w->slot[0];
        x.gx = (int)std::min<long long>(1024, (w->g.inst_stride + 255) / 256);
        x.block0 = blocks;
        blocks += x.gx * n_active(w);
      }
      pos.off = blob_put(blob, d);
    } else {
      return BEC_OK;
    }
    pos.blocks = blocks;
    out.push_back(pos);
  }
  plan.swap(out);
  return BEC_OK;
}

int group_issue(bec_group* grp, const std::vector<GPos>& plan, const unsigned char* d_blob,
                int prec, int D) {
  const hipStream_t st = grp->stream;
  for (const GPos& p : plan) {
    const void* d = d_blob + p.off;
    hipError_t e = hipSuccess;
    if (p.kind == R_X) {
      for (size_t c = 0; c < p.xchunks.size() && e == hipSuccess; ++c) {
        const auto& xc = p.xchunks[c];
        e = prec == BEC_FP32 ? launch_xpass_group<float>(st, D, xc.data(), (int)xc.size(),
                                                         p.xblocks[c], p.threads, p.lds)
                             : launch_xpass_group<double>(st, D, xc.data(), (int)xc.size(),
                                                          p.xblocks[c], p.threads, p.lds);
      }
    } else if (p.kind == R_STRIDED) {
      e = prec == BEC_FP32
              ? launch_strided_group<float>(st, p.mode, (const SGroupDesc*)d, p.nd, p.blocks)
              : launch_strided_group<double>(st, p.mode, (const SGroupDesc*)d, p.nd, p.blocks);
    } else if (p.kind == R_REDUCE) {
      e = launch_reduce_group(st, (const RGroupDesc*)d, p.nd, p.blocks);
    } else {
      e = prec == BEC_FP32 ? launch_scale_group<float>(st, (const CGroupDesc*)d, p.nd, p.blocks)
                           : launch_scale_group<double>(st, (const CGroupDesc*)d, p.nd, p.blocks);
    }
    HIPCHK(e);
  }
  return BEC_OK;
}
}  // namespace

extern "C" {

int bec_group_create(bec_ws* const* wss, int n, bec_group** out) {
  if (!wss || !out || n < 1) return fail(BEC_EINVAL, "bec_group_create: null argument or n < 1");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!wss[i]) return fail(BEC_EINVAL, "bec_group_create: null workspace");
    if (wss[i]->prec != wss[0]->prec || wss[i]->D != wss[0]->D ||
        wss[i]->g.ndim != wss[0]->g.ndim || wss[i]->itp != wss[0]->itp)
      return fail(BEC_EINVAL, "bec_group_create: workspaces of a group share precision, spin, "
                              "dimension and time kind");
    for (int j = 0; j < i; ++j)
      if (wss[j] == wss[i]) return fail(BEC_EINVAL, "bec_group_create: workspace listed twice");
  }
  bec_group* g = new bec_group();
  g->ws.assign(wss, wss + n);
  auto cleanup = [&](const char* m) {
    bec_group_destroy(g);
    return fail(BEC_EDEVICE, m);
  };
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup("hipStreamCreate failed");
  if (hipEventCreateWithFlags(&g->done, hipEventDisableTiming) != hipSuccess)
    return cleanup("hipEventCreate failed");
  g->ready.assign(n, nullptr);
  for (int i = 0; i < n; ++i)
    if (hipEventCreateWithFlags(&g->ready[i], hipEventDisableTiming) != hipSuccess)
      return cleanup("hipEventCreate failed");
  *out = g;
  return BEC_OK;
}

void bec_group_destroy(bec_group* g) {
  if (!g) return;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  for (auto& e : g->graphs) {
    (void)hipGraphExecDestroy(e.exec);
    (void)hipFree(e.d_blob);
  }
  if (g->d_blob) (void)hipFree(g->d_blob);
  for (hipEvent_t e : g->ready)
    if (e) (void)hipEventDestroy(e);
  if (g->done) (void)hipEventDestroy(g->done);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

int bec_group_split_step(bec_group* grp, int n_steps, int use_graph) {
  if (!grp) return fail(BEC_EINVAL, "null group");
  if (n_steps < 0) return fail(BEC_EINVAL, "n_steps < 0");
  grp->grouped = 0;
  if (n_steps == 0) return BEC_OK;
  std::vector<bec_ws*> live;
  std::vector<GPos> plan;
  std::vector<unsigned char> blob;
  CHK(group_plan(grp, n_steps, live, plan, blob));
  if (live.empty()) return BEC_OK;
  if (plan.empty()) {  // no grouped kernel for some position: each workspace on its own
    for (bec_ws* w : live) CHK(use_graph ? bec_split_step_graph(w, n_steps) : bec_split_step(w, n_steps));
    return BEC_OK;
  }
  // the group's stream runs after the work already queued on every member's stream, and
  // the members' streams after the group's launches
  for (size_t i = 0; i < grp->ws.size(); ++i) {
    HIPCHK(hipEventRecord(grp->ready[i], grp->ws[i]->stream));
    HIPCHK(hipStreamWaitEvent(grp->stream, grp->ready[i], 0));
  }
  const int prec = live[0]->prec, D = live[0]->D;
  if (use_graph) {
    bec_group::Graph* hit = nullptr;
    for (auto& e : grp->graphs)
      if (e.n == n_steps && e.blob == blob) hit = &e;
    if (!hit) {
      constexpr size_t kMaxGraphs = 16;
      if (grp->graphs.size() >= kMaxGraphs) {
        HIPCHK(hipStreamSynchronize(grp->stream));
        (void)hipGraphExecDestroy(grp->graphs.front().exec);
        (void)hipFree(grp->graphs.front().d_blob);
        grp->graphs.erase(grp->graphs.begin());
      }
      void* d = nullptr;
      HIPCHK(hipMalloc(&d, std::max<size_t>(blob.size(), 256)));
      if (hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return fail(BEC_EDEVICE, "group descriptor copy failed");
      }
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(grp->stream, hipStreamCaptureModeThreadLocal));
      const int rc = group_issue(grp, plan, (const unsigned char*)d, prec, D);
      const hipError_t ec = hipStreamEndCapture(grp->stream, &graph);
      hipGraphExec_t exec = nullptr;
      hipError_t ei = hipErrorUnknown;
      if (rc == BEC_OK && ec == hipSuccess) ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (rc != BEC_OK || ec != hipSuccess || ei != hipSuccess) {
        (void)hipFree(d);
        if (rc != BEC_OK) return rc;
        HIPCHK(ec);
        HIPCHK(ei);
      }
      grp->graphs.push_back({n_steps, blob, d, exec});
      hit = &grp->graphs.back();
    }
    HIPCHK(hipGraphLaunch(hit->exec, grp->stream));
  } else {
    HIPCHK(hipStreamSynchronize(grp->stream));  // the previous call's launches read d_blob
    if (blob.size() > grp->blob_cap) {
      if (grp->d_blob) (void)hipFree(grp->d_blob);
      grp->d_blob = nullptr;
      grp->blob_cap = 0;
      HIPCHK(hipMalloc(&grp->d_blob, blob.size()));
      grp->blob_cap = blob.size();
    }
    HIPCHK(hipMemcpy(grp->d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice));
    CHK(group_issue(grp, plan, (const unsigned char*)grp->d_blob, prec, D));
  }
  HIPCHK(hipEventRecord(grp->done, grp->stream));
  for (bec_ws* w : grp->ws) HIPCHK(hipStreamWaitEvent(w->stream, grp->done, 0));
  grp->grouped = 1;
  return BEC_OK;
}

int bec_group_grouped(const bec_group* grp) { return grp ? grp->grouped : 0; }

}  // extern "C"
