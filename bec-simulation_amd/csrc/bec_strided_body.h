// Body of the y / z pass kernel (modes and layout: bec_strided.hip), pasted by the
// preprocessor into the kernel itself and into the grouped launch's per-tile function
// (bec_group.hip).  Names in scope at the paste: T, N, MODE, TB, FIELDS, g, tb, data, axis,
// kin_frac, dt_override, itp, ddi_data, and the macros BEC_BX (x chunk and line of the
// other transverse axis), BEC_BY (field group), BEC_BZ (index into the active-instance
// list) and BEC_GX (x tiles of the launch, for the SM_KE partial index).
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int E = fft_e<N>();  // values per thread (8, 12 or 20)
  constexpr int TEAM = N / E;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int inst = tb.active[BEC_BZ];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem);
  double* tabs = reinterpret_cast<double*>(tile + FIELDS * N * TB);  // [2*(N+TB+1)] (re,im) pairs
  double* rsh = tabs + 2 * (N + TB + 1);
  cplx<double>* tws = reinterpret_cast<cplx<double>*>(rsh + 16);  // N twiddles
  const cplx<double>* twg = tw_table<N>(tb);
  const cplx<double>* tw = tws;

  const int pitch = ddi_data ? g.nxh_pitch : g.n[0];
  const int valid = ddi_data ? g.nxh : g.n[0];
  const int n1 = ddi_data ? g.sn[1] : g.n[1];
  const long long fstride = ddi_data ? g.fhat_field_stride : g.nsites;
  const long long istride = ddi_data ? g.fhat_inst_stride : g.inst_stride;
  const long long estride = axis == 1 ? (long long)pitch : (long long)pitch * n1;
  const long long ostride = axis == 1 ? (long long)pitch * n1 : (long long)pitch;
  // zero padding (DDI spectra of a padded grid): a forward transform reads only the first
  // n[axis] entries (the rest are zero), and an inverse one keeps only those outputs
  const bool padded = ddi_data && g.spad;
  const int n_in = (padded && MODE != SM_INV) ? g.n[axis] : N;
  const int n_out = (padded && MODE != SM_FWD) ? g.n[axis] : N;
  const int nxch = (valid + TB - 1) / TB;
  const int xch = BEC_BX % nxch;
  const int other = BEC_BX / nxch;
  const int x0 = xch * TB;
  const size_t base = (size_t)inst * istride + (size_t)BEC_BY * FIELDS * fstride +
                      (size_t)other * ostride + x0;

  // All global loads first (each thread owns exactly E tile elements: the launch has
  // FIELDS*TB*N/E threads), then the LDS writes: one memory latency per block.
  cplx<T> ld[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int idx = tid + k * nthr;
    const int xi = idx % TB, rest = idx / TB, e = rest % N, f = rest / N;
    const bool in = idx < FIELDS * N * TB && x0 + xi < valid && e < n_in;
    // out-of-range lanes read element 0 of their line and store zero
    const cplx<T> v = data[base + (in ? (size_t)f * fstride + (size_t)e * estride + xi : 0)];
    ld[k] = in ? v : mk((T)0, (T)0);
  }
  // N <= TWK * nthr (nthr = FIELDS*TB*N/E); clamped, not guarded
  constexpr int TWK = (E + FIELDS * TB - 1) / (FIELDS * TB);
  cplx<double> twv[TWK];
#pragma unroll
  for (int k = 0; k < TWK; ++k) twv[k] = twg[min(tid + k * nthr, N - 1)];
  // unguarded writes (the launch has exactly FIELDS*N*TB/E threads; a twiddle index past
  // N - 1 rewrites entry N - 1 with the value it holds), so no load is sunk into a branch
#pragma unroll
  for (int k = 0; k < E; ++k) tile[tid + k * nthr] = ld[k];
#pragma unroll
  for (int k = 0; k < TWK; ++k) tws[min(tid + k * nthr, N - 1)] = twv[k];
  // per-axis phase tables for the kinetic multiply: entries [0, N) along the line, [N, N+TB)
  // for the tile's x columns; the factor of the other transverse axis (y for the z pass) and
  // the 1/N of the inverse transform are folded into the line entries, so each element
  // takes one FP64 complex product of two table entries
  if (MODE == SM_KIN) {
    const InstParams& P = tb.params[inst];
    const double dt = (dt_override > 0.0 ? dt_override : P.dt) * kin_frac;
    auto phase = [&](double k, double& re, double& im) __attribute__((always_inline)) {
      const double a = 0.5 * (k * k) * dt;
      if (itp) { re = exp(-a); im = 0.0; }
      else { double s, c; sincos_d(a, &s, &c); re = c; im = -s; }
    };
    double ore, oim;  // y factor of the z pass
    phase((g.ndim == 3 && axis == 2) ? tb.ktab[g.nmax + other] : 0.0, ore, oim);
    for (int i = tid; i < N + TB; i += nthr) {
      double k;
      if (i < N) k = tb.ktab[axis * g.nmax + i];
      else k = (x0 + (i - N) < g.n[0]) ? tb.ktab[x0 + (i - N)] : 0.0;
      double re, im;
      phase(k, re, im);
      if (i < N) {
        const cplx<double> v = (mk(re, im) * mk(ore, oim)) * g.inv_ntot;
        re = v.x;
        im = v.y;
      }
      tabs[2 * i] = re;
      tabs[2 * i + 1] = im;
    }
  }
  __syncthreads();

  const int xi = tid % TB, rest = tid / TB, tt = rest % TEAM, f = rest / TEAM;
  const bool act = f < FIELDS;
  cplx<T>* lbase = tile + (act ? f : 0) * N * TB + xi;
  if (MODE != SM_INV) fft_team<T, N, false, E>(lbase, TB, tt, act, tw);

  if (MODE == SM_KIN) {
    for (int idx = tid; idx < FIELDS * N * TB; idx += nthr) {
      const int xi2 = idx % TB, e = (idx / TB) % N;
      const cplx<double> pa = mk(tabs[2 * e], tabs[2 * e + 1]);
      const cplx<double> px = mk(tabs[2 * (N + xi2)], tabs[2 * (N + xi2) + 1]);
      tile[idx] = mul64(tile[idx], px * pa);
    }
    __syncthreads();
  } else if (MODE == SM_DDI) {
    const InstParams& P = tb.params[inst];
    const double C = P.cdd;
    for (int idx = tid; idx < N * TB; idx += nthr) {
      const int xi2 = idx % TB, e = idx / TB;
      const int x = x0 + xi2;
      // rfftfreq along x (Nyquist +), fftfreq along y and z (ddi.jl:66-68)
      double kx = 0.0;
      if (x < valid) kx = (x == g.sn[0] / 2) ? -tb.sktab[g.sn[0] / 2] : tb.sktab[x];
      double ky, kz;
      if (axis == 2) { ky = tb.sktab[g.smax + other]; kz = tb.sktab[2 * g.smax + e]; }
      else { ky = tb.sktab[g.smax + e]; kz = 0.0; }
      cplx<T> fx = tile[(0 * N + e) * TB + xi2];
      cplx<T> fy = tile[(1 * N + e) * TB + xi2];
      cplx<T> fz = tile[(2 * N + e) * TB + xi2];
      // normalisation of the inverse transform of the (possibly padded) spectral grid
      const double s = C / ((double)g.sn[0] * (double)g.sn[1] * (double)g.sn[2]);
      ddi_qmix<T>(kx, ky, kz, g.ndim, g.ddi_secular != 0, s, fx, fy, fz);
      tile[(0 * N + e) * TB + xi2] = fx;
      tile[(1 * N + e) * TB + xi2] = fy;
      tile[(2 * N + e) * TB + xi2] = fz;
    }
    __syncthreads();
  } else if (MODE == SM_DERIV) {
    // fft_utils.jl:4-19,25-44: component c holds f and takes i k_c / N (c < ndim)
    const int c = BEC_BY;
    for (int idx = tid; idx < N * TB; idx += nthr) {
      const int xi2 = idx % TB, e = idx / TB;
      const int x = x0 + xi2;
      double k = 0.0;
      if (c == 0) k = x < valid ? tb.ktab[x] : 0.0;
      else if (c == 1) k = tb.ktab[g.nmax + (axis == 1 ? e : other)];
      else if (c == 2 && g.ndim == 3) k = tb.ktab[2 * g.nmax + e];
      tile[idx] = mul64(tile[idx], mk(0.0, k * g.inv_ntot));
    }
    __syncthreads();
  } else if (MODE == SM_KE) {
    double acc = 0.0;
    for (int idx = tid; idx < FIELDS * N * TB; idx += nthr) {
      const int xi2 = idx % TB, e = (idx / TB) % N;
      const int x = x0 + xi2;
      if (x >= valid) continue;
      const double kx = tb.ktab[x];
      double k2 = kx * kx;
      if (axis == 2) {
        const double ky = tb.ktab[g.nmax + other], kz = tb.ktab[2 * g.nmax + e];
        k2 += ky * ky;
        k2 += kz * kz;
      } else {
        const double ky = tb.ktab[g.nmax + e];
        k2 += ky * ky;
      }
      acc += k2 * (double)abs2(tile[idx]);
    }
    const double v = block_sum(acc, rsh);
    if (tid == 0)
      *partial_ptr(tb, RS_KIN, inst, BEC_BX + BEC_GX * BEC_BY) = v;
    return;
  }

  if (MODE == SM_INV || MODE == SM_KIN || MODE == SM_DDI || MODE == SM_DERIV)
    fft_team<T, N, true, E>(lbase, TB, tt, act, tw);

  for (int idx = tid; idx < FIELDS * N * TB; idx += nthr) {
    const int xi2 = idx % TB, rest2 = idx / TB, e = rest2 % N, f2 = rest2 / N;
    if (x0 + xi2 < valid && e < n_out)
      data[base + (size_t)f2 * fstride + (size_t)e * estride + xi2] = tile[idx];
  }
