// Body of the x-pass kernel (stages and layout: bec_kernels.hip), pasted by the
// preprocessor into the kernel itself and into the grouped launch's per-tile function
// (bec_group.hip), so that both compile from one text while the single-workspace kernel
// keeps its kernel-argument code generation.  Names in scope at the paste: T, NX, D, g,
// tb, prog, src, dst, fhat, phi, ROWS, and the macros BEC_BX (x tile) and BEC_BY (index
// into the active-instance list), and BEC_REPEAT (1 in the repeat kernel of 1-D lines).
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int EX = xpass_e<NX, D>();
  constexpr int TEAM = NX / EX;
  const int tid = threadIdx.x;
#if BEC_REPEAT
  // The repeat kernel runs one thread per site (at least): the tile's FFT threads
  // (nthr = D ROWS NX / EX, the same assignment of loads, stores and norm partials as the
  // single-step kernel) plus threads that only take part in the per-site operators and
  // elementwise loops (nsite-strided) and in the barriers.
  const int nthr = D * ROWS * NX / EX;
  const bool own = threadIdx.x < nthr;
#else
  const int nthr = blockDim.x;
  constexpr bool own = true;
#endif
  const int nsite = blockDim.x;
  const int inst = tb.active[BEC_BY];
  const InstParams& P = tb.params[inst];
  const double dt = prog.dt_override > 0.0 ? prog.dt_override : P.dt;
  const bool itp = prog.itp != 0;
  const int st = prog.stages;
  const bool has_aux = (st & (XS_LOAD_PHI | XS_WRITE_FHAT)) != 0;
  const int TS = ROWS * NX;  // sites in this tile (a power of two, as are the grid sides)
  const int rows_shift = __ffs(ROWS) - 1;  // rows per block: a power of two
  // LDS tiles are padded (one slot per 8 elements, lds_at<true>): element e of the tile
  // lives at P(e) = e + e/8; lines start at multiples of NX, so P(line*NX + i) =
  // P(line*NX) + P(i).
  auto P8 = [](int e) __attribute__((always_inline)) { return e + (e >> 3); };

  // DDI auxiliary lines: 2 per row, of the spectral x length (2 NX when zero-padded)
  const int AL = g.spad ? 2 * NX : NX;
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem);
  // the padded tile rounded up to an even element count (odd only for 8-point 1-D lines),
  // so that what follows starts 16-byte aligned
  cplx<T>* aux = tile + ((size_t)D * TS / 8 * 9 + 1) / 2 * 2;
  // FP64 Zeeman factors (ROWS*AL/8*9 is even, so this stays 16-byte aligned)
  cplx<double>* zf =
      reinterpret_cast<cplx<double>*>(aux + (has_aux ? 2 * ROWS * AL / 8 * 9 : 0));
  double* rsh = reinterpret_cast<double*>(zf + 2 * kMaxOps * D);
  cplx<double>* tws = reinterpret_cast<cplx<double>*>(rsh + 16);  // NX twiddles

  const cplx<double>* twg = tw_table<NX>(tb);
  const cplx<double>* tw = tws;
  const int row0 = BEC_BX * ROWS;
  const size_t ibase = (size_t)inst * g.inst_stride + (size_t)row0 * NX;

  // Every global load of the prologue is issued before the first wait: each thread loads
  // exactly EX tile elements (the launch has D*ROWS*NX/EX threads) and up to EX/3
  // twiddles (nthr >= 3 NX / EX), so the block waits for one memory latency, not for one
  // per element, and the FFT passes read their twiddles from LDS.
  // (indices are clamped rather than guarded, so the arrays stay in registers)
  constexpr int TWK = (EX + 2) / 3;
  cplx<double> twv[TWK];
#pragma unroll
  for (int k = 0; k < TWK; ++k) twv[k] = twg[min(tid + k * nthr, NX - 1)];
  // The tile (D contiguous chunks of ROWS*NX elements) and the twiddles into LDS, before
  // any code with a call (the Zeeman factors' sin/cos) so that nothing is live across it.
  // No branch separates a load from its LDS write (a guarded write lets the compiler sink
  // the load into the guard and wait for it there): the launch has exactly D*TS/EX threads,
  // so every tile index is in range, and a twiddle index past NX - 1 rewrites entry NX - 1
  // with the value it already holds.
  if (!(st & XS_NO_LOAD) && own) {
    cplx<T> ld[EX];
#pragma unroll
    for (int k = 0; k < EX; ++k) {
      const int e = tid + k * nthr;
      const int c = (e / NX) >> rows_shift, r = e - c * TS;
      ld[k] = src[ibase + (size_t)c * g.nsites + r];
    }
#pragma unroll
    for (int k = 0; k < TWK; ++k) tws[min(tid + k * nthr, NX - 1)] = twv[k];
#pragma unroll
    for (int k = 0; k < EX; ++k) tile[P8(tid + k * nthr)] = ld[k];
  } else if (own) {
#pragma unroll
    for (int k = 0; k < TWK; ++k) tws[min(tid + k * nthr, NX - 1)] = twv[k];
  }

  // Blocks of fewer than 3 NX / EX threads (the scalar gas, F = 0, on one short line: one
  // x line per block) cover the rest of the twiddle table here; elsewhere this loop is empty.
  for (int i = tid + TWK * nthr; i < NX; i += nthr) tws[i] = twg[i];

  // Op lists in LDS (A at [0, kMaxOps), B at [kMaxOps, 2 kMaxOps)), copied with
  // compile-time indices so the by-value kernel argument is never indexed at run time.
  // Slot i is copied by thread i, or by thread 0 when the block is smaller than that (1-D
  // lines of 16 or 32 points run blocks of 6 or 12 threads).
  __shared__ Op s_ops[2 * kMaxOps];
#if BEC_REPEAT
  // per-launch constants of the repeated step: the trap potential of every site and the
  // 1-D kinetic factor of every x (lines of at most 512 points)
  __shared__ double s_vtr[512];
  __shared__ cplx<double> s_kph[512];
#endif
#pragma unroll
  for (int i = 0; i < kMaxOps; ++i) {
    if (tid == (i < nthr ? i : 0)) s_ops[i] = prog.A[i];
    if (tid == (kMaxOps + i < nthr ? kMaxOps + i : 0)) s_ops[kMaxOps + i] = prog.B[i];
  }
  __syncthreads();
  // Zeeman factors of every diagonal op (uniform over the block).
  for (int idx = tid; idx < 2 * kMaxOps * D; idx += nsite) {
    const int i = idx / D, c = idx % D;
    const bool used = i < kMaxOps ? i < prog.nA : (i - kMaxOps) < prog.nB;
    const Op op = s_ops[i];
    if (used && op.kind == OP_DIAG) {
      const double dtf = op.frac * dt;
      if (itp) zf[idx] = mk(exp(-(P.zee[c] - P.zee_shift) * dtf), 0.0);
      else zf[idx] = cis_neg<double>(P.zee[c] * dtf);
    }
  }
  __syncthreads();
  double norm_acc = 0.0, ke_acc = 0.0, phimax = 0.0;

  const int line = tid / TEAM, tt = tid - (tid / TEAM) * TEAM;
  if (st & XS_INV_FFT)
    fft_team<T, NX, true, EX, true>(tile + P8(line * NX), 1, tt, line < D * ROWS, tw);

  // ITP renormalisation factor of the instance (only used when an OP_SCALE is present)
#if BEC_REPEAT
  double sc = 1.0 / sqrt(tb.red[(size_t)inst * RS_COUNT + RS_NORM] * g.dV);
#else
  const double sc = 1.0 / sqrt(tb.red[(size_t)inst * RS_COUNT + RS_NORM] * g.dV);
#endif

  if (st & XS_LOAD_PHI) {
    ddi_lines_in<T>(g, fhat, (size_t)inst * g.fhat_inst_stride, row0, ROWS, AL, aux, tid, nthr);
    __syncthreads();
    ddi_lines_fft<T, NX, true>(tb, aux, ROWS, AL, tid, nthr, tw);
  }

  auto site_coords = [&](int s, int& x, int& y, int& z) {
    const int r = s / NX;
    x = s - r * NX;
    const int row = row0 + r;
    z = div_side(row, g.n[1]);
    y = row - z * g.n[1];
  };
  auto trap_at = [&](int s) -> double {
    int x, y, z;
    site_coords(s, x, y, z);
    if (tb.Varr) return tb.Varr[(size_t)inst * g.nsites + (size_t)row0 * NX + s];
    const double* ax = tb.axtab + (size_t)inst * 3 * g.nmax;
    double v = ax[x];
    if (g.ndim >= 2) v += ax[g.nmax + y];
    if (g.ndim >= 3) v += ax[2 * g.nmax + z];
    return 0.5 * v;
  };

  // Per-site operators, one site loop per operator: each loop body holds a single
  // operator, so the register allocation of the kernel is the largest single operator
  // (not their sum), and every site stays owned by the same thread across operators.
  // When a tile has no more sites than threads (E <= D) the "loop" is a single guarded
  // iteration, so no uniform table load can be hoisted out of it into VGPRs.
  auto for_sites = [&](auto&& body) __attribute__((always_inline)) {
    if constexpr (EX <= D) {
      if (tid < TS) body(tid);
    } else {
      for (int s = tid; s < TS; s += nsite) body(s);
    }
  };
  auto run_ops = [&](const Op* ops, int nops, int zbase) {
    for (int i = 0; i < nops; ++i) {
      const Op op = ops[i];
      const double dtf = op.frac * dt;
      if (op.kind == OP_DIAG) {
        const cplx<double>* zfi = &zf[(zbase + i) * D];
        for_sites([&](int s) __attribute__((always_inline)) {
          // op_diag with the components left in LDS: a density sweep, then the phase
          // products (the barrier makes the second sweep re-read LDS instead of keeping
          // all D values live in VGPRs)
          T n = 0;
#pragma unroll
          for (int c = 0; c < D; ++c) n = density_acc(n, tile[P8(c * TS + s)]);
#if BEC_REPEAT
          const cplx<double> base = diag_base((double)n, s_vtr[s], P, dtf, itp);
#else
          const cplx<double> base = diag_base((double)n, trap_at(s), P, dtf, itp);
#endif
          __asm__ volatile("" ::: "memory");
#pragma unroll
          for (int c = 0; c < D; ++c) {
            const int e = P8(c * TS + s);
            tile[e] = mul_split<T>(tile[e], cmul_fixed(base, zfi[c]));
          }
        });
      } else if (op.kind == OP_SPINMIX && D == 3) {
        for_sites([&](int s) __attribute__((always_inline)) {
          cplx<T> sp[D];
#pragma unroll
          for (int c = 0; c < D; ++c) sp[c] = tile[P8(c * TS + s)];
          op_spin_mixing<T, D>(sp, P, dtf, itp);
#pragma unroll
          for (int c = 0; c < D; ++c) tile[P8(c * TS + s)] = sp[c];
        });
      } else if (op.kind == OP_LOSS) {
        if (P.has_loss) {
          for_sites([&](int s) __attribute__((always_inline)) {
            // components stay in LDS (density sweep, then the per-component factors)
            double n = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) n += (double)density_acc<T>(0, tile[P8(c * TS + s)]);
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int c = 0; c < D; ++c) {
              if (P.loss[c] < 1e-30) continue;
              const int e = P8(c * TS + s);
              tile[e] = scale64(tile[e], loss_factor(P.loss[c], n, dtf));
            }
          });
        }
      } else if (op.kind == OP_NEMATIC && D < 13) {
        for_sites([&](int s) __attribute__((always_inline)) {
          cplx<T> sp[D];
#pragma unroll
          for (int c = 0; c < D; ++c) sp[c] = tile[P8(c * TS + s)];
          op_nematic<T, D>(sp, P.c2, dtf, itp);
#pragma unroll
          for (int c = 0; c < D; ++c) tile[P8(c * TS + s)] = sp[c];
        });
      } else if (op.kind == OP_SCALE) {
        for (int e = tid; e < D * TS; e += nsite) tile[P8(e)] = scale64(tile[P8(e)], sc);
        __syncthreads();  // the scale loop's element ownership differs from the site loops
      }
    }
    __syncthreads();
  };

  if ((st & XS_LOAD_PHI) && !g.spad) {
    // max |Phi| for the imaginary-time overflow guard (ground_state.jl:4-29); the guard
    // reads the unpadded Phi buffers, which the padded convolution does not write
    for (int s = tid; s < TS; s += nthr) {
      const int r = s / NX, x = s - r * NX;
      const cplx<T> a = aux[P8((2 * r) * AL + x)], b = aux[P8((2 * r + 1) * AL + x)];
      phimax = fmax(phimax, fmax(fabs((double)a.x), fmax(fabs((double)a.y), fabs((double)b.x))));
    }
  }

  if (st & XS_STORE_PHI) {
    T* ph = phi + (size_t)inst * 3 * g.nsites + (size_t)row0 * NX;
    for (int s2 = tid; s2 < TS; s2 += nthr) {
      const int r = s2 / NX, x = s2 - r * NX;
      const cplx<T> a = aux[P8((2 * r) * AL + x)], b = aux[P8((2 * r + 1) * AL + x)];
      ph[s2] = a.x;
      ph[g.nsites + s2] = a.y;
      ph[2 * g.nsites + s2] = b.x;
    }
  }

#if BEC_REPEAT
  // Repeated split steps (the repeat kernel, prog.nrep > 1, 1-D lines): the line stays in
  // LDS from one step to the next.  A repetition's renormalisation uses the previous one's
  // norm reduced as the per-step schedule reduces it (the block partial of the NORM stage,
  // which the one-partial reduction returns unchanged), so the steps give the bits of nrep
  // separate launches.  Only the stages ops A, 1-D kinetic, ops B and norm are repeated.
  const int nrep = prog.nrep > 1 ? prog.nrep : 1;
  // the same values as the single-step kernel computes per use
  for (int s = tid; s < TS; s += nsite) s_vtr[s] = trap_at(s);
  if (st & XS_KIN_1D) {
    const double kdt = prog.kin_frac * dt;
    for (int x = tid; x < NX; x += nsite) {
      const double k = tb.ktab[x];
      const double a = 0.5 * (k * k) * kdt;
      cplx<double> ph = itp ? mk(exp(-a), 0.0) : cis_neg<double>(a);
      s_kph[x] = ph * g.inv_ntot;
    }
  }
  __syncthreads();
  __shared__ double s_norm;
  for (int rep = 0; rep < nrep; ++rep) {
  if (rep > 0 && (st & XS_NORM)) {
    const double v = block_sum(norm_acc, rsh);
    if (tid == 0) s_norm = v;
    __syncthreads();
    sc = 1.0 / sqrt(s_norm * g.dV);
    norm_acc = 0.0;
    for (int e = tid; e < D * TS; e += nsite) tile[P8(e)] = scale64(tile[P8(e)], sc);
    __syncthreads();
  }
#endif

  if (st & XS_OPS_A) run_ops(s_ops, prog.nA, 0);

  if (st & XS_KIN_1D) {
    fft_team<T, NX, false, EX, true>(tile + P8(line * NX), 1, tt, line < D * ROWS, tw);
    const double kdt = prog.kin_frac * dt;
    for (int e = tid; e < D * TS; e += nsite) {
      const int x = e % NX;
      const double k = tb.ktab[x];
      cplx<double> ph;
      if (prog.deriv) {  // fft_utils.jl:4-19 along x: i k_x / N on component 0
        ph = mk(0.0, e < TS ? k * g.inv_ntot : 0.0);
      } else {
#if BEC_REPEAT
        ph = s_kph[x];
#else
        const double a = 0.5 * (k * k) * kdt;
        ph = itp ? mk(exp(-a), 0.0) : cis_neg<double>(a);
        ph = ph * g.inv_ntot;
#endif
      }
      tile[P8(e)] = mul64(tile[P8(e)], ph);
    }
    __syncthreads();
    fft_team<T, NX, true, EX, true>(tile + P8(line * NX), 1, tt, line < D * ROWS, tw);
  }

  if (st & XS_OPS_B) run_ops(s_ops + kMaxOps, prog.nB, kMaxOps);

  if (st & XS_WRITE_FHAT) {
    for (int s = tid; s < TS; s += nthr) {
      cplx<T> sp[D];
#pragma unroll
      for (int c = 0; c < D; ++c) sp[c] = tile[P8(c * TS + s)];
      T fx, fy, fz;
      spin_vec<T, D>(sp, fx, fy, fz);
      const int r = s / NX, x = s - r * NX;
      aux[P8((2 * r) * AL + x)] = mk(fx, fy);
      aux[P8((2 * r + 1) * AL + x)] = mk(fz, (T)0);
    }
    if (AL > NX) ddi_lines_zero_upper<T, NX>(aux, ROWS, AL, tid, nthr);
    __syncthreads();
    ddi_lines_fft<T, NX, false>(tb, aux, ROWS, AL, tid, nthr, tw);
    ddi_lines_out<T>(g, fhat, (size_t)inst * g.fhat_inst_stride, row0, ROWS, AL, aux, tid, nthr);
  }

  if ((st & XS_NORM) && own) {
    for (int s = tid; s < TS; s += nthr) {
#pragma unroll
      for (int c = 0; c < D; ++c) norm_acc += (double)abs2(tile[P8(c * TS + s)]);
    }
  }
#if BEC_REPEAT
  __syncthreads();
  }  // repetitions
#endif

  if (st & XS_FWD_FFT)
    fft_team<T, NX, false, EX, true>(tile + P8(line * NX), 1, tt, line < D * ROWS, tw);

  if ((st & XS_KE_1D) && own) {
    for (int e = tid; e < D * TS; e += nthr) {
      const double k = tb.ktab[e % NX];
      ke_acc += (k * k) * (double)abs2(tile[P8(e)]);
    }
  }

  if (!(st & XS_NO_STORE) && own) {  // EX elements per thread (exact launch shape)
#pragma unroll
    for (int k = 0; k < EX; ++k) {
      const int e = tid + k * nthr;
      const int c = (e / NX) >> rows_shift, r = e - c * TS;
      dst[ibase + (size_t)c * g.nsites + r] = tile[P8(e)];
    }
  }

  // per-block partial sums
  if (st & XS_LOAD_PHI) {
    const double m = block_max(phimax, rsh);
    if (tid == 0) *partial_ptr(tb, RS_PHIMAX, inst, BEC_BX) = m;
  }
  if (st & XS_NORM) {
    const double v = block_sum(norm_acc, rsh);
    if (tid == 0) *partial_ptr(tb, RS_NORM, inst, BEC_BX) = v;
  }
  if (st & XS_KE_1D) {
    const double v = block_sum(ke_acc, rsh);
    if (tid == 0) *partial_ptr(tb, RS_KIN, inst, BEC_BX) = v;
  }
