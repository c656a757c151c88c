// Device building blocks for the split-step hot path on gfx950:
//   * cplx<T>          : plain complex arithmetic (no library types)
//   * fft_team<T,N>    : Stockham radix-8/4/2 FFT of one length-N line held in LDS,
//                        computed by a team of N/8 threads (8 values per thread in
//                        VGPRs, one LDS round trip per radix pass)
//   * per-site spinor propagators: diagonal (trap/Zeeman/contact/LHY),
//     spin-1 Rodrigues rotation, general-F Euler rotation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "bec_types.h"

namespace bec {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cplx<T> operator*(cplx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cplx<T> conjc(cplx<T> a) { return mk(a.x, -a.y); }
template <typename T>
__device__ __forceinline__ T abs2(cplx<T> a) { return a.x * a.x + a.y * a.y; }

// Arithmetic of operators that run in more than one kernel (the diagonal step runs in the
// x pass and in the lane-per-site pass, depending on its neighbours in the schedule).
// Multiply-adds are fused within these expressions only and never with the caller's
// operations after inlining, so every kernel gives the same bits for the same site and a
// result does not depend on how a run is split into launches.
template <typename T>
__device__ __forceinline__ T density_acc(T n, cplx<T> a) {
#pragma clang fp contract(on)
  return n + (a.x * a.x + a.y * a.y);
}
__device__ __forceinline__ cplx<double> cmul_fixed(cplx<double> a, cplx<double> b) {
#pragma clang fp contract(on)
  return cplx<double>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

// Element i of an array at a wave-uniform base, with the per-lane part as a 32-bit byte
// offset (i * sizeof(T) < 4 GiB): the access takes the scalar-base addressing form, one VGPR
// for the lane offset instead of a 64-bit address per array.
template <typename T>
__device__ __forceinline__ T* lane_elem(T* base, int i) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + (unsigned)i * (unsigned)sizeof(T));
}

// Coefficients of sincos_d: the Taylor coefficients of sin (1/17! down to -1/3!, alternating
// signs), of cos (-1/18! down to -1/2!), 2/pi and pi/2 split into three doubles.
struct SinCosCoef {
  double v[20] = {2.8114572543455206e-15, -7.647163731819816e-13, 1.6059043836821613e-10,
                  -2.505210838544172e-08, 2.7557319223985893e-06, -0.0001984126984126984,
                  0.008333333333333333,   -0.16666666666666666,
                  -1.5619206968586225e-16, 4.779477332387385e-14, -1.1470745597729725e-11,
                  2.08767569878681e-09,    -2.755731922398589e-07, 2.48015873015873e-05,
                  -0.001388888888888889,   0.041666666666666664,  -0.5,
                  0.6366197723675814,      1.5707963267948966,    6.123233995736766e-17};
};
template <int I = 0>
__constant__ SinCosCoef kSinCosConst = SinCosCoef{};
using const_f64_ptr = const __attribute__((address_space(4))) double*;

// Out-of-line library sine/cosine for the rare arguments |a| > 1e6 (and NaN), so that
// its large-argument reduction does not set the register budget of the calling kernels.
static __device__ __noinline__ void sincos_lib(double a, double* s, double* c) { sincos(a, s, c); }

// sin and cos of a double.  Cody-Waite reduction a = k pi/2 + x with pi/2 split into three
// doubles (FMA steps, exact enough for |k| < 2^20), then the Taylor series of sin (to x^17)
// and cos (to x^18) on |x| <= pi/4 in Horner form (truncation < 1e-19); the quadrant k mod 4
// selects and signs the pair.  Measured against long-double sinl/cosl: <= 1.5 ulp for
// |a| <= 1e6.  The phase arguments of the split-step operators are far below that bound.
// The coefficients are read from constant memory through a pointer the compiler cannot see
// through: as literals, every FP64 coefficient is a register pair that the compiler hoists
// out of the operator loops of the site kernels, where they spill.
__device__ __forceinline__ void sincos_d(double a, double* s, double* c) {
  if (!(fabs(a) <= 1.0e6)) {
    sincos_lib(a, s, c);
    return;
  }
  const_f64_ptr cf = (const_f64_ptr)kSinCosConst<>.v;
  asm volatile("" : "+s"(cf));
  // |a| <= pi/4 (every phase argument of a small step): k = 0 and x = a exactly, so the
  // reduction and the quadrant selection are skipped with the same result
  const bool small = fabs(a) <= 0.7853981633974483;
  const double k = small ? 0.0 : rint(a * cf[17]);  // 2/pi
  double x = a;
  if (!small) {
    x = fma(-k, cf[18], a);  // pi/2 = hi + mid + lo
    x = fma(-k, cf[19], x);
    x = fma(-k, -1.4973849048591698e-33, x);
  }
  const double z = x * x;
  double ps = cf[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) ps = fma(ps, z, cf[i]);
  const double sn = fma(x * z, ps, x);
  double pc = cf[8];
#pragma unroll
  for (int i = 9; i < 17; ++i) pc = fma(pc, z, cf[i]);
  const double cs = fma(z, pc, 1.0);
  if (small) {
    *s = sn;
    *c = cs;
    return;
  }
  const int q = (int)(long long)k & 3;
  double rs = (q & 1) ? cs : sn, rc = (q & 1) ? sn : cs;
  if (q == 1 || q == 2) rc = -rc;
  if (q & 2) rs = -rs;
  *s = rs;
  *c = rc;
}

// cis(-a) for a phase argument that is always formed in FP64
template <typename T>
__device__ __forceinline__ cplx<T> cis_neg(double a) {
  double s, c;
  sincos_d(a, &s, &c);
  return mk<T>((T)c, (T)(-s));
}

// Multiply by a real FP64 factor with one rounding to T per component (an FP32 copy of
// the factor would bias every element of a renormalised FP32 state the same way).
template <typename T>
__device__ __forceinline__ cplx<T> scale64(cplx<T> a, double f) {
  return mk((T)((double)a.x * f), (T)((double)a.y * f));
}

// Multiply by an FP64 complex factor with one rounding to T per component.  Unit-modulus
// factors (twiddles, kinetic and potential phases) rounded to FP32 first have |w| != 1
// with a bias that does not average out: the FP32 norm would drift every step.
template <typename T>
__device__ __forceinline__ cplx<T> mul64(cplx<T> a, cplx<double> w) {
  const double ax = a.x, ay = a.y;
  return mk((T)(ax * w.x - ay * w.y), (T)(ax * w.y + ay * w.x));
}

// The same product for register-heavy callers: in FP32 the factor is split into
// hi + lo floats and applied with fused multiply-adds (as unbiased as the FP64 form,
// without FP64 temporaries per component).
template <typename T>
__device__ __forceinline__ cplx<T> mul_split(cplx<T> a, cplx<double> f) {
#pragma clang fp contract(on)
  if constexpr (sizeof(T) == 8) {
    return a * f;
  } else {
    const float hx = (float)f.x, hy = (float)f.y;
    const float lx = (float)(f.x - (double)hx), ly = (float)(f.y - (double)hy);
    return mk<T>(fmaf(a.x, hx, fmaf(-a.y, hy, fmaf(a.x, lx, -a.y * ly))),
                 fmaf(a.x, hy, fmaf(a.y, hx, fmaf(a.x, ly, a.y * lx))));
  }
}

// Sum |s_c|^2 in FP64.
template <typename T, int D>
__device__ __forceinline__ double site_norm2(const cplx<T>* s) {
  double n = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) n += (double)s[c].x * s[c].x + (double)s[c].y * s[c].y;
  return n;
}

// FP32-bulk, real time: a per-site unitary step (spin rotation, tensor step) preserves
// sum_c |psi_c|^2 exactly, but FP32 copies of its fixed matrices are not exactly
// unitary, and a state whose spin points the same way at every site picks up the same
// norm error everywhere.  Restore the per-site norm (one FP64 factor).
template <typename T, int D>
__device__ __forceinline__ void restore_norm(cplx<T>* s, double n0) {
  if constexpr (sizeof(T) == 4) {
    const double n1 = site_norm2<T, D>(s);
    if (n1 > 0.0) {
      const double f = sqrt(n0 / n1);
#pragma unroll
      for (int c = 0; c < D; ++c) s[c] = scale64(s[c], f);
    }
  }
}

// x / sqrt(2) with a single rounding (the FP32 constant 1/sqrt(2) is 1.7e-8 low).
template <typename T>
__device__ __forceinline__ T mul_rsqrt2(T x) {
  return (T)((double)x * 0.70710678118654752440);
}

// ---------------------------------------------------------------------------
// In-register small DFTs
// ---------------------------------------------------------------------------
template <typename T, bool INV>
__device__ __forceinline__ void dft2(cplx<T>& a, cplx<T>& b) {
  cplx<T> t = a - b;
  a = a + b;
  b = t;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cplx<T>& v0, cplx<T>& v1, cplx<T>& v2, cplx<T>& v3) {
  cplx<T> a0 = v0 + v2, a1 = v0 - v2, a2 = v1 + v3, a3 = mul_mi<T, INV>(v1 - v3);
  v0 = a0 + a2;
  v2 = a0 - a2;
  v1 = a1 + a3;
  v3 = a1 - a3;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cplx<T>* v) {
  // even/odd split: E = DFT4(v0,v2,v4,v6), O = DFT4(v1,v3,v5,v7)
  dft4<T, INV>(v[0], v[2], v[4], v[6]);
  dft4<T, INV>(v[1], v[3], v[5], v[7]);
  // twiddle odd part by W8^k, W8 = exp(-+ i pi/4)
  cplx<T> o1 = v[3], o2 = v[5], o3 = v[7];
  auto r = [](T x) __attribute__((always_inline)) { return mul_rsqrt2<T>(x); };
  // W8^1 = (1 -+ i)/sqrt2
  o1 = INV ? mk(r(o1.x - o1.y), r(o1.x + o1.y)) : mk(r(o1.x + o1.y), r(o1.y - o1.x));
  o2 = mul_mi<T, INV>(o2);
  // W8^3 = (-1 -+ i)/sqrt2
  o3 = INV ? mk(r(-o3.x - o3.y), r(o3.x - o3.y)) : mk(r(o3.y - o3.x), r(-o3.x - o3.y));
  cplx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = e0 + o0;
  v[4] = e0 - o0;
  v[1] = e1 + o1;
  v[5] = e1 - o1;
  v[2] = e2 + o2;
  v[6] = e2 - o2;
  v[3] = e3 + o3;
  v[7] = e3 - o3;
}

// Radix 3 (w = exp(-+2 pi i/3)): y0 = v0 + (v1 + v2), y1,2 = v0 - (v1 + v2)/2 -+ i sin(pi/3) (v1 - v2);
// every product by a constant and each output formed in FP64 with one rounding.
template <typename T, bool INV>
__device__ __forceinline__ void dft3(cplx<T>* v) {
  constexpr double S3 = 0.86602540378443864676;  // sin(pi/3)
  const double ax = (double)v[1].x + (double)v[2].x, ay = (double)v[1].y + (double)v[2].y;
  const double bx = ((double)v[1].x - (double)v[2].x) * S3, by = ((double)v[1].y - (double)v[2].y) * S3;
  const double mx = (double)v[0].x - 0.5 * ax, my = (double)v[0].y - 0.5 * ay;
  // -i b (forward) or +i b (inverse)
  const double tx = INV ? -by : by, ty = INV ? bx : -bx;
  v[0] = mk((T)((double)v[0].x + ax), (T)((double)v[0].y + ay));
  v[1] = mk((T)(mx + tx), (T)(my + ty));
  v[2] = mk((T)(mx - tx), (T)(my - ty));
}

// Radix 5 (w = exp(-+2 pi i/5)), with a_k = v_k + v_{5-k}, b_k = v_k - v_{5-k}:
// y_{1,4} = v0 + c1 a1 + c2 a2 -+ i (s1 b1 + s2 b2), y_{2,3} = v0 + c2 a1 + c1 a2 -+ i (s2 b1 - s1 b2)
// (c_k = cos(2 pi k/5), s_k = sin(2 pi k/5)); FP64 inside, one rounding per output.
template <typename T, bool INV>
__device__ __forceinline__ void dft5(cplx<T>* v) {
  constexpr double C1 = 0.30901699437494742410, C2 = -0.80901699437494742410;
  constexpr double S1 = 0.95105651629515357212, S2 = 0.58778525229247312917;
  const double x0 = v[0].x, y0 = v[0].y;
  const double a1x = (double)v[1].x + (double)v[4].x, a1y = (double)v[1].y + (double)v[4].y;
  const double a2x = (double)v[2].x + (double)v[3].x, a2y = (double)v[2].y + (double)v[3].y;
  const double b1x = (double)v[1].x - (double)v[4].x, b1y = (double)v[1].y - (double)v[4].y;
  const double b2x = (double)v[2].x - (double)v[3].x, b2y = (double)v[2].y - (double)v[3].y;
  const double m1x = x0 + C1 * a1x + C2 * a2x, m1y = y0 + C1 * a1y + C2 * a2y;
  const double m2x = x0 + C2 * a1x + C1 * a2x, m2y = y0 + C2 * a1y + C1 * a2y;
  const double n1x = S1 * b1x + S2 * b2x, n1y = S1 * b1y + S2 * b2y;
  const double n2x = S2 * b1x - S1 * b2x, n2y = S2 * b1y - S1 * b2y;
  // -i n (forward) or +i n (inverse)
  const double t1x = INV ? -n1y : n1y, t1y = INV ? n1x : -n1x;
  const double t2x = INV ? -n2y : n2y, t2y = INV ? n2x : -n2x;
  v[0] = mk((T)(x0 + a1x + a2x), (T)(y0 + a1y + a2y));
  v[1] = mk((T)(m1x + t1x), (T)(m1y + t1y));
  v[4] = mk((T)(m1x - t1x), (T)(m1y - t1y));
  v[2] = mk((T)(m2x + t2x), (T)(m2y + t2y));
  v[3] = mk((T)(m2x - t2x), (T)(m2y - t2y));
}

// Radix 7, the same symmetric form: a_k = v_k + v_{7-k}, b_k = v_k - v_{7-k} (k = 1..3),
// y_j = v0 + sum_k cos(2 pi jk/7) a_k -+ i sum_k sin(2 pi jk/7) b_k, y_{7-j} with the other
// sign (jk reduced mod 7: the cosines are C1..C3 permuted, the sines +-S1..S3).
template <typename T, bool INV>
__device__ __forceinline__ void dft7(cplx<T>* v) {
  constexpr double C1 = 0.62348980185873353053, C2 = -0.22252093395631440429,
                   C3 = -0.90096886790241912624;
  constexpr double S1 = 0.78183148246802980871, S2 = 0.97492791218182360702,
                   S3 = 0.43388373911755812048;
  const double x0 = v[0].x, y0 = v[0].y;
  double ax[4], ay[4], bx[4], by[4];
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    ax[k] = (double)v[k].x + (double)v[7 - k].x;
    ay[k] = (double)v[k].y + (double)v[7 - k].y;
    bx[k] = (double)v[k].x - (double)v[7 - k].x;
    by[k] = (double)v[k].y - (double)v[7 - k].y;
  }
  constexpr double cc[3][3] = {{C1, C2, C3}, {C2, C3, C1}, {C3, C1, C2}};
  constexpr double ss[3][3] = {{S1, S2, S3}, {S2, -S3, -S1}, {S3, -S1, S2}};
  cplx<T> out[7];
  out[0] = mk((T)(x0 + ax[1] + ax[2] + ax[3]), (T)(y0 + ay[1] + ay[2] + ay[3]));
#pragma unroll
  for (int j = 1; j <= 3; ++j) {
    const double mx = x0 + cc[j - 1][0] * ax[1] + cc[j - 1][1] * ax[2] + cc[j - 1][2] * ax[3];
    const double my = y0 + cc[j - 1][0] * ay[1] + cc[j - 1][1] * ay[2] + cc[j - 1][2] * ay[3];
    const double nx = ss[j - 1][0] * bx[1] + ss[j - 1][1] * bx[2] + ss[j - 1][2] * bx[3];
    const double ny = ss[j - 1][0] * by[1] + ss[j - 1][1] * by[2] + ss[j - 1][2] * by[3];
    // -i n (forward) or +i n (inverse)
    const double tx = INV ? -ny : ny, ty = INV ? nx : -nx;
    out[j] = mk((T)(mx + tx), (T)(my + ty));
    out[7 - j] = mk((T)(mx - tx), (T)(my - ty));
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = out[k];
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dftR(cplx<T>* v) {
  if constexpr (R == 2) dft2<T, INV>(v[0], v[1]);
  else if constexpr (R == 3) dft3<T, INV>(v);
  else if constexpr (R == 4) dft4<T, INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5) dft5<T, INV>(v);
  else if constexpr (R == 7) dft7<T, INV>(v);
  else dft8<T, INV>(v);
}

// Radix plan for a length-N line: radices (2, 3, 4, 5, 7 or 8) whose product is N, and E,
// the values each thread holds (N/E threads per line); every radix divides E.  Lines of
// 2^k points use radix 8 first and E = 8 (or 16 where a caller asks for it); lines of
// 3 * 2^k, 5 * 2^k and 7 * 2^k points (24 ... 384) use radix 4 passes with one radix-3, -5
// or -7 pass and E = 12, 20 or 28.
template <int N>
struct RadixPlan;
template <> struct RadixPlan<8>   { static constexpr int np = 1, e = 8;  static constexpr int r[5] = {8}; };
template <> struct RadixPlan<16>  { static constexpr int np = 2, e = 8;  static constexpr int r[5] = {8, 2}; };
template <> struct RadixPlan<32>  { static constexpr int np = 2, e = 8;  static constexpr int r[5] = {8, 4}; };
template <> struct RadixPlan<64>  { static constexpr int np = 2, e = 8;  static constexpr int r[5] = {8, 8}; };
template <> struct RadixPlan<128> { static constexpr int np = 3, e = 8;  static constexpr int r[5] = {8, 4, 4}; };
template <> struct RadixPlan<256> { static constexpr int np = 3, e = 8;  static constexpr int r[5] = {8, 8, 4}; };
template <> struct RadixPlan<512> { static constexpr int np = 3, e = 8;  static constexpr int r[5] = {8, 8, 8}; };
template <> struct RadixPlan<24>  { static constexpr int np = 3, e = 12; static constexpr int r[5] = {4, 2, 3}; };
template <> struct RadixPlan<48>  { static constexpr int np = 3, e = 12; static constexpr int r[5] = {4, 4, 3}; };
template <> struct RadixPlan<96>  { static constexpr int np = 4, e = 12; static constexpr int r[5] = {4, 4, 2, 3}; };
template <> struct RadixPlan<192> { static constexpr int np = 4, e = 12; static constexpr int r[5] = {4, 4, 4, 3}; };
template <> struct RadixPlan<384> { static constexpr int np = 5, e = 12; static constexpr int r[5] = {4, 4, 4, 2, 3}; };
template <> struct RadixPlan<40>  { static constexpr int np = 3, e = 20; static constexpr int r[5] = {4, 2, 5}; };
template <> struct RadixPlan<80>  { static constexpr int np = 3, e = 20; static constexpr int r[5] = {4, 4, 5}; };
template <> struct RadixPlan<160> { static constexpr int np = 4, e = 20; static constexpr int r[5] = {4, 4, 2, 5}; };
template <> struct RadixPlan<320> { static constexpr int np = 4, e = 20; static constexpr int r[5] = {4, 4, 4, 5}; };
template <> struct RadixPlan<56>  { static constexpr int np = 3, e = 28; static constexpr int r[5] = {4, 2, 7}; };
template <> struct RadixPlan<112> { static constexpr int np = 3, e = 28; static constexpr int r[5] = {4, 4, 7}; };
template <> struct RadixPlan<224> { static constexpr int np = 4, e = 28; static constexpr int r[5] = {4, 4, 2, 7}; };

// Values per thread of a length-N line FFT: the plan's E, or E2 (a multiple of 8) for a
// power-of-two line whose caller holds more values per thread.
template <int N, int E2 = 8>
constexpr int fft_e() {
  return RadixPlan<N>::e == 8 ? E2 : RadixPlan<N>::e;
}

// One Stockham pass: butterflies j = t + b*(N/E), b < E/R.
// in[j + r*N/R] -> twiddle W_N^{(j mod Ns) r N/(Ns R)} -> DFT_R -> out[(j/Ns)Ns R + j mod Ns + r Ns]
// The line lives at base[e*ES]; tw[m] = exp(-2 pi i m / N) in FP64 for every T.
// Index of element i of a line: i * ES, or with PAD (unit-stride lines) i + i/8, one pad
// slot per 8 elements so that the stride-R writes of the first pass hit distinct LDS banks.
template <bool PAD>
__device__ __forceinline__ int lds_at(int i, int ES) {
  return PAD ? i + (i >> 3) : i * ES;
}

template <typename T, int N, int R, bool INV, int E = 8, bool PAD = false>
__device__ __forceinline__ void stockham_pass(cplx<T>* base, int ES, int t, bool active, int Ns,
                                              const cplx<double>* __restrict__ tw) {
  constexpr int NB = E / R;      // butterflies per thread
  constexpr int TEAM = N / E;
  cplx<T> v[E];
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j = t + b * TEAM;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b * R + r] = base[lds_at<PAD>(j + r * (N / R), ES)];
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j = t + b * TEAM;
      int k = j % Ns;
      if (Ns > 1) {
        // W_N^{st r} for r = 1..R-1 straight from the table (st r < N): the callers keep
        // the table in LDS, so this costs LDS reads instead of FP64 products, and every
        // twiddle is the correctly rounded value rather than an accumulated power
        const int st = k * (N / (Ns * R));
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<double> w = tw[st * r];
          if (INV) w.y = -w.y;
          v[b * R + r] = mul64(v[b * R + r], w);
        }
      }
      dftR<T, R, INV>(&v[b * R]);
      int o = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) base[lds_at<PAD>(o + r * Ns, ES)] = v[b * R + r];
    }
  }
  __syncthreads();
}

// Full length-N FFT (unnormalised) of the line at base[e*ES], in place in LDS.
// Every thread of the block must call it (it synchronises); `active` marks the
// threads that own a (line, team index) pair.
template <typename T, int N, bool INV, int E, bool PAD, int I>
__device__ __forceinline__ void fft_passes(cplx<T>* base, int ES, int t, bool active, int Ns,
                                           const cplx<double>* __restrict__ tw) {
  using P = RadixPlan<N>;
  if constexpr (I < P::np) {
    static_assert(E % P::r[I] == 0, "every radix of the plan divides the values per thread");
    stockham_pass<T, N, P::r[I], INV, E, PAD>(base, ES, t, active, Ns, tw);
    fft_passes<T, N, INV, E, PAD, I + 1>(base, ES, t, active, Ns * P::r[I], tw);
  }
}

template <typename T, int N, bool INV, int E = 8, bool PAD = false>
__device__ __forceinline__ void fft_team(cplx<T>* base, int ES, int t, bool active,
                                         const cplx<double>* __restrict__ tw) {
  static_assert(N % E == 0, "line shorter than the per-thread register block");
  fft_passes<T, N, INV, E, PAD, 0>(base, ES, t, active, 1, tw);
}

// ---------------------------------------------------------------------------
// Per-site spinor propagators
// ---------------------------------------------------------------------------

// propagators.jl:277-314: n from the pre-update spinor; ITP uses real
// exponentials with the Zeeman shift min_c E_c, real time uses cis.
// Common factor of all components: cis(-(V + c0 n + c_lhy n^1.5) dt) or its ITP exponential.
__device__ __forceinline__ cplx<double> diag_base(double nd, double Vtrap, const InstParams& P,
                                                  double dtf, bool itp) {
#pragma clang fp contract(on)
  // without LHY (c_lhy = 0, uniform over the block) the term is fma(0, sqrt(n), x) = x, so
  // skipping the FP64 square root gives the same bits
  const double a = P.c_lhy != 0.0 ? (Vtrap + P.c0 * nd + P.c_lhy * nd * sqrt(nd)) * dtf
                                  : (Vtrap + P.c0 * nd) * dtf;
  if (itp) return mk(exp(-a), 0.0);
  double sn, cs;
  sincos_d(a, &sn, &cs);
  return mk(cs, -sn);
}

template <typename T, int D>
__device__ __forceinline__ void op_diag(cplx<T>* s, double Vtrap, const InstParams& P, double dtf,
                                        bool itp, const cplx<double>* zf) {
  T n = 0;
#pragma unroll
  for (int c = 0; c < D; ++c) n = density_acc(n, s[c]);
  const cplx<double> base = diag_base((double)n, Vtrap, P, dtf, itp);
#pragma unroll
  for (int c = 0; c < D; ++c) s[c] = mul_split<T>(s[c], cmul_fixed(base, zf[c]));
}

// observables.jl:68-88 spin density: Fz = sum m|psi|^2,
// Fx + iFy = sum_{c>=1} f+(m_c) conj(psi_{c-1}) psi_c.
template <typename T, int D>
__device__ __forceinline__ void spin_vec(const cplx<T>* s, T& fx, T& fy, T& fz) {
  constexpr int F = (D - 1) / 2;
  fz = 0;
  fx = 0;
  fy = 0;
#pragma unroll
  for (int c = 0; c < D; ++c) fz += (T)(F - c) * abs2(s[c]);
#pragma unroll
  for (int c = 1; c < D; ++c) {
    const T m = (T)(F - c);
    const T fp = (T)sqrt((double)(F * (F + 1)) - (double)m * ((double)m + 1.0));
    cplx<T> pr = conjc(s[c - 1]) * s[c];
    fx += fp * pr.x;
    fy += fp * pr.y;
  }
}

// spin_mixing.jl:71-98: U = I - i sin(th) nF + (cos th - 1) nF^2 (ITP: sinh/cosh, no i),
// applied as U s = s - i sin(th) (nF s) + (cos th - 1) nF (nF s) with the spin-1 matrices
// written out (nF = [[c, u*, 0], [u, 0, u*], [0, u, -c]], u = (a + i b)/sqrt2).
template <typename T>
__device__ __forceinline__ void op_rodrigues(cplx<T>* s, double c1, double dtf, bool itp) {
  T fx, fy, fz;
  spin_vec<T, 3>(s, fx, fy, fz);
  T fmag = sqrt(fx * fx + fy * fy + fz * fz);
  if (!(fmag >= (T)1e-30)) return;
  const T is2 = (T)0.70710678118654752440;
  const T inv = (T)1 / fmag;  // one division for the three components of the axis
  const T c = fz * inv;
  const cplx<T> u = mk(fx * inv * is2, fy * inv * is2), uc = conjc(u);
  auto nF = [&](const cplx<T>* v, cplx<T>* o) __attribute__((always_inline)) {
    o[0] = v[0] * c + uc * v[1];
    o[1] = u * v[0] + uc * v[2];
    o[2] = u * v[1] - v[2] * c;
  };
  cplx<T> a[3], b[3];
  nF(s, a);
  nF(a, b);
  const double th = c1 * (double)fmag * dtf;
  double s1, c1m;
  if (itp) {
    s1 = sinh(th);
    c1m = cosh(th) - 1.0;
  } else {
    double sn, cs;
    sincos_d(th, &sn, &cs);
    s1 = sn;
    c1m = cs - 1.0;
  }
  // combined in FP64 with one rounding per component (FP32-bulk: no systematic bias)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    // - i s1 a  (real time)   or   - s1 a  (imaginary time)
    const double tx = itp ? -(double)a[i].x * s1 : (double)a[i].y * s1;
    const double ty = itp ? -(double)a[i].y * s1 : -(double)a[i].x * s1;
    s[i] = mk((T)((double)s[i].x + tx + (double)b[i].x * c1m),
              (T)((double)s[i].y + ty + (double)b[i].y * c1m));
  }
}

template <int F>
inline constexpr tens::DHalf<F> kDHalf{};

template <int F, bool TRANS>
inline constexpr tens::DHalfParity<F, TRANS> kDHalfPar{};

// w = d v (TRANS false) or w = d^T v with d = d^F(pi/2): a real matrix with compile-time
// entries.  An even row takes the F + 1 sums v_c + v_{D-1-c} (c <= F), an odd row the F
// differences v_c - v_{D-1-c} (c < F): about half the products of the plain row sums, each
// an FMA pair on literals (zeros skipped).
template <typename T, int D, bool TRANS>
__device__ __forceinline__ void dhalf_apply(const cplx<T>* v, cplx<T>* w) {
  constexpr int F = (D - 1) / 2;
  constexpr int NP = F > 0 ? F : 1;
  cplx<T> ps[NP], ms[NP];
  static_for<0, F>([&](auto jc) __attribute__((always_inline)) {
    constexpr int j = decltype(jc)::value;
    ps[j] = v[j] + v[D - 1 - j];
    ms[j] = v[j] - v[D - 1 - j];
  });
  static_for<0, D>([&](auto ic) __attribute__((always_inline)) {
    constexpr int i = decltype(ic)::value;
    constexpr int par = kDHalfPar<F, TRANS>.s[i];
    cplx<T> acc = mk((T)0, (T)0);
    if constexpr (par == 0) {
      static_for<0, D>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        constexpr T c = (T)kDHalf<F>.v[TRANS ? j * D + i : i * D + j];
        if constexpr (c != (T)0) acc = mk(acc.x + c * v[j].x, acc.y + c * v[j].y);
      });
    } else {
      if constexpr (par == 1) {
        constexpr T c = (T)kDHalf<F>.v[TRANS ? F * D + i : i * D + F];
        if constexpr (c != (T)0) acc = mk(acc.x + c * v[F].x, acc.y + c * v[F].y);
      }
      static_for<0, F>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        constexpr T c = (T)kDHalf<F>.v[TRANS ? j * D + i : i * D + j];
        const cplx<T> u = par == 1 ? ps[j] : ms[j];
        if constexpr (c != (T)0) acc = mk(acc.x + c * u.x, acc.y + c * u.y);
      });
    }
    w[i] = acc;
  });
}

// v_c *= e^{i m x} for m = F - c (Rz(-x) of the reference's convention): the factors of m and
// -m are conjugates, so the F powers e^{i m x}, m = 1..F, come from one recurrence on e.
// The products are explicit FMAs, so that every kernel that inlines the rotation (the fused
// x/site pass and the lane-per-site pass) rounds them alike.
template <typename T, int D>
__device__ __forceinline__ void phase_sweep(cplx<T>* v, cplx<T> e) {
  constexpr int F = (D - 1) / 2;
  auto fm = [](T x, T y, T z) __attribute__((always_inline)) {
    if constexpr (sizeof(T) == 4) return fmaf(x, y, z);
    else return fma(x, y, z);
  };
  auto cm = [&](cplx<T> a, cplx<T> b) __attribute__((always_inline)) {
    return mk<T>(fm(a.x, b.x, -(a.y * b.y)), fm(a.x, b.y, a.y * b.x));
  };
  cplx<T> p = e;
#pragma unroll
  for (int m = 1; m <= F; ++m) {
    v[F - m] = cm(p, v[F - m]);
    v[F + m] = cm(conjc(p), v[F + m]);
    if (m < F) p = cm(p, e);
  }
}

// spinor_utils.jl:82-166: exp(-i dt phi.F) = Rz(alpha) Ry(beta) Dz(theta) Ry(-beta) Rz(-alpha),
// beta = acos(phi_z/|phi|), alpha = atan2(phi_y, phi_x), theta = |phi| dt; ITP: Dz = exp(-m theta).
// The reference forms Ry(+-beta) from a numerical eigenbasis V_Fy of F_y; here the eigenbasis
// is the closed form U = Rz(pi/2) d^F(pi/2), so that, with a = alpha + pi/2,
//   R = Rz(a) d Rz(beta) d^T Dz(theta) d Rz(-beta) d^T Rz(-a),   Rz(x) = diag(e^{-i x m}),
// four real D x D matvecs and five phase sweeps (cis recurrences as in the reference, one
// power per |m|).
// Returns 1 if the rotation was applied (0: |phi| |dt| < 1e-14, identity).
template <typename T, int D>
__device__ __forceinline__ int op_euler(cplx<T>* s, double phx, double phy, double phz, double dt,
                                        bool itp) {
  constexpr int F = (D - 1) / 2;
  const double pm = sqrt(phx * phx + phy * phy + phz * phz);
  if (pm * fabs(dt) < 1e-14) return 0;
  const double n0 = (sizeof(T) == 4 && !itp) ? site_norm2<T, D>(s) : 0.0;
  // cos/sin of beta and alpha from phi itself (no acos/atan2): cos beta = phi_z/|phi| (clamped),
  // sin beta = rho/|phi| >= 0, (cos alpha, sin alpha) = (phi_x, phi_y)/rho, with atan2's
  // signed-zero convention when rho = 0.
  const double rho = sqrt(phx * phx + phy * phy);
  const double ipm = 1.0 / pm;  // one division per radius
  double cb = phz * ipm;
  cb = cb < -1.0 ? -1.0 : (cb > 1.0 ? 1.0 : cb);
  const double sb = rho * ipm;
  double cal, sal;
  if (rho > 0.0) {
    const double irho = 1.0 / rho;
    cal = phx * irho;
    sal = phy * irho;
  } else {
    cal = signbit(phx) ? -1.0 : 1.0;
    sal = 0.0;
  }
  const double ca = -sal, sa = cal;  // a = alpha + pi/2
  const double theta = pm * dt;
  cplx<T> v[D], w[D];
#pragma unroll
  for (int c = 0; c < D; ++c) v[c] = s[c];
  phase_sweep<T, D>(v, mk((T)ca, (T)sa));  // Rz(-a): e^{+i a m}
  dhalf_apply<T, D, true>(v, w);
  phase_sweep<T, D>(w, mk((T)cb, (T)sb));  // Rz(-beta)
  dhalf_apply<T, D, false>(w, v);
  if (itp) {
    const double et = exp(theta);
    double eF = 1.0;
#pragma unroll
    for (int k = 0; k < F; ++k) eF *= et;
    T dz = (T)(1.0 / eF);
    const T st = (T)et;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      v[c] = v[c] * dz;
      dz *= st;
    }
  } else {
    double st, ct;
    sincos_d(theta, &st, &ct);
    phase_sweep<T, D>(v, mk((T)ct, (T)(-st)));  // e^{-i theta m}
  }
  dhalf_apply<T, D, true>(v, w);
  phase_sweep<T, D>(w, mk((T)cb, (T)(-sb)));  // Rz(beta)
  dhalf_apply<T, D, false>(w, v);
  phase_sweep<T, D>(v, mk((T)ca, (T)(-sa)));  // Rz(a)
#pragma unroll
  for (int c = 0; c < D; ++c) s[c] = v[c];
  if (!itp) restore_norm<T, D>(s, n0);
  return 1;
}

// Returns the number of Euler rotations applied (0 or 1; Rodrigues steps count 0).
template <typename T, int D>
__device__ __forceinline__ int op_spin_mixing(cplx<T>* s, const InstParams& P, double dtf, bool itp) {
  if (fabs(P.c1) < 1e-30) return 0;  // spin_mixing.jl:9
  if constexpr (D == 3) {
    op_rodrigues<T>(s, P.c1, dtf, itp);
    return 0;
  } else {
    T fx, fy, fz;
    spin_vec<T, D>(s, fx, fy, fz);
    return op_euler<T, D>(s, P.c1 * (double)fx, P.c1 * (double)fy, P.c1 * (double)fz, dtf, itp);
  }
}


// losses.jl:25-49: psi_c *= exp(-rate_c n dt / 2) with n from the pre-update spinor;
// components with rate_c < 1e-30 are left unchanged.
__device__ __forceinline__ double loss_factor(double rate, double n, double dtf) {
  return exp(-rate * n * dtf / 2.0);
}
template <typename T, int D>
__device__ __forceinline__ void op_loss(cplx<T>* s, const InstParams& P, double dtf) {
  if (!P.has_loss) return;
  double n = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) n += (double)density_acc<T>(0, s[c]);
#pragma unroll
  for (int c = 0; c < D; ++c)
    if (!(P.loss[c] < 1e-30)) s[c] = scale64(s[c], loss_factor(P.loss[c], n, dtf));
}

// nematic.jl:34-119: A00 = sum_c (-1)^c psi_c psi_{D-1-c} / sqrt(D) from the pre-update
// spinor (the sign (-1)^(F-m) is (-1)^c), V = c2 (-1)^c A00 / sqrt(D) per (c, D-1-c) pair:
// real time psi_c <- cos(|V|dt) psi_c - i (V/|V|) sin(|V|dt) conj(psi_pair),
// ITP       psi_c <- cosh(|V|dt) psi_c - (V/|V|) sinh(|V|dt) conj(psi_pair); |V| < 1e-30 skips.
template <typename T, int D>
__device__ __forceinline__ void op_nematic(cplx<T>* s, double c2, double dtf, bool itp) {
  if (fabs(c2) < 1e-30) return;  // nematic.jl:43
  const double isd = 1.0 / sqrt((double)D);
  double ar = 0.0, ai = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const double sg = (c % 2 == 0) ? 1.0 : -1.0;
    const double xr = s[c].x, xi = s[c].y, yr = s[D - 1 - c].x, yi = s[D - 1 - c].y;
    ar += sg * (xr * yr - xi * yi);
    ai += sg * (xr * yi + xi * yr);
  }
  const double vr0 = c2 * ar * isd * isd, vi0 = c2 * ai * isd * isd;
  const double av = sqrt(vr0 * vr0 + vi0 * vi0);  // |V| is the same for every pair
  if (av < 1e-30) return;
  double cc, ss;
  if (itp) {
    cc = cosh(av * dtf);
    ss = sinh(av * dtf);
  } else {
    sincos_d(av * dtf, &ss, &cc);
  }
#pragma unroll
  for (int c = 0; c < (D + 1) / 2; ++c) {
    const int cp = D - 1 - c;
    const double sg = (c % 2 == 0) ? 1.0 : -1.0;
    const double pr = sg * vr0 / av, pi = sg * vi0 / av;  // V/|V|
    // coef = ph * sh (ITP) or i * ph * sin (real time)
    const double kr = itp ? pr * ss : -pi * ss, ki = itp ? pi * ss : pr * ss;
    const double mr = s[c].x, mi = s[c].y, nr = s[cp].x, ni = s[cp].y;
    // x <- cc x - coef conj(y)
    auto upd = [&](double xr, double xi, double yr, double yi) __attribute__((always_inline)) {
      return mk((T)(cc * xr - (kr * yr + ki * yi)), (T)(cc * xi - (ki * yr - kr * yi)));
    };
    if (c == cp) {
      s[c] = upd(mr, mi, mr, mi);
    } else {
      s[c] = upd(mr, mi, nr, ni);
      s[cp] = upd(nr, ni, mr, mi);
    }
  }
}

// ---------------------------------------------------------------------------
// General-F tensor interaction (tensor_interaction.jl:130-240)
// ---------------------------------------------------------------------------
template <int B, int E, int S = 1, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  if constexpr (B < E) {
    fn(std::integral_constant<int, B>{});
    static_for<B + S, E, S>(fn);
  }
}

// Clebsch-Gordan table of spin F in tens:: order, evaluated by the compiler.
template <int F>
inline constexpr tens::CGTab<F> kCG{};
// The same values in constant memory.  VALU literals are 32 bits, so an FP64 CG value used
// as an instruction operand costs two s_mov_b32 per use; read from here, the compiler
// fetches eight values per s_load_dwordx16 instead.
template <int F>
__constant__ tens::CGTab<F> kCGConst = tens::CGTab<F>{};

// CG value at table index I as an operand of type T (FP32: literal; FP64: loaded through
// cgp, a pointer to kCGConst<F> that the caller makes opaque on every matvec, so the loads
// stay next to their use instead of being hoisted out of the Taylor loop and spilled).
template <typename T, int F, int I>
__device__ __forceinline__ T cg_operand(const_f64_ptr cgp) {
  if constexpr (sizeof(T) == 8) return cgp[I];
  else return (T)kCG<F>.v[I];
}

// w = h v with the mean-field matrix of the spinor psi,
//   h_{m m'} = sum_S g_S sum_M CG(m, M-m | S M) CG(m', M-m' | S M) conj(psi_{M-m}) psi_{M-m'},
// evaluated in factorised form (no D x D matrix is formed):
//   B_SM = sum_{m'} CG(m', M-m') psi_{M-m'} v_{m'},  (h v)_m = sum_{S,M} g_S CG(m, M-m) conj(psi_{M-m}) B_SM,
// using CG(m1, m2) = CG(m2, m1) for even S to visit each unordered pair once.
// cgt: table in tens:: order; every index below is a compile-time constant.
template <typename T, int D, bool SYM = false>
__device__ __forceinline__ void tensor_matvec(const cplx<T>* psi, const cplx<T>* v, cplx<T>* w,
                                              const T* g, const_f64_ptr cgp) {
  constexpr int F = (D - 1) / 2;
#pragma unroll
  for (int c = 0; c < D; ++c) w[c] = mk((T)0, (T)0);
  static_for<-2 * F, 2 * F + 1>([&](auto Mc) __attribute__((always_inline)) {
    constexpr int M = decltype(Mc)::value;
    constexpr int S0 = tens::s0(M);
    constexpr int NS = tens::ns(F, M);
    constexpr int LO = tens::m1lo(F, M);
    constexpr int HI = tens::m1hi(M);
    cplx<T> Bs[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) Bs[k] = mk((T)0, (T)0);
    static_for<LO, HI + 1>([&](auto m1c) __attribute__((always_inline)) {
      constexpr int m1 = decltype(m1c)::value;
      constexpr int m2 = M - m1;
      cplx<T> p = psi[F - m2] * v[F - m1];
      if constexpr (m1 != m2) {
        if constexpr (SYM) p = p + p;  // v = psi
        else p = p + psi[F - m1] * v[F - m2];
      }
      static_for<0, NS>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        constexpr int I = tens::off(F, M, m1, S0) + k;
        if constexpr ((T)kCG<F>.v[I] != (T)0) Bs[k] = Bs[k] + p * cg_operand<T, F, I>(cgp);
      });
    });
#pragma unroll
    for (int k = 0; k < NS; ++k) Bs[k] = Bs[k] * g[S0 / 2 + k];
    static_for<LO, HI + 1>([&](auto m1c) __attribute__((always_inline)) {
      constexpr int m1 = decltype(m1c)::value;
      constexpr int m2 = M - m1;
      cplx<T> C = mk((T)0, (T)0);
      static_for<0, NS>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        constexpr int I = tens::off(F, M, m1, S0) + k;
        if constexpr ((T)kCG<F>.v[I] != (T)0) C = C + Bs[k] * cg_operand<T, F, I>(cgp);
      });
      w[F - m1] = w[F - m1] + conjc(psi[F - m2]) * C;
      if constexpr (m1 != m2) w[F - m2] = w[F - m2] + conjc(psi[F - m1]) * C;
    });
  });
}

// FP64 variant of tensor_matvec with the channels folded per instance:
//   P_i = sum over the pair i = (m1, m2) of psi_{m2} v_{m1} (+ psi_{m1} v_{m2}, m1 != m2),
//   C_i = sum_j K_M[i][j] P_j,   (h v)_{m1} += conj(psi_{m2}) C_i (and m1 <-> m2),
// which is the factorised form with K_M = CG_M^T diag(g) CG_M (4.3 instead of 6.1 kflop
// per matvec at F = 6).  Its operands come from memory in both precisions' cost model for
// FP64 (no 64-bit literals), so fewer of them pay off; kp points at this instance's
// kernel table in constant memory, made opaque by the caller per matvec.
template <typename T, int D, bool SYM = false>
__device__ __forceinline__ void tensor_matvec_folded(const cplx<T>* psi, const cplx<T>* v,
                                                     cplx<T>* w, const_f64_ptr kp) {
  constexpr int F = (D - 1) / 2;
  // Every product is written as a chain of fused multiply-adds (a complex multiply-add is
  // four FMAs): 2.2k VALU instructions per matvec at F = 6 instead of 2.8k with the
  // separate multiplies and adds that the complex operators leave after contraction.
#pragma unroll
  for (int c = 0; c < D; ++c) w[c] = mk((T)0, (T)0);
  static_for<-2 * F, 2 * F + 1>([&](auto Mc) __attribute__((always_inline)) {
    constexpr int M = decltype(Mc)::value;
    constexpr int LO = tens::m1lo(F, M);
    constexpr int NP = tens::npairs(F, M);
    constexpr int KB = tens::kfold_off(F, M);
    cplx<T> Pp[NP];
    static_for<0, NP>([&](auto ic) __attribute__((always_inline)) {
      constexpr int m1 = LO + decltype(ic)::value;
      constexpr int m2 = M - m1;
      const cplx<T> a = psi[F - m2], b = v[F - m1];
      T pr = fma(a.x, b.x, -(a.y * b.y));
      T pi = fma(a.x, b.y, a.y * b.x);
      if constexpr (m1 != m2) {
        if constexpr (SYM) {  // v = psi: psi_m1 v_m2 = psi_m2 v_m1
          pr = pr + pr;
          pi = pi + pi;
        } else {
          const cplx<T> c = psi[F - m1], d = v[F - m2];
          pr = fma(c.x, d.x, pr);
          pr = fma(-c.y, d.y, pr);
          pi = fma(c.x, d.y, pi);
          pi = fma(c.y, d.x, pi);
        }
      }
      Pp[decltype(ic)::value] = mk(pr, pi);
    });
    static_for<0, NP>([&](auto ic) __attribute__((always_inline)) {
      constexpr int i = decltype(ic)::value;
      constexpr int m1 = LO + i;
      constexpr int m2 = M - m1;
      const T k0 = (T)kp[KB + i * NP];
      T cr = Pp[0].x * k0, ci = Pp[0].y * k0;
      static_for<1, NP>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        const T kj = (T)kp[KB + i * NP + j];
        cr = fma(Pp[j].x, kj, cr);
        ci = fma(Pp[j].y, kj, ci);
      });
      // w_{m1} += conj(psi_{m2}) C (and m1 <-> m2)
      auto acc = [&](int c, cplx<T> a) __attribute__((always_inline)) {
        T wr = fma(a.x, cr, w[c].x);
        wr = fma(a.y, ci, wr);
        T wi = fma(a.x, ci, w[c].y);
        wi = fma(-a.y, cr, wi);
        w[c] = mk(wr, wi);
      };
      acc(F - m1, psi[F - m2]);
      if constexpr (m1 != m2) acc(F - m2, psi[F - m1]);
    });
  });
}

template <typename T>
__device__ __forceinline__ T cabs1(cplx<T> a) { return fabs(a.x) + fabs(a.y); }

// Storage of the Taylor sum of op_tensor.  RegAcc keeps it in VGPRs; LdsAcc keeps it in the
// calling thread's own D slots of an LDS tile (slot of component c: at(c)), which takes
// 2D VGPRs off the step's live set (the FP64 spin-6 step then fits 256 VGPRs).  Both give
// the same bits: the update below is written with one rounding per operation.
template <typename T, int D>
struct RegAcc {
  cplx<T> a[D];
  __device__ __forceinline__ cplx<T> get(int c) const { return a[c]; }
  __device__ __forceinline__ void set(int c, cplx<T> v) { a[c] = v; }
};
template <typename T, typename At>
struct LdsAcc {
  cplx<T>* base;
  At at;
  __device__ __forceinline__ cplx<T> get(int c) const { return base[at(c)]; }
  __device__ __forceinline__ void set(int c, cplx<T> v) { base[at(c)] = v; }
};

// psi <- exp(-i h dt) psi (real time) or exp(-h dt) psi (imaginary time), h built from the
// incoming spinor.  The reference diagonalises h (LAPACK eigen); here the action of the
// exponential is summed as a Taylor series to working precision, with the step split into
// nsub pieces when the norm bound theta = 2 |dt| n sum_S |g_S| exceeds 1/2.  Sites with
// n < 1e-20 are left unchanged (tensor_interaction.jl:212).
// op_tensor returns the matvecs it ran, plus 1 << kSymShift when the first of them took the
// symmetric pair sums (every site that runs the series); the profiling counters of the
// calling kernels split the two.
constexpr int kSymShift = 24;
template <typename T, int D, typename Acc>
__device__ __forceinline__ int op_tensor(cplx<T>* s, const InstParams& P, double dtf, bool itp,
                                         Acc& acc, const double* kfold) {
  constexpr int F = (D - 1) / 2;
  double n = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) n += (double)abs2(s[c]);
  if (n < 1e-20) return 0;
  T g[F + 1];
  double gsum = 0.0;
#pragma unroll
  for (int k = 0; k <= F; ++k) {
    g[k] = (T)P.gS[k];
    gsum += fabs(P.gS[k]);
  }
  const double theta = 2.0 * fabs(dtf) * n * gsum;
  int nsub = 1;
  if (theta > 0.5) nsub = (int)ceil(theta / 0.5);
  const double hstep = dtf / nsub;
  // Taylor terms shrink by at least th_sub / (k + 1) <= 1/4 per order (th_sub bounds
  // |h hstep| in the component 1-norm used below, with a factor 2 to spare), so after term k
  // the rest of the series is below 2 |v_k| th_sub / (k + 1): the sum stops once that is
  // under eps |acc| (eps: the unit round-off, or 3e-9 for FP32 psi)
  const double th_sub = theta / nsub;
  const T eps = sizeof(T) == 8 ? (T)1.1102230246251565e-16 : (T)3e-9;
  const double n_in = (sizeof(T) == 4 && !itp) ? site_norm2<T, D>(s) : 0.0;
  cplx<T> psi0[D], v[D], w[D];
  int nmv = 0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    psi0[c] = s[c];
    acc.set(c, s[c]);
  }
  // one Taylor term: v <- (-i or -1) hstep/k w, acc += v; true once the rest is negligible
  auto term = [&](int k) __attribute__((always_inline)) {
    const T f = (T)(hstep / k);
    T nv = 0, na = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
#pragma clang fp contract(off)
      v[c] = itp ? mk(-w[c].x * f, -w[c].y * f) : mk(w[c].y * f, -w[c].x * f);
      const cplx<T> a = acc.get(c) + v[c];
      acc.set(c, a);
      nv += cabs1(v[c]);
      na += cabs1(a);
    }
    return (T)(2.0 * th_sub / (k + 1)) * nv <= eps * na;
  };
  // The first term of the first sub-step has v = psi, whose pair sums are twice one product
  // (peeled before the loop, so the loop keeps one matvec body)
  bool first_done;
  if constexpr (sizeof(T) == 8) {
    const_f64_ptr kp = (const_f64_ptr)kfold;
    asm volatile("" : "+s"(kp));
    tensor_matvec_folded<T, D, true>(psi0, psi0, w, kp);
  } else {
    tensor_matvec<T, D, true>(psi0, psi0, w, g, (const_f64_ptr)kCGConst<F>.v);
  }
  ++nmv;
  first_done = term(1);
  for (int sub = 0; sub < nsub; ++sub) {
    int k = 1;
    if (sub == 0) {
      if (first_done) continue;
      k = 2;  // v holds the first term
    } else {
#pragma unroll
      for (int c = 0; c < D; ++c) v[c] = acc.get(c);
    }
    for (; k <= 60; ++k) {
      if constexpr (sizeof(T) == 8) {
        const_f64_ptr kp = (const_f64_ptr)kfold;
        asm volatile("" : "+s"(kp));
        tensor_matvec_folded<T, D>(psi0, v, w, kp);
      } else {
        tensor_matvec<T, D>(psi0, v, w, g, (const_f64_ptr)kCGConst<F>.v);
      }
      ++nmv;
      if (term(k)) break;
    }
  }
#pragma unroll
  for (int c = 0; c < D; ++c) s[c] = acc.get(c);
  if (!itp) restore_norm<T, D>(s, n_in);
  return nmv + (1 << kSymShift);
}

template <typename T, int D>
__device__ __forceinline__ int op_tensor(cplx<T>* s, const InstParams& P, double dtf, bool itp,
                                         const double* kfold) {
  RegAcc<T, D> acc;
  return op_tensor<T, D>(s, P, dtf, itp, acc, kfold);
}

// Tensor interaction energy density 1/2 sum_S g_S sum_M |A_SM|^2 with the pair amplitude
// A_SM = sum_{m1+m2=M} CG(m1, m2 | S M) psi_m1 psi_m2 (tensor_interaction.jl:247-276).
template <typename T, int D>
__device__ __forceinline__ double tensor_energy_site(const cplx<T>* s, const InstParams& P) {
  constexpr int F = (D - 1) / 2;
  double e = 0.0;
  static_for<-2 * F, 2 * F + 1>([&](auto Mc) __attribute__((always_inline)) {
    constexpr int M = decltype(Mc)::value;
    constexpr int S0 = tens::s0(M);
    constexpr int NS = tens::ns(F, M);
    constexpr int LO = tens::m1lo(F, M);
    constexpr int HI = tens::m1hi(M);
    double ar[NS], ai[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) ar[k] = ai[k] = 0.0;
    static_for<LO, HI + 1>([&](auto m1c) __attribute__((always_inline)) {
      constexpr int m1 = decltype(m1c)::value;
      constexpr int m2 = M - m1;
      const cplx<T> a = s[F - m1], b = s[F - m2];
      double pr = (double)a.x * b.x - (double)a.y * b.y;
      double pi = (double)a.x * b.y + (double)a.y * b.x;
      if constexpr (m1 != m2) { pr *= 2.0; pi *= 2.0; }
      static_for<0, NS>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        constexpr double c = kCG<F>.v[tens::off(F, M, m1, S0) + k];
        ar[k] += c * pr;
        ai[k] += c * pi;
      });
    });
#pragma unroll
    for (int k = 0; k < NS; ++k) e += 0.5 * P.gS[S0 / 2 + k] * (ar[k] * ar[k] + ai[k] * ai[k]);
  });
  return e;
}

}  // namespace bec
