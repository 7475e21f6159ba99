// Device-side building blocks of the fused Allen-Cahn explicit step (gfx950).
//
// Physics follows jax_am/phase_field/allen_cahn.py:53-85 (state_rhs), :106-114
// (explicit_euler) and :117-128 (liquid reset), rewritten as one pass per cell:
//
//   eta_new_i = eta_i - c1 * eta_i * (eta_i^2 * B + A) + c2 * a_i * lap_i
//   c1 = dt*L(T)*m_g/ad_hoc,  c2 = dt*L(T)*kappa_g*ad_hoc,  B = 1 - 4*gamma,
//   A  = 4*gamma*S - 1 + 2*(1-zeta)^2,  S = sum_j eta_j^2,
//
// which is d/d eta_i of the reference's local energy (allen_cahn.py:42-51)
// m_g*(eta^3 - eta + 4*gamma*eta*(S - eta^2) + 2*(1-zeta)^2*eta) in closed form.
// a_i is the cubic anisotropy factor (allen_cahn.py:28-39) in the arccos-free form
// a = 1 + eps*(q^2 + (1-q)^2), q = max_k (d_k.e)^2 / |e|^2 clamped to 1 (q = 1 when
// |e| = 0, the reference's NaN -> angle 0 rule).
//
// Per-cell quantities (temperature, liquid indicator zeta, Arrhenius mobility, liquid
// reset) are evaluated in FP64, except that with FP32 arithmetic a cell whose mobility is
// negligible (make_cold) takes an FP32 mobility; per-orientation
// arithmetic is done in the compute type TC (float for F32/BF16 storage, double for F64).
// All device code is built with -ffp-contract=off: every fused multiply-add below is
// written out, so each kernel that shares these functions produces bitwise the same
// result for the same cell.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "pfhip.h"

#ifndef PF_STORE_MODE
#define PF_STORE_MODE 2  // 0: direct per-cell stores, 1: LDS-staged coalesced rows, 2: staged + non-temporal
#endif

namespace pf {

// ------------------------------------------------------------------ storage types
struct bf16_t {
  uint16_t u;
};

template <class TS> struct Traits;
template <> struct Traits<float> {
  using TC = float;
  static constexpr int dtype = PF_F32;
};
template <> struct Traits<double> {
  using TC = double;
  static constexpr int dtype = PF_F64;
};
template <> struct Traits<bf16_t> {
  using TC = float;
  static constexpr int dtype = PF_BF16;
};

template <class T, int N> struct alignas(sizeof(T) * N) Vec {
  T v[N];
};
// the same with a given alignment (global chunks that start on 4- or 8-byte boundaries)
template <class T, int N, int A> struct alignas(A) VecA {
  T v[N];
};

// raw vector type of a 2/4/8/16-byte piece
template <int BYTES> struct RawT;
template <> struct RawT<16> { typedef unsigned int type __attribute__((ext_vector_type(4))); };
// 16-byte piece at a 2-byte-aligned global address (global accesses run in unaligned mode)
typedef unsigned int raw16a2_t __attribute__((ext_vector_type(4), aligned(2)));
template <> struct RawT<8> { typedef unsigned int type __attribute__((ext_vector_type(2))); };
template <> struct RawT<4> { typedef unsigned int type; };
template <> struct RawT<2> { typedef unsigned short type; };

__device__ __forceinline__ float to_tc(float x) { return x; }
__device__ __forceinline__ double to_tc(double x) { return x; }
__device__ __forceinline__ float to_tc(bf16_t x) { return __uint_as_float(((uint32_t)x.u) << 16); }

template <class TS> __device__ __forceinline__ TS from_tc(typename Traits<TS>::TC x);
template <> __device__ __forceinline__ float from_tc<float>(float x) { return x; }
template <> __device__ __forceinline__ double from_tc<double>(double x) { return x; }
template <> __device__ __forceinline__ bf16_t from_tc<bf16_t>(float x) {
  // round-to-nearest-even (v_cvt_pk_bf16_f32 on gfx950)
  bf16_t r;
  r.u = __builtin_bit_cast(uint16_t, (__bf16)x);
  return r;
}

// fused multiply-add in the compute type
__device__ __forceinline__ float fmac(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmac(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float rcp_approx(float x) { return __builtin_amdgcn_rcpf(x); }

// ------------------------------------------------------------------ kernel arguments
// Static description of one instance: grid, physics and temperature model.  Built on the
// host from pf_params_t (make_stepdev) and passed as a kernel argument (single step) or
// kept in device memory (batched engine, reduction).
struct StepDev {
  const void* T;       // PF_T_BUFFER temperature
  const void* axes;    // [3][K][3] in TC
  int nz, ny, nx, k;
  int tiles_x, tiles_y, tiles_z, zc;
  int zs, zsplit;      // planes [0, zsplit) in chunks of zs, [zsplit, nz) in chunks of zc (zsplit 0: no tail)
  int tmodel, t_f64, force_liquid, pad0;
  float ihx2f, ihy2f, ihz2f, Bf;          // FP32 copies of ihx2, ihy2, ihz2, B
  double hx, hy, hz;
  double ihx2, ihy2, ihz2;
  double m_over_ad, kappa_ad, gamma4, B;  // m_g/ad_hoc, kappa_g*ad_hoc, 4*gamma, 1-4*gamma
  double L0, Qg, gas_const, T_l, T_clip, theta, zeta_window, eps, liquid_T;
  double stiff;  // m_g/ad_hoc (4 + 4|gamma|) + 4 kappa_g ad_hoc (1 + eps) sum 1/h^2 (cold-path threshold)
  // laser
  double T_amb, p0x, p0y, p0z, dirx, diry, dirz, vel;
  double q_coef, pe, T_cap;          // Rosenthal: Q/(2 pi kappa_T), vel/(2 alpha)
  double kappa_T, alpha;             // Rosenthal: the source's conductivity and diffusivity
  double dT_peak, g_sxy, g_sz;       // Gaussian: 1/(2 sigma_xy^2), 1/(2 sigma_z^2)
  double T_ref, grad_T, cool_rate;   // quench
};

// Rosenthal coefficients Q / (2 pi kappa_T) and vel / (2 alpha) (simple/example.py:35), one
// expression for the host and the device.
__host__ __device__ inline double rosenthal_q(double Q, double kappa_T) { return Q / (2.0 * M_PI * kappa_T); }
__host__ __device__ inline double rosenthal_pe(double vel, double alpha) { return vel / (2.0 * alpha); }

// FP32 constants of one step: the temperature model at t and the cold-path mobility.
struct ColdPath {
  float px, py, pz;           // source position at t
  float hx, hy, hz;
  float hs, T_lim;            // |hstep * stiff|; cold-path cells have T < T_lim and L(T) * hs < 1e-6
  float L0, nql2;             // Arrhenius: L0, -Qg / gas_const * log2(e)
  float hm, hk, eps1, eps2;   // hstep*m_g/ad_hoc, hstep*kappa_g*ad_hoc, 1+eps, 2*eps
  float gamma4;
  float T_amb, dT, gxy, gz, q, pe, dxr, dyr, dzr, T_cap, T_ref, grad_T, cool_t;
};

// One step of one instance: state buffers, time, the FP32 step constants and the step's
// stochastic-rounding seed.
struct StepDyn {
  const void* in;
  void* out;
  double t_prev, hstep;
  ColdPath cold;
  uint32_t seed;
};

// ------------------------------------------------------------------ stochastic rounding
// FP32 and BF16 storage round each Euler update stochastically instead of to nearest.  With
// round-to-nearest, a value near eta = 1 whose per-step change lies below half an ulp (2^-25)
// keeps its old value step after step while the FP64 solution moves on, so the error grows
// linearly, by up to 2^-25 per step (measured on a 2.4 M-cell, 822-step solve: 2.3e-5 against
// FP64, the same in a plain C FP32 restatement).  Rounding up with probability (distance to the
// lower neighbour) / (gap) is unbiased, so the error only random-walks (same solve: 2.5e-6 after
// 500 steps instead of 1.4e-5).  The random bits depend only on (cell index, channel, start time
// of the step), so every kernel (tile shape, lane split, generic kernel, batch engine) produces
// bitwise the same result.  The bits of channel i of a cell are key + i * 0x9e3779b9 (mod 2^32)
// with key a murmur3-style hash of (cell, step): each channel's bits are uniform and independent
// from step to step and from cell to cell, which is what unbiased rounding needs; within a cell
// the channels' bits form a Weyl sequence (evenly spread, not independent).
__host__ __device__ inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352dU;
  h ^= h >> 15;
  h *= 0x846ca68bU;
  h ^= h >> 16;
  return h;
}
__host__ __device__ inline uint32_t step_seed(double t) {
  const uint64_t b = __builtin_bit_cast(uint64_t, t);
  return mix32((uint32_t)b ^ mix32((uint32_t)(b >> 32) + 0x632be5abU));
}
// Key of cell (z, j) of a step, j = y * Nx + x its index within the z-plane (below 2^31): a hash
// of the step seed and z (uniform across a plane, so the march forms it in scalar registers)
// mixed with j.  Channel i of the cell takes key + i * kSrChannel.
__host__ __device__ inline uint32_t sr_plane_key(uint32_t seed, int z) {
  return mix32(seed ^ ((uint32_t)z * 0x85ebca6bU));
}
__device__ __forceinline__ uint32_t sr_cell_key(uint32_t plane_key, int j) { return mix32((uint32_t)j ^ plane_key); }
constexpr uint32_t kSrChannel = 0x9e3779b9U;
// c + d rounded stochastically to FP32.  The sum is formed in FP64 (exact while the exponents of c
// and d differ by at most 29, as for every small update; rounded at 2^-52 relative otherwise); 29
// random bits are added below the FP32 mantissa of that FP64 value and cut off, so the FP32
// magnitude rounds up with probability (cut-off part) / 2^29 and the expected result is c + d.  The
// FP64 -> FP32 conversion is then exact for normal FP32 results.  Infinities stay infinite (the
// added bits never carry into the exponent of a zero mantissa) and NaNs stay NaN.  Nine VALU
// instructions (two conversions, an FP64 add, a 64-bit integer add, a mask, a conversion back and
// the random bits) against twelve for an FP32 Fast2Sum residual compared with a random threshold.
// The random bits are (h + bits of d) mod 2^29: h is uniform and independent of d, so the sum is
// too, and it cannot be formed before d exists (so it holds no register across the channel loop).
__device__ __forceinline__ float sr_add_f32(float c, float d, uint32_t h) {
  const double x = (double)c + (double)d;
  const uint64_t b = (uint64_t)__double_as_longlong(x) + (uint64_t)((h + __float_as_uint(d)) & 0x1fffffffU);
  return (float)__longlong_as_double((long long)(b & ~0x1fffffffULL));
}
// x rounded stochastically to the BF16 grid (returned as the FP32 value of that BF16): 16
// random bits (the top half of h + bits of x, as in sr_add_f32) added below the BF16 mantissa,
// then truncated.  Arithmetic NaNs are quiet (bit
// 22 set), so they stay NaN; a carry past the largest finite value gives infinity.
__device__ __forceinline__ float sr_bf16(float x, uint32_t h) {
  const uint32_t xb = __float_as_uint(x);
  return __uint_as_float((xb + ((h + xb) >> 16)) & 0xffff0000U);
}
// rounding mode of the Euler update by storage type: 0 none (FP64), 1 FP32, 2 BF16
template <class TS> struct SrMode { static constexpr int value = 0; };
template <> struct SrMode<float> { static constexpr int value = 1; };
template <> struct SrMode<bf16_t> { static constexpr int value = 2; };

// Cold path: a cell whose FP32 mobility satisfies L(T) * |hstep * stiff| < 1e-6 changes by
// less than 1e-6 * |eta| per step, so the relative error of its FP32 mobility (hardware
// reciprocal and exp2: under 2e-5 while the exponent is above -126, below which the mobility
// is flushed to zero) moves it by under 2e-11 * |eta|.  Cold cells must also lie at least 10 K below T_liquidus, the
// liquid-reset threshold and T_clip, so zeta = 1 and no reset are the exact FP64 decisions.
// No transcendental is evaluated here, so host and device build bitwise the same constants.
__host__ __device__ inline ColdPath make_cold(const StepDev& a, double t, double hstep) {
  ColdPath k;
  const double vt = a.vel * t;
  k.px = (float)(a.p0x + vt * a.dirx);
  k.py = (float)(a.p0y + vt * a.diry);
  k.pz = (float)(a.p0z + vt * a.dirz);
  k.hx = (float)a.hx;
  k.hy = (float)a.hy;
  k.hz = (float)a.hz;
  k.hs = (float)fabs(hstep * a.stiff);
  k.T_lim = (float)(fmin(a.T_l, fmin(a.liquid_T, a.T_clip)) - 10.0);
  k.L0 = (float)a.L0;
  k.nql2 = (float)(-(a.Qg / a.gas_const) * 1.4426950408889634);
  k.hm = (float)(hstep * a.m_over_ad);
  k.hk = (float)(hstep * a.kappa_ad);
  k.eps1 = (float)(1.0 + a.eps);
  k.eps2 = (float)(2.0 * a.eps);
  k.gamma4 = (float)a.gamma4;
  k.T_amb = (float)a.T_amb;
  k.dT = (float)a.dT_peak;
  k.gxy = (float)a.g_sxy;
  k.gz = (float)a.g_sz;
  k.q = (float)a.q_coef;
  k.pe = (float)a.pe;
  k.dxr = (float)a.dirx;
  k.dyr = (float)a.diry;
  k.dzr = (float)a.dirz;
  k.T_cap = (float)a.T_cap;
  k.T_ref = (float)a.T_ref;
  k.grad_T = (float)a.grad_T;
  k.cool_t = (float)(a.cool_rate * t);
  return k;
}

__host__ __device__ inline StepDyn make_dyn(const StepDev& a, const void* in, void* out, double t_prev,
                                            double hstep) {
  StepDyn d;
  d.in = in;
  d.out = out;
  d.t_prev = t_prev;
  d.hstep = hstep;
  d.cold = make_cold(a, t_prev, hstep);
  d.seed = step_seed(t_prev);
  return d;
}

// ------------------------------------------------------------------ per-cell physics
// The Gaussian spot factorises: T = T_ambient + dT_peak * (exy * ez) with
// exy = exp(-(X^2 + Y^2) / (2 sigma_xy^2)) of the cell column and ez = exp(-Z^2 / (2 sigma_z^2))
// of the plane.  Every kernel evaluates it in this form, so the march can take exy once per
// column and ez once per plane and still agree bitwise with the per-cell evaluation.
__device__ __forceinline__ double gauss_exy64(const StepDev& a, int ix, int iy, double t) {
  const double vt = a.vel * t;
  const double X = (((double)ix + 0.5) * a.hx - a.p0x) - vt * a.dirx;
  const double Y = (((double)iy + 0.5) * a.hy - a.p0y) - vt * a.diry;
  return exp(-((X * X + Y * Y) * a.g_sxy));
}
__device__ __forceinline__ double gauss_ez64(const StepDev& a, int iz, double t) {
  const double Z = (((double)iz + 0.5) * a.hz - a.p0z) - a.vel * t * a.dirz;
  return exp(-(Z * Z * a.g_sz));
}
__device__ __forceinline__ double gauss_T64(const StepDev& a, double exy, double ez) {
  return a.T_amb + a.dT_peak * (exy * ez);
}
__device__ __forceinline__ float gauss_exy32(const ColdPath& k, int ix, int iy) {
  const float X = ((float)ix + 0.5f) * k.hx - k.px;
  const float Y = ((float)iy + 0.5f) * k.hy - k.py;
  return expf(-((X * X + Y * Y) * k.gxy));
}
__device__ __forceinline__ float gauss_ez32(const ColdPath& k, int iz) {
  const float Z = ((float)iz + 0.5f) * k.hz - k.pz;
  return expf(-(Z * Z * k.gz));
}
__device__ __forceinline__ float gauss_T32(const ColdPath& k, float exy, float ez) { return k.T_amb + k.dT * (exy * ez); }

// Global load whose value is forced into its register on the spot.  The compiler's wait
// analysis does not follow branch conditions: a load left outstanding on the temperature-buffer
// branch would make the other temperature models wait, where the branches join, for every
// outstanding memory operation, the march's prefetched plane and row stores included.
template <class T>
__device__ __forceinline__ T load_now(const T* p) {
  T v = *p;
  asm volatile("" : "+v"(v));
  return v;
}

// The same object, read through an offset the compiler cannot prove to be zero (an SGPR set by an
// asm statement).  Fields read through it are loaded (scalar loads) where they are used instead of
// being hoisted to the kernel's start: the FP64 hot branch reads some 30 uniform FP64 constants, and
// hoisted they stay live in scalar registers across the whole z-march.  On the FP32 K = 20 8x8
// lane-pair kernel that spilled 105 SGPRs to VGPR lanes, reloaded by ~90 v_readlane per plane on
// the common path; read here they cost scalar loads inside the branch only (18 SGPRs spilled,
// 163 -> 150 VGPRs; a 256x256x64, K = 20 step 0.1440 -> 0.1379 ms in FP32, 0.1154 -> 0.1114 ms
// with BF16 storage; bitwise the same results).
template <class T>
__device__ __forceinline__ const T& reload(const T& r) {
  int zero;
  asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
  return *(const T*)((const char*)&r + zero);
}

// Temperature at the centroid of cell (ix, iy, iz) at time t (FP64); cell index c for the buffer mode.
__device__ __forceinline__ double cell_temperature(const StepDev& a, int ix, int iy, int iz, long long c,
                                                   double t) {
  if (a.tmodel == PF_T_BUFFER) {
    return a.t_f64 ? load_now((const double*)a.T + c) : (double)load_now((const float*)a.T + c);
  }
  const double cx = ((double)ix + 0.5) * a.hx;
  const double cy = ((double)iy + 0.5) * a.hy;
  const double cz = ((double)iz + 0.5) * a.hz;
  if (a.tmodel == PF_T_QUENCH) {
    // solidification/example.py:32  T = T_l + G*z - G*v*t
    return a.T_ref + a.grad_T * cz - a.cool_rate * t;
  }
  const double vt = a.vel * t;
  const double X = (cx - a.p0x) - vt * a.dirx;
  const double Y = (cy - a.p0y) - vt * a.diry;
  const double Z = (cz - a.p0z) - vt * a.dirz;
  if (a.tmodel == PF_T_ROSENTHAL) {
    // simple/example.py:31-35, multi_scan/example.py:41-47
    const double R = sqrt(X * X + Y * Y + Z * Z);
    const double proj = X * a.dirx + Y * a.diry + Z * a.dirz;
    double T = a.T_amb + a.q_coef / R * exp(-a.pe * (R + proj));
    if (a.T_cap > 0.0 && T > a.T_cap) T = a.T_cap;
    return T;
  }
  // Gaussian moving spot
  (void)X;
  (void)Y;
  (void)Z;
  return gauss_T64(a, gauss_exy64(a, ix, iy, t), gauss_ez64(a, iz, t));
}

// The same temperature in FP32 (cold-path test and value).
__device__ __forceinline__ float cell_temperature_f32(const StepDev& a, const ColdPath& k, int ix, int iy, int iz,
                                                      long long c) {
  if (a.tmodel == PF_T_BUFFER) return a.t_f64 ? (float)load_now((const double*)a.T + c) : load_now((const float*)a.T + c);
  const float cz = ((float)iz + 0.5f) * k.hz;
  if (a.tmodel == PF_T_QUENCH) return k.T_ref + k.grad_T * cz - k.cool_t;
  const float X = ((float)ix + 0.5f) * k.hx - k.px;
  const float Y = ((float)iy + 0.5f) * k.hy - k.py;
  const float Z = cz - k.pz;
  if (a.tmodel == PF_T_ROSENTHAL) {
    const float R = sqrtf(X * X + Y * Y + Z * Z);
    const float proj = X * k.dxr + Y * k.dyr + Z * k.dzr;
    float T = k.T_amb + k.q / R * expf(-k.pe * (R + proj));
    if (k.T_cap > 0.0f && T > k.T_cap) T = k.T_cap;
    return T;
  }
  return gauss_T32(k, gauss_exy32(k, ix, iy), gauss_ez32(k, iz));
}

template <class TC> struct CellCoef {
  TC c1, c2, A, w1, w2;  // w1 = c2*(1+eps), w2 = 2*eps*c2 (anisotropic factor folded into c2)
  bool liquid;
};

// Per-cell coefficients of the update (allen_cahn.py:55-58,83-84,124-126), FP64 evaluation.
template <class TC>
__device__ __forceinline__ CellCoef<TC> cell_coef64(const StepDev& a, double T, TC S, double hstep) {
  CellCoef<TC> r;
  const double Tc = T > a.T_clip ? a.T_clip : T;                 // :56
  // :57  zeta = 0.5*(1 - tanh(theta*(Tc/T_l - 1))).  In FP64, tanh(x) rounds to exactly +-1
  // for |x| >= 20, i.e. outside a window of width ~2e-9*T_l around T_l for theta = 1e10;
  // zeta_window (host side) is a safe bound on that window, so outside it a compare
  // gives the identical value and the division + tanh run only inside it.
  double zeta;
  if (fabs(Tc - a.T_l) > a.zeta_window) {
    zeta = Tc < a.T_l ? 1.0 : 0.0;
  } else {
    const double x = a.theta * (Tc / a.T_l - 1.0);
    const double th = fabs(x) >= 20.0 ? copysign(1.0, x) : tanh(x);
    zeta = 0.5 * (1.0 - th);
  }
  const double omz = 1.0 - zeta;
  const double Lg = a.L0 * exp(-a.Qg / (Tc * a.gas_const));      // :83
  const double hL = hstep * Lg;
  const double c1 = hL * a.m_over_ad;
  const double c2 = hL * a.kappa_ad;
  r.c1 = (TC)c1;
  r.c2 = (TC)c2;
  r.w1 = (TC)(c2 * (1.0 + a.eps));
  r.w2 = (TC)(2.0 * a.eps * c2);
  r.A = fmac((TC)a.gamma4, S, (TC)(2.0 * omz * omz - 1.0));
  r.liquid = a.force_liquid && (T > a.liquid_T);                 // :124-126 (unclipped T)
  return r;
}

// Per-cell coefficients from the cell's FP32 temperature T32 and a callable giving its FP64
// temperature.  F64 storage: always the FP64 evaluation.  FP32 arithmetic: each cold-path cell
// (make_cold) takes an FP32 mobility (zeta = 1, no liquid reset: exact there); every other cell
// takes the FP64 evaluation.  The choice is made per cell from the cell's own temperature, so
// results do not depend on which cells share a wavefront; the FP64 temperature is evaluated,
// and the FP64 branch entered, only by wavefronts holding a hot cell.
template <class TC, class F64>
__device__ __forceinline__ CellCoef<TC> cell_coef_with(const StepDev& a, const ColdPath& k, float T32,
                                                       const F64& t64, double hstep, TC S) {
  if constexpr (sizeof(TC) == 8) {
    (void)k;
    (void)T32;
    return cell_coef64<TC>(a, t64(), S, hstep);
  } else {
    CellCoef<float> r;
    const float Lg = k.L0 * __builtin_amdgcn_exp2f(k.nql2 * __builtin_amdgcn_rcpf(T32));
    const bool hot = !(T32 < k.T_lim && fabsf(Lg) * k.hs < 1e-6f);  // NaN -> FP64 path
    r.c1 = k.hm * Lg;
    r.c2 = k.hk * Lg;
    r.w1 = r.c2 * k.eps1;
    r.w2 = r.c2 * k.eps2;
    r.A = fmac(k.gamma4, S, -1.0f);
    r.liquid = false;
    if (__ballot(hot)) {
      if (hot) r = cell_coef64<float>(reload(a), t64(), S, hstep);  // constants loaded on this branch only
    }
    return r;
  }
}

// Per-cell coefficients of cell (ix, iy, iz) (index c) at time t.
template <class TC>
__device__ __forceinline__ CellCoef<TC> cell_coef(const StepDev& a, const ColdPath& k, int ix, int iy, int iz,
                                                  long long c, double t, double hstep, TC S) {
  const float T32 = sizeof(TC) == 8 ? 0.0f : cell_temperature_f32(a, k, ix, iy, iz, c);
  return cell_coef_with<TC>(a, k, T32, [&] { return cell_temperature(a, ix, iy, iz, c, t); }, hstep, S);
}

// A cell whose coefficients are not finite (non-finite S, or a non-finite temperature that makes
// zeta and the mobility NaN): the reference's update of every one of its channels is then NaN,
// a channel that is zero on its whole stencil included (0 * NaN in the local term), so such a
// cell never takes the zero-channel skip.
template <class TC>
__device__ __forceinline__ bool coef_nonfinite(const CellCoef<TC>& cf) { return !isfinite(cf.A + cf.c1 + cf.c2); }

// Bitwise "non-zero" test of a channel's 7-point stencil (a -0.0 counts as non-zero), or of `keep`
// (1 in a cell with non-finite coefficients, which must not skip; 0 otherwise).  One OR chain, so the
// ballot on it is a single compare (a wave-uniform flag ORed into the ballot's mask instead costs two
// scalar instructions per channel and measured 0.3-0.7 % slower).
__device__ __forceinline__ bool any_nz(float c, float a, float b, float d, float e, float f, float g, uint32_t keep) {
  return (__float_as_uint(c) | __float_as_uint(a) | __float_as_uint(b) | __float_as_uint(d) | __float_as_uint(e) |
          __float_as_uint(f) | __float_as_uint(g) | keep) != 0u;
}
__device__ __forceinline__ bool any_nz(double c, double a, double b, double d, double e, double f, double g,
                                       uint32_t keep) {
  return (__double_as_longlong(c) | __double_as_longlong(a) | __double_as_longlong(b) | __double_as_longlong(d) |
          __double_as_longlong(e) | __double_as_longlong(f) | __double_as_longlong(g) | (long long)keep) != 0;
}

// One orientation channel: centre value c, x/y/z neighbours (edge-replicated at the
// domain faces by the caller), returns eta_new before the liquid reset.
// With RHS = true it returns -L*(grad + laplace) instead (c1, c2 then carry no dt),
// the reference's state_rhs value.
// The anisotropy factor is skipped when |grad eta|^2 == 0 on every lane of the wavefront:
// there the full evaluation gives q = 1 and ca = w1 exactly, which is what the skip uses.
// ax: the rotated cube axes of this orientation, axis k at ax[4k .. 4k+2].
// With coefficients that carry dt (the step's), RHS = true returns the Euler increment
// eta_new - eta, which FP32 storage adds to eta with stochastic rounding (update_sr).
template <class TC, bool ANISO, bool RHS = false>
__device__ __forceinline__ TC update_one(const CellCoef<TC>& cf, TC B, TC ihx2, TC ihy2, TC ihz2,
                                         TC c, TC xm, TC xp, TC ym, TC yp, TC zm, TC zp,
                                         const TC* __restrict__ ax) {
  const TC dx = fmac(c, (TC)-2, xm + xp);
  const TC dy = fmac(c, (TC)-2, ym + yp);
  const TC dz = fmac(c, (TC)-2, zm + zp);
  const TC lap = fmac(dz, ihz2, fmac(dy, ihy2, dx * ihx2));
  TC ca;
  if constexpr (ANISO) {
    const TC ex = xp - xm, ey = yp - ym, ez = zp - zm;
    const TC n2 = fmac(ez, ez, fmac(ey, ey, ex * ex));
    ca = cf.w1;
    if (__ballot(n2 != (TC)0)) {
      const TC* a0 = ax;
      const TC* a1 = ax + 4;
      const TC* a2 = ax + 8;
      const TC p0 = fmac(a0[2], ez, fmac(a0[1], ey, a0[0] * ex));
      const TC p1 = fmac(a1[2], ez, fmac(a1[1], ey, a1[0] * ex));
      const TC p2 = fmac(a2[2], ez, fmac(a2[1], ey, a2[0] * ex));
      TC q;
      if constexpr (sizeof(TC) == 4) {
        const TC m = fmaxf(fmaxf(fabsf(p0), fabsf(p1)), fabsf(p2));
        q = fminf(m * m * rcp_approx(n2), 1.0f);   // n2 == 0 -> NaN -> 1 (minNum)
      } else {
        const TC m = fmax(fmax(fabs(p0), fabs(p1)), fabs(p2));
        q = n2 > 0.0 ? fmin((m * m) / n2, 1.0) : 1.0;
      }
      ca = fmac(cf.w2, fmac(q, q, -q), cf.w1);   // w1 + w2 q (q - 1); exactly w1 at q = 1
    }
  } else {
    (void)ax;
    ca = cf.c2;
  }
  const TC e2 = c * c;
  const TC t = fmac(e2, B, cf.A);
  const TC u = c * t;
  if constexpr (RHS) {
    return fmac(ca, lap, -cf.c1 * u);
  } else {
    const TC r = fmac(-cf.c1, u, c);
    return fmac(ca, lap, r);
  }
}

// The Euler update of one channel in the storage type's rounding: FP64 to nearest; FP32 adds the
// increment to eta with stochastic rounding (sr_add_f32); BF16 rounds the FP32 update
// stochastically to the BF16 grid (sr_bf16).  h: the channel's random bits.
template <class TS, bool ANISO>
__device__ __forceinline__ typename Traits<TS>::TC update_sr(const CellCoef<typename Traits<TS>::TC>& cf,
                                                             typename Traits<TS>::TC B, typename Traits<TS>::TC ihx2,
                                                             typename Traits<TS>::TC ihy2, typename Traits<TS>::TC ihz2,
                                                             typename Traits<TS>::TC c, typename Traits<TS>::TC xm,
                                                             typename Traits<TS>::TC xp, typename Traits<TS>::TC ym,
                                                             typename Traits<TS>::TC yp, typename Traits<TS>::TC zm,
                                                             typename Traits<TS>::TC zp,
                                                             const typename Traits<TS>::TC* __restrict__ ax, uint32_t h) {
  using TC = typename Traits<TS>::TC;
  if constexpr (SrMode<TS>::value == 1) {
    return sr_add_f32(c, update_one<TC, ANISO, true>(cf, B, ihx2, ihy2, ihz2, c, xm, xp, ym, yp, zm, zp, ax), h);
  } else if constexpr (SrMode<TS>::value == 2) {
    return sr_bf16(update_one<TC, ANISO, false>(cf, B, ihx2, ihy2, ihz2, c, xm, xp, ym, yp, zm, zp, ax), h);
  } else {
    (void)h;
    return update_one<TC, ANISO, false>(cf, B, ihx2, ihy2, ihz2, c, xm, xp, ym, yp, zm, zp, ax);
  }
}

template <class TC> __device__ __forceinline__ TC lap_B(const StepDev& a);
template <> __device__ __forceinline__ float lap_B<float>(const StepDev& a) { return a.Bf; }
template <> __device__ __forceinline__ double lap_B<double>(const StepDev& a) { return a.B; }
template <class TC> __device__ __forceinline__ void inv_h2(const StepDev& a, TC& x, TC& y, TC& z);
template <> __device__ __forceinline__ void inv_h2<float>(const StepDev& a, float& x, float& y, float& z) {
  x = a.ihx2f;
  y = a.ihy2f;
  z = a.ihz2f;
}
template <> __device__ __forceinline__ void inv_h2<double>(const StepDev& a, double& x, double& y, double& z) {
  x = a.ihx2;
  y = a.ihy2;
  z = a.ihz2;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// s_waitcnt immediate of gfx9 for vmcnt(0) alone (expcnt and lgkmcnt fields at their maxima)
constexpr int kWaitVm0 = 0x0F70;

// Raw buffer stores: the fourth descriptor word of an untyped gfx9 buffer, and an offset
// that the hardware's range check drops (any buffer of the march is under 2^31 bytes).
constexpr int kRawBufferDword3 = 0x00020000;
constexpr unsigned kOob = 0x80000000u;

// N bytes from src (N = 1, 2, 4, 8 or 16, aligned to N) to byte offset off of buffer r;
// NT marks the store non-temporal.
template <int N, bool NT>
__device__ __forceinline__ void buffer_store(const void* src, __amdgpu_buffer_rsrc_t r, unsigned off) {
  constexpr int aux = NT ? 2 : 0;
  typedef unsigned U2 __attribute__((ext_vector_type(2)));
  typedef unsigned U4 __attribute__((ext_vector_type(4)));
  if constexpr (N == 16) __builtin_amdgcn_raw_buffer_store_b128(*(const U4*)src, r, (int)off, 0, aux);
  else if constexpr (N == 8) __builtin_amdgcn_raw_buffer_store_b64(*(const U2*)src, r, (int)off, 0, aux);
  else if constexpr (N == 4) __builtin_amdgcn_raw_buffer_store_b32(*(const unsigned*)src, r, (int)off, 0, aux);
  else if constexpr (N == 2) __builtin_amdgcn_raw_buffer_store_b16(*(const unsigned short*)src, r, (int)off, 0, aux);
  else __builtin_amdgcn_raw_buffer_store_b8(*(const unsigned char*)src, r, (int)off, 0, aux);
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also releases global
// memory, which makes every wave wait for all of its outstanding global loads and stores
// (s_waitcnt vmcnt(0)) at the barrier; the march's prefetched plane and its row stores
// must stay in flight across the barriers, and no wave reads global data written by
// another wave of the same launch (eta_in and eta_out never alias).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Copy the rotated cube axes ([3][K][3] in global memory: axis k of orientation i at
// (k*K + i)*3) into the LDS layout used by the march: orientation i at ax[12 i], axis k at
// ax[12 i + 4 k].  Reads through ordinary per-lane loads once per workgroup; the caller's
// next barrier publishes the table.
template <class TC, int K>
__device__ __forceinline__ void stage_axes(const TC* __restrict__ axes, TC* __restrict__ ax, int tid, int nt) {
  for (int w = tid; w < 9 * K; w += nt) {
    const int i = w / 9, r = w - 9 * (w / 9);
    const int k = r / 3, j = r - 3 * (r / 3);
    ax[12 * i + 4 * k + j] = axes[(k * K + i) * 3 + j];
  }
}

// ------------------------------------------------------------------ tile geometry
constexpr int p2div(int n, int cap) {
  int v = cap;
  while (v > 1 && n % v) v /= 2;
  return v;
}

// Bank conflicts (extra LDS cycles) of one wavefront's neighbour reads: lane l reads LV
// elements of EB bytes at cell (l / H) of a TX-wide tile row, channel offset (l % H) * KH,
// cell stride KP elements.  NB banks of 4 bytes; a read of W bytes per lane is served
// 4 * NB / W lanes at a time.
constexpr int lds_conflicts(int KP, int KH, int LV, int H, int TX, int EB, int NB) {
  const int W = LV * EB;
  int per = (4 * NB) / W;
  if (per < 1) per = 1;
  if (per > 64) per = 64;
  int total = 0;
  for (int p0 = 0; p0 < 64; p0 += per) {
    int cnt[64] = {};
    for (int l = p0; l < p0 + per && l < 64; ++l) {
      const int m = l / H, h = l % H;
      const int lx = m % TX, ly = m / TX;
      const int word = ((((ly + 1) * (TX + 2) + lx + 1) * KP + h * KH) * EB) / 4;
      for (int w = 0; w < W / 4; ++w) cnt[(word + w) % NB] += 1;
    }
    int mx = 0;
    for (int b = 0; b < NB; ++b) mx = cnt[b] > mx ? cnt[b] : mx;
    total += mx - 1;
  }
  return total;
}

// Channel layout of one tile shape.  H threads per cell: H = 1, or H = 2 (even K: the two
// lanes of a pair (2m, 2m+1) hold channels [0, K/2) and [K/2, K) of the same cell).
template <class TS, int K, int TX, int TY, int H> struct Tile {
  using TC = typename Traits<TS>::TC;
  static_assert(H == 1 || (H == 2 && K % 2 == 0), "H = 2 needs even K");
  static constexpr int KH = K / H;               // channels per thread
  // widest global chunk: 16 bytes of TS and of TC
  static constexpr int VCAP = (16 / (int)sizeof(TC)) < (16 / (int)sizeof(TS)) ? 16 / (int)sizeof(TC)
                                                                              : 16 / (int)sizeof(TS);
  // K not a multiple of VCAP: a cell is loaded as K / VCAP chunks of VCAP elements plus one
  // tail chunk of K % VCAP elements, at the cell's own alignment (GA = sizeof(TS) bytes;
  // gfx950 global accesses run in unaligned mode, so these stay single dwordx2/x3/x4 loads).
  static constexpr bool SPLIT = K % VCAP != 0 && K > VCAP;
  static constexpr int VEC = SPLIT ? VCAP : p2div(K, VCAP);
  static constexpr int TAIL = SPLIT ? K % VCAP : 0;
  static constexpr int GA = SPLIT ? (int)sizeof(TS) : VEC * (int)sizeof(TS);
  // LDS read vector (TC elements): 16 bytes for H = 1 (the last group of a cell may read
  // padding), else the widest power of two dividing KH; NG groups per neighbour
  static constexpr int LV = H == 1 ? 16 / (int)sizeof(TC) : p2div(KH, 16 / (int)sizeof(TC));
  static constexpr int NG = (KH + LV - 1) / LV;
  // cell stride KP (TC elements): the multiple of LV covering every lane's reads
  // ((H - 1) * KH + NG * LV >= K), among the next five, with the fewest bank conflicts of the
  // neighbour reads for 32 and 64 banks
  static constexpr int KMIN = (H - 1) * KH + NG * LV;
  static constexpr int pick_kp() {
    int best = KMIN, bc = 1 << 30;
    for (int c = 0; c < 5; ++c) {
      const int kp = KMIN + c * LV;
      const int n = lds_conflicts(kp, KH, LV, H, TX, (int)sizeof(TC), 32) +
                    lds_conflicts(kp, KH, LV, H, TX, (int)sizeof(TC), 64);
      if (n < bc) {
        bc = n;
        best = kp;
      }
    }
    return best;
  }
  static constexpr int KP = pick_kp();
  static constexpr int WV = p2div(KP, VEC);                   // LDS write vector of a load chunk
  static constexpr int SVS = p2div(KH, 16 / (int)sizeof(TS));  // staging store vector (TS elements)
  static constexpr int NC = TX * TY;             // cells per workgroup
  static constexpr int NT = NC * H;              // threads per workgroup
  static constexpr int PX = TX + 2, PY = TY + 2;
  static constexpr int CPC = K / VEC;            // full chunks per cell
  static constexpr int NCH = PY * PX * CPC;      // full chunks per plane
  static constexpr int CPT = (NCH + NT - 1) / NT;
  static constexpr int NCT = TAIL ? PY * PX : 0;  // tail chunks per plane (one per cell)
  static constexpr int CPTT = (NCT + NT - 1) / NT;
  static constexpr int PLANE = PY * PX * KP;     // TC elements per LDS plane
  // LDS plane stride (TC elements): the plane plus one chunk slot at PLANE, which takes the
  // hand-over's writes of the lanes past the plane's last chunk (so that every lane writes
  // every chunk without a per-lane condition); 16-byte aligned
  static constexpr int PSTR = (PLANE + VEC + 3) & ~3;
  // BF16 lane-pair tiles keep a ring of three LDS planes and store their results straight
  // from registers: one barrier per plane (256x256x64, K = 20 BF16 step: -3 %).  Every other
  // tile keeps two planes and stages its results in LDS for coalesced row stores, two barriers
  // per plane: FP32 pieces stored per lane (8 bytes at an 80-byte cell stride) cost more than
  // the barrier saves (the same step in FP32: 0.142 -> 0.182 ms), and for one-lane tiles a
  // third plane would not fit two workgroups per CU.
  static constexpr bool RING3 = H == 2 && sizeof(TS) == 2;
  static constexpr int NBUF = RING3 ? 3 : 2;
  static constexpr int STAGE = PF_STORE_MODE != 0 && !RING3 ? NC * K : 8;  // TS elements of output staging
  static constexpr int AXN = 12 * K;                                        // TC elements of the LDS axis table
  static constexpr size_t LDS_BYTES = NBUF * PSTR * sizeof(TC) + STAGE * sizeof(TS) + AXN * sizeof(TC) + 768;
};

// The sum of squares S = sum_i eta_i^2 of a cell is formed as two sequential FMA chains, over
// channels [0, K/2) and [K/2, K), added low + high.  Every kernel uses this order, so a cell's
// S is bitwise the same whether one thread or a lane pair holds its channels.
template <class TC>
__device__ __forceinline__ TC sum_sq_halves(TC lo, TC hi) { return lo + hi; }

// The value of the other lane of a lane pair (lanes 2m and 2m + 1): a DPP quad permutation
// [1, 0, 3, 2] on the VALU, where a shuffle would round-trip through the LDS crossbar
// (ds_bpermute) on the path to the cell's coefficients.
__device__ __forceinline__ float pair_swap(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pair_swap(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Gaussian z-factors of the planes of one march (LDS), for marches of up to kZTab planes.
constexpr int kZTab = 64;
struct ZTab {
  double e64[kZTab];
  float e32[kZTab];
};

// First plane and length of z-chunk tile_z (0 = the lowest): the tail [0, zsplit) in chunks of zs,
// then chunks of zc up to the top (set_tiles).
__device__ __forceinline__ void zchunk_range(const StepDev& a, int tile_z, int& zb, int& len) {
  const int nb = a.zsplit > 0 ? (a.zsplit + a.zs - 1) / a.zs : 0;
  if (tile_z < nb) {
    zb = tile_z * a.zs;
    len = min(zb + a.zs, a.zsplit) - zb;
  } else {
    zb = a.zsplit + (tile_z - nb) * a.zc;
    len = min(zb + a.zc, a.nz) - zb;
  }
}

// ------------------------------------------------------------------ 2.5-D march of one tile
// A workgroup owns a TX x TY column of cells (H threads per cell) and marches it upward through
// the z-planes [zb, zb + len).  Planes (tile + one-cell halo ring, edge-replicated at the domain
// faces) are double-buffered in LDS; the plane two ahead is prefetched into registers while the
// current plane is computed.  Each thread keeps its KH channels of its cell at z-1, z and z+1
// in three register arrays whose roles rotate by renaming (the z loop is unrolled by three).
// With H = 2 the two lanes of a pair exchange their partial sums of squares with a lane
// shuffle.  With the Gaussian source each thread takes its column factor once and each plane's
// factor comes from an LDS table filled at the start (gauss_exy*, gauss_ez*).
// Results are staged in LDS and written back as contiguous row segments (PF_STORE_MODE 1, 2)
// or stored per cell (0).  The instance and step arguments are read through references (kernel
// argument or device memory), so uniform values stay in scalar registers.
template <int K, class TS, bool ANISO, int TX, int TY, int H>
struct March {
  using TC = typename Traits<TS>::TC;
  using G = Tile<TS, K, TX, TY, H>;
  static constexpr int VEC = G::VEC, KP = G::KP, LV = G::LV, NG = G::NG, WV = G::WV, TAIL = G::TAIL;
  static constexpr int KH = G::KH, KS = K / 2;
  static constexpr int TV = TAIL > 0 ? TAIL : 1;
  using SV = VecA<TS, VEC, G::GA>;
  using TLv = VecA<TS, TV, G::GA>;
  using WVv = Vec<TC, WV>;
  using LVv = Vec<TC, LV>;
  using SVv = Vec<TS, G::SVS>;

  const StepDev& a;
  const StepDyn& d;
  TC* __restrict__ buf0;
  TC* __restrict__ buf1;
  TC* __restrict__ buf2;  // third plane of the ring (RING3)
  TS* __restrict__ stage;
  const TC* __restrict__ ax;  // LDS axis table (ANISO)
  const ZTab* __restrict__ zt;  // LDS Gaussian z-factors of this march's planes (gsep)
  unsigned src_off[G::CPT];  // byte offsets within an input plane (unsigned 32-bit: base + offset addressing)
  int dst_off[G::CPT];
  SV pre[G::CPT];
  unsigned tsrc_off[G::CPTT > 0 ? G::CPTT : 1];
  int tdst_off[G::CPTT > 0 ? G::CPTT : 1];
  TLv pret[G::CPTT > 0 ? G::CPTT : 1];
  // Row write-back (store_rows): each contiguous (y, z) row segment of the tile goes out in
  // CH-byte pieces, NJO pieces per thread; opos[j] is the byte offset of this thread's
  // piece j within an output plane (kNoPiece: none), fixed for the whole march.
  static constexpr int CELLB = K * (int)sizeof(TS);
  static constexpr int ROWB = TX * CELLB;
  // rows whose full length is a multiple of 16 bytes go out in 16-byte pieces whatever the
  // cell alignment (a partial row's last piece in units of U bytes); otherwise the widest
  // power-of-two piece dividing a cell
  static constexpr int CA = CELLB & -CELLB;
  static constexpr bool WIDE = ROWB % 16 == 0 && CA < 16;
  static constexpr int CH = WIDE ? 16 : (CA >= 16 ? 16 : CA);
  static constexpr int U = CA < 4 ? CA : 4;
  static constexpr int CPR = ROWB / CH;
  static constexpr int NCHO = TY * CPR;
  static constexpr int NJO = (NCHO + G::NT - 1) / G::NT;
  static constexpr unsigned kNoPiece = kOob;
  // FP64 tiles write their rows back through buffer stores that every lane issues (see
  // store_rows): a 256x256x64, K = 20 FP64 step 0.3355 -> 0.3247 ms.  FP32 rows measured slower
  // that way (the same step in FP32 0.1351 -> 0.1367 ms), and so did the three-plane ring's
  // per-cell BF16 stores (0.1006 -> 0.1252 ms): there, waiting for the previous plane's stores
  // at each hand-over is the faster schedule.
  static constexpr bool BUF_ROWS = PF_STORE_MODE != 0 && !WIDE && sizeof(TC) == 8;
  unsigned opos[PF_STORE_MODE != 0 ? NJO : 1];
  int gx, gy, cl, ch0, own, oxm, oxp, oym, oyp, zb, len, x0, y0;
  bool active, gsep;
  double exy64;
  float exy32;

  __device__ __forceinline__ March(const StepDev& a_, const StepDyn& d_, int tile_x, int tile_y, int tile_z,
                                   TC* smem, TS* stage_, TC* ax_, ZTab* zt_)
      : a(a_), d(d_) {
    const int tid = threadIdx.x;
    cl = H == 1 ? tid : tid >> 1;
    ch0 = H == 1 ? 0 : (tid & 1) * KH;
    const int lx = cl % TX, ly = cl / TX;
    x0 = tile_x * TX;
    y0 = tile_y * TY;
    gx = x0 + lx;
    gy = y0 + ly;
    active = (gx < a.nx) && (gy < a.ny);
    zchunk_range(a, tile_z, zb, len);
    buf0 = smem;
    buf1 = smem + G::PSTR;
    buf2 = smem + (G::RING3 ? 2 : 1) * G::PSTR;
    stage = stage_;
    ax = ax_;
    if constexpr (ANISO) stage_axes<TC, K>((const TC*)a.axes, ax_, tid, G::NT);
    zt = zt_;
    gsep = a.tmodel == PF_T_GAUSSIAN && len <= kZTab;
    exy64 = 0.0;
    exy32 = 0.0f;
    if (gsep) {
      exy64 = gauss_exy64(a, gx, gy, d.t_prev);
      exy32 = gauss_exy32(d.cold, gx, gy);
      for (int w = tid; w < len; w += G::NT) {
        zt_->e64[w] = gauss_ez64(a, zb + w, d.t_prev);
        zt_->e32[w] = gauss_ez32(d.cold, zb + w);
      }
    }
#pragma unroll
    for (int j = 0; j < G::CPT; ++j) {
      const int w = tid + j * G::NT;
      const int ww = w < G::NCH ? w : G::NCH - 1;
      const int cs = ww / G::CPC, v = ww % G::CPC;
      const int r = cs / G::PX, cc = cs % G::PX;
      const int sy = clampi(y0 - 1 + r, 0, a.ny - 1);
      const int sx = clampi(x0 - 1 + cc, 0, a.nx - 1);
      src_off[j] = (unsigned)((sy * a.nx + sx) * K + v * VEC) * (unsigned)sizeof(TS);
      dst_off[j] = (w < G::NCH) ? cs * KP + v * VEC : G::PLANE;
    }
    if constexpr (TAIL > 0) {
#pragma unroll
      for (int j = 0; j < G::CPTT; ++j) {
        const int w = tid + j * G::NT;
        const int cs = w < G::NCT ? w : G::NCT - 1;
        const int r = cs / G::PX, cc = cs % G::PX;
        const int sy = clampi(y0 - 1 + r, 0, a.ny - 1);
        const int sx = clampi(x0 - 1 + cc, 0, a.nx - 1);
        tsrc_off[j] = (unsigned)((sy * a.nx + sx) * K + G::CPC * VEC) * (unsigned)sizeof(TS);
        tdst_off[j] = (w < G::NCT) ? cs * KP + G::CPC * VEC : G::PLANE;
      }
    }
#if PF_STORE_MODE != 0
    {
      const int rowb = min(TX, a.nx - x0) * CELLB;
#pragma unroll
      for (int j = 0; j < NJO; ++j) {
        const int w = tid + j * G::NT;
        const int r = w / CPR;
        const int off = (w - r * CPR) * CH;
        opos[j] = (w < NCHO && y0 + r < a.ny && off < rowb)
                      ? (unsigned)((y0 + r) * a.nx + x0) * (unsigned)CELLB + (unsigned)off
                      : kNoPiece;
      }
    }
#endif
    own = ((ly + 1) * G::PX + (lx + 1)) * KP + ch0;
    oxm = own - KP;
    oxp = own + KP;
    oym = own - G::PX * KP;
    oyp = own + G::PX * KP;
  }

  __device__ __forceinline__ long long plane_elems() const { return (long long)a.ny * a.nx * K; }

  __device__ __forceinline__ void load_plane(int z) {
    const TS* p = (const TS*)d.in + (long long)z * plane_elems();
#pragma unroll
    for (int j = 0; j < G::CPT; ++j) pre[j] = *(const SV*)((const char*)p + src_off[j]);
    if constexpr (TAIL > 0) {
#pragma unroll
      for (int j = 0; j < G::CPTT; ++j) pret[j] = *(const TLv*)((const char*)p + tsrc_off[j]);
    }
  }
  // LDS chunk w sits at w * VEC when cells are unpadded (KP = K, no tail chunk): the offsets
  // are then immediates off one per-thread base instead of a register per chunk.
  static constexpr bool DLIN = KP == G::CPC * VEC && TAIL == 0;
  // Every lane writes every chunk: the chunks past the plane's end (lanes of the last pass
  // beyond it, which loaded a copy of the last chunk) land in the slot at PLANE.  With the
  // last chunk's write behind a per-lane condition, the compiler waited for everything in
  // flight (the previous plane's row stores included) before requesting the next plane
  // (256x256x64, K = 20 step: BF16 0.1035 -> 0.1006 ms, FP64 0.3355 -> 0.3318, FP32 0.1355 -> 0.1351).
  __device__ __forceinline__ void store_plane(TC* b) {
#pragma unroll
    for (int j = 0; j < G::CPT; ++j) {
      TC* dst = !DLIN                 ? b + dst_off[j]
                : j * G::NT + G::NT <= G::NCH ? b + (int)threadIdx.x * VEC + j * G::NT * VEC
                                              : b + min((int)threadIdx.x + j * G::NT, G::NCH) * VEC;
#pragma unroll
      for (int q = 0; q < VEC / WV; ++q) {
        WVv c;
#pragma unroll
        for (int e = 0; e < WV; ++e) c.v[e] = to_tc(pre[j].v[q * WV + e]);
        *(WVv*)(dst + q * WV) = c;
      }
    }
    if constexpr (TAIL > 0) {
#pragma unroll
      for (int j = 0; j < G::CPTT; ++j) {
#pragma unroll
        for (int e = 0; e < TAIL; ++e) b[tdst_off[j] + e] = to_tc(pret[j].v[e]);
      }
    }
  }
  __device__ __forceinline__ void read_center(const TC* b, TC (&dst)[KH]) {
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const LVv v = *(const LVv*)(b + own + g * LV);
#pragma unroll
      for (int l = 0; l < LV; ++l)
        if (g * LV + l < KH) dst[g * LV + l] = v.v[l];
    }
  }

  // S of this thread's cell (order: sum_sq_halves of the two chains [0, KS) and [KS, K))
  __device__ __forceinline__ TC cell_sum_sq(const TC (&c)[KH]) {
    if constexpr (H == 1) {
      TC lo = 0, hi = 0;
#pragma unroll
      for (int i = 0; i < KS; ++i) lo = fmac(c[i], c[i], lo);
#pragma unroll
      for (int i = KS; i < K; ++i) hi = fmac(c[i], c[i], hi);
      return sum_sq_halves(lo, hi);
    } else {
      TC part = 0;
#pragma unroll
      for (int i = 0; i < KH; ++i) part = fmac(c[i], c[i], part);
      const TC oth = pair_swap(part);
      return ch0 == 0 ? sum_sq_halves(part, oth) : sum_sq_halves(oth, part);
    }
  }

  // plane z + d clamped to the domain (edge replication in z)
  __device__ __forceinline__ int zstep(int z, int d) const { return clampi(z + d, 0, a.nz - 1); }

  // Compute plane z = zb + n from (lo, c, hi) = planes z-1, z, z+1, then (if another plane
  // follows) put plane z+2 into LDS and read its centre into `lo`.  cb holds plane z.  Two
  // planes: plane z+2 replaces plane z in cb after a barrier.  Ring of three (RING3): plane
  // z+2 goes to tb, the plane read while computing z-1 (done before the previous barrier).
  __device__ __forceinline__ void body(int n, TC (&lo)[KH], TC (&c)[KH], TC (&hi)[KH], TC* cb, TC* tb) {
    const int z = zb + n;
    TC res[KH];
    long long cell = 0;
    if (active) {
      cell = ((long long)z * a.ny + gy) * a.nx + gx;
      const TC S = cell_sum_sq(c);
      float T32 = 0.0f;
      if constexpr (sizeof(TC) == 4)
        T32 = gsep ? gauss_T32(d.cold, exy32, zt->e32[n]) : cell_temperature_f32(a, d.cold, gx, gy, z, cell);
      const CellCoef<TC> cf = cell_coef_with<TC>(
          a, d.cold, T32,
          [&] {
            // With FP32 arithmetic the FP64 temperature runs only in wavefronts holding a hot cell.
            // Its column terms do not change with z, but hoisted out of the march they would occupy
            // (or spill) registers in every wavefront: the opaque copies keep them inside this branch.
            int gxo = gx, gyo = gy;
            if constexpr (sizeof(TC) == 4) asm volatile("" : "+v"(gxo), "+v"(gyo));
            const StepDev& ah = reload(a);  // the temperature model's constants, loaded here
            return gsep ? gauss_T64(ah, exy64, zt->e64[n]) : cell_temperature(ah, gxo, gyo, z, cell, reload(d).t_prev);
          },
          d.hstep, S);
      const TC B = lap_B<TC>(a);
      TC ihx2, ihy2, ihz2;
      inv_h2<TC>(a, ihx2, ihy2, ihz2);
      const uint32_t keep = coef_nonfinite(cf) ? 1u : 0u;
      constexpr bool SR = SrMode<TS>::value != 0;
      uint32_t key = SR ? sr_cell_key(sr_plane_key(d.seed, z), gy * a.nx + gx) + (uint32_t)ch0 * kSrChannel : 0u;
      // FP32 storage: opaque, otherwise the compiler reassociates key + i * kSrChannel and keeps the
      // plane-invariant sums ch0 * kSrChannel + i * kSrChannel in registers for the whole march, which
      // spills the lane-pair kernels (BF16 storage has the registers to spare)
      if constexpr (SrMode<TS>::value == 1) asm volatile("" : "+v"(key));
      // The channels of neighbour-vector group g (LV channels, read as four LDS vectors).
      auto group = [&](int g, const LVv& vxm, const LVv& vxp, const LVv& vym, const LVv& vyp) {
#pragma unroll
        for (int l = 0; l < LV; ++l) {
          const int i = g * LV + l;
          if (i < KH) {
            // A channel that is exactly zero on the whole stencil of every lane stays
            // exactly +0 (the full update evaluates to +0 there, finite coefficients given);
            // skip it wave-wide.
            if (__ballot(any_nz(c[i], vxm.v[l], vxp.v[l], vym.v[l], vyp.v[l], lo[i], hi[i], keep))) {
              res[i] = update_sr<TS, ANISO>(cf, B, ihx2, ihy2, ihz2, c[i], vxm.v[l], vxp.v[l], vym.v[l], vyp.v[l],
                                            lo[i], hi[i], ax + 12 * (ch0 + i), key + (uint32_t)i * kSrChannel);
            } else {
              res[i] = (TC)0;
            }
          }
        }
      };
      // FP32-arithmetic lane pairs of up to 10 channels per lane request the next group's four
      // neighbour vectors before this group's channels run (their branches would otherwise keep the
      // reads behind them): the LDS latency of a group overlaps the previous group's arithmetic
      // (256x256x64, K = 20 step: FP32 -0.3 %, BF16 -1.9 %, for 4-6 more VGPRs).  Elsewhere the extra live
      // vectors spill: one-lane tiles read 16-byte vectors (16 more VGPRs), FP64 pairs hold 8-byte
      // elements, and larger halves leave no room; there each group is read as it runs.
      if constexpr (H == 2 && sizeof(TC) == 4 && KH <= 10) {
        LVv nb[2][4];
        nb[0][0] = *(const LVv*)(cb + oxm);
        nb[0][1] = *(const LVv*)(cb + oxp);
        nb[0][2] = *(const LVv*)(cb + oym);
        nb[0][3] = *(const LVv*)(cb + oyp);
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          if (g + 1 < NG) {
            nb[(g + 1) & 1][0] = *(const LVv*)(cb + oxm + (g + 1) * LV);
            nb[(g + 1) & 1][1] = *(const LVv*)(cb + oxp + (g + 1) * LV);
            nb[(g + 1) & 1][2] = *(const LVv*)(cb + oym + (g + 1) * LV);
            nb[(g + 1) & 1][3] = *(const LVv*)(cb + oyp + (g + 1) * LV);
          }
          group(g, nb[g & 1][0], nb[g & 1][1], nb[g & 1][2], nb[g & 1][3]);
        }
      } else {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const LVv vxm = *(const LVv*)(cb + oxm + g * LV);
          const LVv vxp = *(const LVv*)(cb + oxp + g * LV);
          const LVv vym = *(const LVv*)(cb + oym + g * LV);
          const LVv vyp = *(const LVv*)(cb + oyp + g * LV);
          group(g, vxm, vxp, vym, vyp);
        }
      }
      if constexpr (!G::RING3) {
#if PF_STORE_MODE == 0
        TS* o = (TS*)d.out + cell * K + ch0;
#else
        TS* o = stage + cl * K + ch0;
#endif
        // the store on each side of the liquid branch, so that the common side stores the result
        // registers as they are (one store after the join costs a register copy of every result)
        if (__ballot(cf.liquid)) {
          TC rl[KH];
#pragma unroll
          for (int i = 0; i < KH; ++i) rl[i] = cf.liquid ? (TC)0 : res[i];
          store_cell(o, rl);
        } else {
          store_cell(o, res);
        }
      } else if (__ballot(cf.liquid)) {
#pragma unroll
        for (int i = 0; i < KH; ++i) res[i] = cf.liquid ? (TC)0 : res[i];
      }
    }
    if constexpr (G::RING3) {
      // Plane z+2 into tb and plane z+3 requested before this plane's stores are issued (loads
      // and stores share one completion counter), then the results go out from registers.
      if (n + 1 < len) {
        store_plane(tb);
        if (n + 2 < len) load_plane(zstep(z, 3));
      }
      if (active) store_cell((TS*)d.out + cell * K + ch0, res);
      lds_barrier();
      if (n + 1 < len) read_center(tb, lo);
    } else {
      lds_barrier();
      // Plane z+2 (prefetched one plane ago) replaces plane z and plane z+3 is requested before
      // this plane's row stores are issued: loads and stores share one completion counter, so
      // a wait for the prefetched data (or for its registers) placed after the stores would
      // also wait for the stores.
      if (n + 1 < len) {
        store_plane(cb);
        if (n + 2 < len) load_plane(zstep(z, 3));
      }
#if PF_STORE_MODE != 0
      store_rows(z);
#endif
      if (n + 1 < len) {
        lds_barrier();
        read_center(cb, lo);
      }
    }
  }

  // This thread's KH results of one cell to o (global memory or the LDS staging rows).
  __device__ __forceinline__ void store_cell(TS* o, const TC (&res)[KH]) {
#pragma unroll
    for (int j = 0; j < KH / G::SVS; ++j) {
      SVv s;
#pragma unroll
      for (int e = 0; e < G::SVS; ++e) s.v[e] = from_tc<TS>(res[j * G::SVS + e]);
      *(SVv*)(o + j * G::SVS) = s;
    }
  }

  // Output plane z as a raw buffer: a store at offset kOob or beyond is dropped by the
  // hardware's range check (the plane is under 2^31 bytes).
  __device__ __forceinline__ __amdgpu_buffer_rsrc_t out_plane(int z) const {
    return __builtin_amdgcn_make_buffer_rsrc((TS*)d.out + (long long)z * plane_elems(), 0,
                                             (int)(plane_elems() * (long long)sizeof(TS)), kRawBufferDword3);
  }

  // Coalesced write-back of plane z of the tile from the LDS staging rows: consecutive
  // lanes store consecutive CH-byte pieces of each contiguous (y, z) row segment.  BUF_ROWS:
  // every lane issues one buffer store per pass (pieces never straddle a row's end unless
  // WIDE), lanes without a piece at an out-of-range offset, so that every wave has the same
  // stores in flight however the domain's faces cut the tile, and the compiler's waits at the
  // next hand-over cover the prefetched plane only.  Otherwise lanes without a piece skip
  // their store, and since the compiler takes the path with the fewest operations in flight,
  // the next hand-over also waits for this plane's stores.
  __device__ __forceinline__ void store_rows(int z) {
    using CT = typename RawT<CH>::type;
    if constexpr (BUF_ROWS) {
      const __amdgpu_buffer_rsrc_t r = out_plane(z);
#pragma unroll
      for (int j = 0; j < NJO; ++j) {
        const int w = min((int)threadIdx.x + j * G::NT, NCHO - 1);
        buffer_store<CH, PF_STORE_MODE == 2>((const char*)stage + w * CH, r, opos[j]);
      }
    } else {
      using GT = typename std::conditional<WIDE, raw16a2_t, CT>::type;
      using UT = typename RawT<U>::type;
      const int rowb = min(TX, a.nx - x0) * CELLB;
      TS* out = (TS*)d.out + (long long)z * plane_elems();
#pragma unroll
      for (int j = 0; j < NJO; ++j) {
        if (opos[j] != kNoPiece) {
          const int w = threadIdx.x + j * G::NT;
          const int off = (w % CPR) * CH;
          const char* src = (const char*)stage + w * CH;  // = row (w / CPR) of the staging rows, byte off
          char* dst = (char*)out + opos[j];
          if (!WIDE || off + CH <= rowb) {
            const GT v = *(const CT*)src;
#if PF_STORE_MODE == 2
            __builtin_nontemporal_store(v, (GT*)dst);
#else
            *(GT*)dst = v;
#endif
          } else {
            for (int q = 0; q < rowb - off; q += U) *(UT*)(dst + q) = *(const UT*)(src + q);
          }
        }
      }
    }
  }

  // The stores of one plane's write-back, all dropped (offset kOob): issued after the
  // prologue's prefetch, they give the z loop's first hand-over the same loads and stores in
  // flight as every later one, so that the compiler computes one exact set of waits for all.
  __device__ __forceinline__ void dropped_stores() {
    if constexpr (BUF_ROWS) {
      const __amdgpu_buffer_rsrc_t r = out_plane(zb);
      const unsigned zero[4] __attribute__((aligned(16))) = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int j = 0; j < NJO; ++j) buffer_store<CH, PF_STORE_MODE == 2>(zero, r, kOob + (unsigned)(j * CH));
    }
  }

  __device__ __forceinline__ void run() {
    TC r0[KH], r1[KH], r2[KH];
    // prologue: plane zb-1 (edge-replicated) straight into registers, then planes zb, zb+1
    // into LDS.  The register loads go first so that the waits for the LDS planes' data also
    // retire them: no load of the prologue is then outstanding when the z loop starts, and the
    // loop's waits only concern the prefetched plane.
    const int z0 = zb;
    {
      const int zbh = zstep(z0, -1);
      const int sxo = min(gx, a.nx - 1), syo = min(gy, a.ny - 1);
      const TS* p =
          (const TS*)d.in + (long long)zbh * plane_elems() + (long long)(syo * a.nx + sxo) * K + ch0;
      using HV = VecA<TS, VEC, (int)sizeof(TS)>;
#pragma unroll
      for (int j = 0; j < KH / VEC; ++j) {
        const HV s = *(const HV*)(p + j * VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) r0[j * VEC + e] = to_tc(s.v[e]);
      }
#pragma unroll
      for (int e = (KH / VEC) * VEC; e < KH; ++e) r0[e] = to_tc(p[e]);
    }
    load_plane(z0);
    store_plane(buf0);
    load_plane(zstep(z0, 1));
    store_plane(buf1);
    // nothing of the prologue stays in flight into the z loop (the compiler may issue the
    // plane z-1 loads after the LDS planes' loads)
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    lds_barrier();
    read_center(buf0, r1);
    read_center(buf1, r2);
    if (len > 1) {
      load_plane(zstep(z0, 2));
      dropped_stores();
    }
    for (int n = 0; n < len; n += 3) {
      if constexpr (G::RING3) {  // plane zb + m lives in buffer m mod 3
        body(n, r0, r1, r2, buf0, buf2);
        if (n + 1 < len) body(n + 1, r1, r2, r0, buf1, buf0);
        if (n + 2 < len) body(n + 2, r2, r0, r1, buf2, buf1);
      } else {  // plane zb + m lives in buffer m mod 2
        TC* const e = (n & 1) ? buf1 : buf0;
        TC* const o = (n & 1) ? buf0 : buf1;
        body(n, r0, r1, r2, e, e);
        if (n + 1 < len) body(n + 1, r1, r2, r0, o, o);
        if (n + 2 < len) body(n + 2, r2, r0, r1, e, e);
      }
    }
  }
};

template <int K, class TS, bool ANISO, int TX, int TY, int H>
__device__ __forceinline__ void march_tile(const StepDev& a, const StepDyn& d, int tile_x, int tile_y,
                                           int tile_z, typename Traits<TS>::TC* __restrict__ smem,
                                           TS* __restrict__ stage, typename Traits<TS>::TC* __restrict__ ax,
                                           ZTab* __restrict__ zt) {
  March<K, TS, ANISO, TX, TY, H> m(a, d, tile_x, tile_y, tile_z, smem, stage, ax, zt);
  m.run();
}

// ------------------------------------------------------------------ generic-K march
// One thread per cell, K at run time, neighbours straight from global memory.  Used for
// K without a register-tiled instantiation.  Same per-cell and per-channel arithmetic
// as march_tile.  RHS = true writes state_rhs instead of the Euler update.
template <class TS, bool ANISO, bool RHS = false>
__device__ __forceinline__ void generic_cell(const StepDev& a, const StepDyn& dy, long long cell) {
  using TC = typename Traits<TS>::TC;
  const int nx = a.nx, ny = a.ny, nz = a.nz, K = a.k;
  const int ix = (int)(cell % nx);
  const int iy = (int)((cell / nx) % ny);
  const int iz = (int)(cell / ((long long)nx * ny));
  const TS* __restrict__ in = (const TS*)dy.in;
  TS* __restrict__ out = (TS*)dy.out;
  const long long sx = K, sy = (long long)nx * K, sz = (long long)nx * ny * K;
  const long long base = cell * K;
  const long long oxm = ix > 0 ? -sx : 0, oxp = ix < nx - 1 ? sx : 0;
  const long long oym = iy > 0 ? -sy : 0, oyp = iy < ny - 1 ? sy : 0;
  const long long ozm = iz > 0 ? -sz : 0, ozp = iz < nz - 1 ? sz : 0;
  TC slo = 0, shi = 0;  // the two chains of sum_sq_halves
  for (int i = 0; i < K / 2; ++i) {
    const TC c = to_tc(in[base + i]);
    slo = fmac(c, c, slo);
  }
  for (int i = K / 2; i < K; ++i) {
    const TC c = to_tc(in[base + i]);
    shi = fmac(c, c, shi);
  }
  const TC S = sum_sq_halves(slo, shi);
  const CellCoef<TC> cf = cell_coef<TC>(a, dy.cold, ix, iy, iz, cell, dy.t_prev, dy.hstep, S);
  const TC B = lap_B<TC>(a);
  TC ihx2, ihy2, ihz2;
  inv_h2<TC>(a, ihx2, ihy2, ihz2);
  const TC* __restrict__ axes = (const TC*)a.axes;
  const uint32_t keep = coef_nonfinite(cf) ? 1u : 0u;
  const uint32_t key = sr_cell_key(sr_plane_key(dy.seed, iz), (int)(cell - (long long)iz * nx * ny));
  for (int i = 0; i < K; ++i) {
    TC ax[12];
    if constexpr (ANISO) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) ax[4 * k + j] = axes[(k * K + i) * 3 + j];
    }
    const long long e = base + i;
    const TC c = to_tc(in[e]), xm = to_tc(in[e + oxm]), xp = to_tc(in[e + oxp]), ym = to_tc(in[e + oym]),
             yp = to_tc(in[e + oyp]), zm = to_tc(in[e + ozm]), zp = to_tc(in[e + ozp]);
    TC v;
    if (RHS || __ballot(any_nz(c, xm, xp, ym, yp, zm, zp, keep))) {
      if constexpr (RHS)
        v = update_one<TC, ANISO, true>(cf, B, ihx2, ihy2, ihz2, c, xm, xp, ym, yp, zm, zp, ax);
      else
        v = update_sr<TS, ANISO>(cf, B, ihx2, ihy2, ihz2, c, xm, xp, ym, yp, zm, zp, ax, key + (uint32_t)i * kSrChannel);
    } else {
      v = (TC)0;
    }
    if (!RHS && cf.liquid) v = (TC)0;
    out[e] = from_tc<TS>(v);
  }
}

// Tile order of a launch: layers of z-chunks from the top of the domain down.  Workgroups are
// dispatched in index order and dealt round-robin to the 8 XCDs, so the top layer, where a
// surface heat source makes cells hot and the per-cell FP64 path slow, starts first and the
// cheaper layers below fill in behind it.  Within a layer, when the tile rows divide by 8,
// workgroup t goes to tile row 8m + (t mod 8): every XCD then walks whole rows of tiles, so
// the x-neighbours of a tile (whose halo columns it re-reads) run on the same XCD and share its
// L2, and each XCD still gets an equal share of every layer.  Otherwise x fastest, then y.
__device__ __forceinline__ void tile_coords(int t, int tiles_x, int tiles_y, int tiles_z, int& tx, int& ty,
                                            int& tz) {
  const int nl = tiles_x * tiles_y;
  const int layer = t / nl, r = t - layer * nl;
  tz = tiles_z - 1 - layer;
  if (tiles_y % 8 == 0) {
    const int k = r >> 3;
    const int m = k / tiles_x;
    tx = k - m * tiles_x;
    ty = 8 * m + (r & 7);
  } else {
    tx = r % tiles_x;
    ty = r / tiles_x;
  }
}

}  // namespace pf
