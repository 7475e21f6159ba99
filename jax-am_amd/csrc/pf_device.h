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
// Per-cell quantities (temperature, liquid indicator zeta, Arrhenius mobility,
// liquid reset) are evaluated in FP64 for every storage dtype so that the
// liquid/solid decision matches an FP64 run; per-orientation arithmetic is done
// in the compute type TC (float for F32/BF16 storage, double for F64).
// All device code is built with -ffp-contract=off: every fused multiply-add
// below is written out, so each kernel that shares these functions produces
// bitwise the same result for the same cell.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pfhip.h"

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
// Everything a tile march needs about one instance.  Built on the host from
// pf_params_t (pf_host.h) and passed by value (single step) or read from device
// memory (batched engine).
struct StepDev {
  const void* in;
  void* out;
  const void* T;       // PF_T_BUFFER temperature
  const void* axes;    // [3][K][3] in TC
  int nz, ny, nx, k;
  int tiles_x, tiles_y, tiles_z, zc;
  int tmodel, t_f64, force_liquid;
  double hx, hy, hz;
  double ihx2, ihy2, ihz2;
  double m_over_ad, kappa_ad, gamma4, B;  // m_g/ad_hoc, kappa_g*ad_hoc, 4*gamma, 1-4*gamma
  double L0, Qg, gas_const, T_l, T_clip, theta, zeta_window, eps, liquid_T;
  double t_prev, hstep;
  // laser
  double T_amb, p0x, p0y, p0z, dirx, diry, dirz, vel;
  double q_coef, pe, T_cap;          // Rosenthal: Q/(2 pi kappa_T), vel/(2 alpha)
  double dT_peak, g_sxy, g_sz;       // Gaussian: 1/(2 sigma_xy^2), 1/(2 sigma_z^2)
  double T_ref, grad_T, cool_rate;   // quench
};

// ------------------------------------------------------------------ per-cell physics (FP64)
// Temperature at the centroid of cell (ix, iy, iz) at time t; cell index c for the buffer mode.
__device__ __forceinline__ double cell_temperature(const StepDev& a, int ix, int iy, int iz,
                                                   long long c, double t) {
  if (a.tmodel == PF_T_BUFFER) {
    return a.t_f64 ? ((const double*)a.T)[c] : (double)((const float*)a.T)[c];
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
  return a.T_amb + a.dT_peak * exp(-((X * X + Y * Y) * a.g_sxy + Z * Z * a.g_sz));
}

template <class TC> struct CellCoef {
  TC c1, c2, A, w1, w2;  // w1 = c2*(1+eps), w2 = 2*eps*c2 (anisotropic factor folded into c2)
  bool liquid;
};

// Per-cell coefficients of the update (allen_cahn.py:55-58,83-84,124-126).
template <class TC>
__device__ __forceinline__ CellCoef<TC> cell_coef(const StepDev& a, double T, TC S) {
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
  const double hL = a.hstep * Lg;
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

// One orientation channel: centre value c, x/y/z neighbours (edge-replicated at the
// domain faces by the caller), returns eta_new.
// With RHS = true it returns -L*(grad + laplace) instead (c1, c2 then carry no dt and the
// liquid reset is not applied), the reference's state_rhs value.
template <class TC, bool ANISO, bool RHS = false>
__device__ __forceinline__ TC update_one(const CellCoef<TC>& cf, TC B, TC ihx2, TC ihy2, TC ihz2,
                                         TC c, TC xm, TC xp, TC ym, TC yp, TC zm, TC zp,
                                         const TC* __restrict__ ax /* 9 values: axis k, comp d at [3k+d] */) {
  const TC dx = fmac(c, (TC)-2, xm + xp);
  const TC dy = fmac(c, (TC)-2, ym + yp);
  const TC dz = fmac(c, (TC)-2, zm + zp);
  const TC lap = fmac(dz, ihz2, fmac(dy, ihy2, dx * ihx2));
  TC ca;
  if constexpr (ANISO) {
    const TC ex = xp - xm, ey = yp - ym, ez = zp - zm;
    const TC n2 = fmac(ez, ez, fmac(ey, ey, ex * ex));
    const TC p0 = fmac(ax[2], ez, fmac(ax[1], ey, ax[0] * ex));
    const TC p1 = fmac(ax[5], ez, fmac(ax[4], ey, ax[3] * ex));
    const TC p2 = fmac(ax[8], ez, fmac(ax[7], ey, ax[6] * ex));
    TC m;
    TC q;
    if constexpr (sizeof(TC) == 4) {
      m = fmaxf(fmaxf(fabsf(p0), fabsf(p1)), fabsf(p2));
      q = fminf(m * m * rcp_approx(n2), 1.0f);   // n2 == 0 -> NaN -> 1 (minNum)
    } else {
      m = fmax(fmax(fabs(p0), fabs(p1)), fabs(p2));
      q = n2 > 0.0 ? fmin((m * m) / n2, 1.0) : 1.0;
    }
    ca = fmac(cf.w2 * q, q - (TC)1, cf.w1);
  } else {
    ca = cf.c2;
  }
  const TC e2 = c * c;
  const TC t = fmac(e2, B, cf.A);
  const TC u = c * t;
  if constexpr (RHS) {
    return fmac(ca, lap, -cf.c1 * u);
  } else {
    const TC r = fmac(-cf.c1, u, c);
    const TC v = fmac(ca, lap, r);
    return cf.liquid ? (TC)0 : v;
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------------------ tile geometry
template <class TS, int K> struct Layout {
  using TC = typename Traits<TS>::TC;
  // elements per global load chunk (must divide K; at most 16 bytes of storage and 16 bytes of TC)
  static constexpr int VEC_MAX = (16 / (int)sizeof(TC)) < (16 / (int)sizeof(TS)) ? (16 / (int)sizeof(TC))
                                                                                   : (16 / (int)sizeof(TS));
  static constexpr int VEC = (K % VEC_MAX == 0) ? VEC_MAX
                             : (VEC_MAX >= 2 && K % 2 == 0) ? 2 : 1;
  // LDS cell stride: 16-byte groups, odd number of groups per cell (bank-conflict-free b128 reads)
  static constexpr int LV = 16 / (int)sizeof(TC);
  static constexpr int KG = (K + LV - 1) / LV;
  static constexpr int KP = (KG % 2 == 1 ? KG : KG + 1) * LV;
  static constexpr int NG = KG;  // groups read per neighbour
};

template <class TS, int K, int TX, int TY> struct Tile {
  using L = Layout<TS, K>;
  using TC = typename L::TC;
  static constexpr int NT = TX * TY;
  static constexpr int PX = TX + 2, PY = TY + 2;
  static constexpr int CPC = K / L::VEC;         // chunks per cell
  static constexpr int NCH = PY * PX * CPC;      // chunks per plane
  static constexpr int CPT = (NCH + NT - 1) / NT;
  static constexpr int PLANE = PY * PX * L::KP;  // TC elements per LDS plane
  static constexpr size_t LDS_BYTES = 2 * PLANE * sizeof(TC);
};

// ------------------------------------------------------------------ 2.5-D march of one tile
// A workgroup of TX*TY threads owns a TX x TY column of cells and marches it
// upward through z-planes [zb, ze).  Planes (tile + one-cell halo ring,
// edge-replicated at the domain faces) are double-buffered in LDS; the plane two
// ahead is prefetched into registers while the current plane is computed.  Each
// thread keeps the K channels of its own cell at z-1, z, z+1 in three register
// arrays whose roles rotate by renaming (the z loop is unrolled by three).
template <int K, class TS, bool ANISO, int TX, int TY>
struct March {
  using TC = typename Traits<TS>::TC;
  using L = Layout<TS, K>;
  using G = Tile<TS, K, TX, TY>;
  static constexpr int VEC = L::VEC, KP = L::KP, LV = L::LV, NG = L::NG;
  using SV = Vec<TS, VEC>;
  using CV = Vec<TC, VEC>;
  using LVv = Vec<TC, LV>;

  const StepDev a;
  TC* __restrict__ buf0;
  TC* __restrict__ buf1;
  const TS* __restrict__ in;
  TS* __restrict__ out;
  int src_off[G::CPT];
  int dst_off[G::CPT];
  SV pre[G::CPT];
  int gx, gy, own, oxm, oxp, oym, oyp, zb, len;
  long long plane_elems;
  bool active;

  __device__ __forceinline__ March(const StepDev& a_, int tile_x, int tile_y, int tile_z, TC* smem) : a(a_) {
    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    gx = tile_x * TX + lx;
    gy = tile_y * TY + ly;
    active = (gx < a.nx) && (gy < a.ny);
    zb = tile_z * a.zc;
    len = min(zb + a.zc, a.nz) - zb;
    plane_elems = (long long)a.ny * a.nx * K;
    in = (const TS*)a.in;
    out = (TS*)a.out;
    buf0 = smem;
    buf1 = smem + G::PLANE;
#pragma unroll
    for (int j = 0; j < G::CPT; ++j) {
      const int w = tid + j * G::NT;
      const int ww = w < G::NCH ? w : G::NCH - 1;
      const int cs = ww / G::CPC, v = ww % G::CPC;
      const int r = cs / G::PX, cc = cs % G::PX;
      const int sy = clampi(tile_y * TY - 1 + r, 0, a.ny - 1);
      const int sx = clampi(tile_x * TX - 1 + cc, 0, a.nx - 1);
      src_off[j] = (sy * a.nx + sx) * K + v * VEC;
      dst_off[j] = (w < G::NCH) ? cs * KP + v * VEC : -1;
    }
    own = ((ly + 1) * G::PX + (lx + 1)) * KP;
    oxm = own - KP;
    oxp = own + KP;
    oym = own - G::PX * KP;
    oyp = own + G::PX * KP;
  }

  __device__ __forceinline__ void load_plane(int z) {
    const TS* p = in + (long long)z * plane_elems;
#pragma unroll
    for (int j = 0; j < G::CPT; ++j) pre[j] = *(const SV*)(p + src_off[j]);
  }
  __device__ __forceinline__ void store_plane(TC* b) {
#pragma unroll
    for (int j = 0; j < G::CPT; ++j) {
      if (dst_off[j] >= 0) {
        CV c;
#pragma unroll
        for (int e = 0; e < VEC; ++e) c.v[e] = to_tc(pre[j].v[e]);
        *(CV*)(b + dst_off[j]) = c;
      }
    }
  }
  __device__ __forceinline__ void read_center(const TC* b, TC (&dst)[K]) {
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const LVv v = *(const LVv*)(b + own + g * LV);
#pragma unroll
      for (int l = 0; l < LV; ++l)
        if (g * LV + l < K) dst[g * LV + l] = v.v[l];
    }
  }

  // Compute plane zb+n from (lo, c, hi) = planes z-1, z, z+1, then (if another plane
  // follows) rotate the LDS ring and read plane z+2 into `lo`.
  __device__ __forceinline__ void body(int n, TC (&lo)[K], TC (&c)[K], TC (&hi)[K]) {
    const int z = zb + n;
    TC* cb = (n & 1) ? buf1 : buf0;
    if (active) {
      const long long cell = ((long long)z * a.ny + gy) * a.nx + gx;
      TC S = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) S = fmac(c[i], c[i], S);
      const double T = cell_temperature(a, gx, gy, z, cell, a.t_prev);
      const CellCoef<TC> cf = cell_coef<TC>(a, T, S);
      const TC B = (TC)a.B, ihx2 = (TC)a.ihx2, ihy2 = (TC)a.ihy2, ihz2 = (TC)a.ihz2;
      const TC* __restrict__ axes = (const TC*)a.axes;
      if constexpr (ANISO) asm volatile("" : "+s"(axes));  // keep the axis table out of VGPRs: reload per plane
      TC res[K];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const LVv vxm = *(const LVv*)(cb + oxm + g * LV);
        const LVv vxp = *(const LVv*)(cb + oxp + g * LV);
        const LVv vym = *(const LVv*)(cb + oym + g * LV);
        const LVv vyp = *(const LVv*)(cb + oyp + g * LV);
#pragma unroll
        for (int l = 0; l < LV; ++l) {
          const int i = g * LV + l;
          if (i < K) {
            TC ax[9];
            if constexpr (ANISO) {
#pragma unroll
              for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                for (int d = 0; d < 3; ++d) ax[3 * kk + d] = axes[(kk * K + i) * 3 + d];
            }
            res[i] = update_one<TC, ANISO>(cf, B, ihx2, ihy2, ihz2, c[i], vxm.v[l], vxp.v[l], vym.v[l],
                                           vyp.v[l], lo[i], hi[i], ax);
          }
        }
      }
      TS* o = out + cell * K;
#pragma unroll
      for (int j = 0; j < K / VEC; ++j) {
        SV s;
#pragma unroll
        for (int e = 0; e < VEC; ++e) s.v[e] = from_tc<TS>(res[j * VEC + e]);
        *(SV*)(o + j * VEC) = s;
      }
    }
    if (n + 1 < len) {
      __syncthreads();
      store_plane(cb);  // plane z+2 replaces plane z
      if (n + 2 < len) load_plane(min(z + 3, a.nz - 1));
      __syncthreads();
      read_center(cb, lo);
    }
  }

  __device__ __forceinline__ void run() {
    TC r0[K], r1[K], r2[K];
    // prologue: planes zb, zb+1 into LDS; plane zb-1 (edge-replicated) straight into registers
    load_plane(zb);
    store_plane(buf0);
    load_plane(min(zb + 1, a.nz - 1));
    store_plane(buf1);
    {
      const int zbh = max(zb - 1, 0);
      const int sxo = min(gx, a.nx - 1), syo = min(gy, a.ny - 1);
      const TS* p = in + (long long)zbh * plane_elems + (long long)(syo * a.nx + sxo) * K;
#pragma unroll
      for (int j = 0; j < K / VEC; ++j) {
        const SV s = *(const SV*)(p + j * VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) r0[j * VEC + e] = to_tc(s.v[e]);
      }
    }
    __syncthreads();
    read_center(buf0, r1);
    read_center(buf1, r2);
    if (len > 1) load_plane(min(zb + 2, a.nz - 1));
    for (int n = 0; n < len; n += 3) {
      body(n, r0, r1, r2);
      if (n + 1 < len) body(n + 1, r1, r2, r0);
      if (n + 2 < len) body(n + 2, r2, r0, r1);
    }
  }
};

template <int K, class TS, bool ANISO, int TX, int TY>
__device__ __forceinline__ void march_tile(const StepDev& a, int tile_x, int tile_y, int tile_z,
                                           typename Traits<TS>::TC* __restrict__ smem) {
  March<K, TS, ANISO, TX, TY> m(a, tile_x, tile_y, tile_z, smem);
  m.run();
}

// ------------------------------------------------------------------ generic-K march
// One thread per cell, K at run time, neighbours straight from global memory.  Used for
// K without a register-tiled instantiation.  Same per-cell and per-channel arithmetic
// as march_tile.  RHS = true writes state_rhs instead of the Euler update.
template <class TS, bool ANISO, bool RHS = false>
__device__ __forceinline__ void generic_cell(const StepDev& a, long long cell) {
  using TC = typename Traits<TS>::TC;
  const int nx = a.nx, ny = a.ny, nz = a.nz, K = a.k;
  const int ix = (int)(cell % nx);
  const int iy = (int)((cell / nx) % ny);
  const int iz = (int)(cell / ((long long)nx * ny));
  const TS* __restrict__ in = (const TS*)a.in;
  TS* __restrict__ out = (TS*)a.out;
  const long long sx = K, sy = (long long)nx * K, sz = (long long)nx * ny * K;
  const long long base = cell * K;
  const long long oxm = ix > 0 ? -sx : 0, oxp = ix < nx - 1 ? sx : 0;
  const long long oym = iy > 0 ? -sy : 0, oyp = iy < ny - 1 ? sy : 0;
  const long long ozm = iz > 0 ? -sz : 0, ozp = iz < nz - 1 ? sz : 0;
  TC S = 0;
  for (int i = 0; i < K; ++i) {
    const TC c = to_tc(in[base + i]);
    S = fmac(c, c, S);
  }
  const double T = cell_temperature(a, ix, iy, iz, cell, a.t_prev);
  const CellCoef<TC> cf = cell_coef<TC>(a, T, S);
  const TC B = (TC)a.B, ihx2 = (TC)a.ihx2, ihy2 = (TC)a.ihy2, ihz2 = (TC)a.ihz2;
  const TC* __restrict__ axes = (const TC*)a.axes;
  for (int i = 0; i < K; ++i) {
    const long long e = base + i;
    TC ax[9];
    if constexpr (ANISO) {
      for (int kk = 0; kk < 3; ++kk)
        for (int d = 0; d < 3; ++d) ax[3 * kk + d] = axes[(kk * K + i) * 3 + d];
    }
    const TC v = update_one<TC, ANISO, RHS>(cf, B, ihx2, ihy2, ihz2, to_tc(in[e]), to_tc(in[e + oxm]),
                                       to_tc(in[e + oxp]), to_tc(in[e + oym]), to_tc(in[e + oyp]),
                                       to_tc(in[e + ozm]), to_tc(in[e + ozp]), ax);
    out[e] = from_tc<TS>(v);
  }
}

// XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so give each
// XCD a contiguous range of logical tiles (x fastest, then y, then z) to keep the
// halo reuse of neighbouring tiles inside one L2.
__device__ __forceinline__ int xcd_swizzle(int b, int n) {
  constexpr int NX = 8;
  const int per = n / NX, rem = n % NX;
  const int x = b % NX, l = b / NX;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

}  // namespace pf
