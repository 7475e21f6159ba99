// C ABI of libpfhip.so (include/pfhip.h): single-instance step, temperature,
// initial condition, argmax and FP64 diagnostics.  The batched engine is in
// pf_batch.hip, the register-tiled instantiations in pf_inst_*.hip.
#include <math.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "pf_host.h"

namespace pf {

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  return PF_E_HIP;
}

size_t dtype_size(int dtype) { return dtype == PF_F64 ? 8 : dtype == PF_F32 ? 4 : dtype == PF_BF16 ? 2 : 0; }
bool valid_dtype(int dtype) { return dtype == PF_F64 || dtype == PF_F32 || dtype == PF_BF16; }

// ------------------------------------------------------------------ parameters
int make_stepdev(const pf_params_t* p, int dtype, const void* T_dev, const void* axes_dev, int zc, StepDev& a) {
  if (!p) return fail(PF_E_ARG, "params pointer is NULL");
  if (!valid_dtype(dtype)) return fail(PF_E_ARG, "dtype must be PF_F64, PF_F32 or PF_BF16");
  if (p->nz < 1 || p->ny < 1 || p->nx < 1 || p->num_oris < 1)
    return fail(PF_E_ARG, "grid dimensions and num_oris must be >= 1");
  const long long n = (long long)p->nz * p->ny * p->nx;
  if (n * p->num_oris >= (1LL << 40)) return fail(PF_E_ARG, "grid too large");
  if ((long long)p->ny * p->nx * p->num_oris >= (1LL << 31))
    return fail(PF_E_ARG, "one z-plane must hold fewer than 2^31 values");
  for (int d = 0; d < 3; ++d)
    if (!(p->h[d] > 0.0)) return fail(PF_E_ARG, "mesh_h_xyz must be positive");
  if (!(p->ad_hoc != 0.0)) return fail(PF_E_ARG, "ad_hoc must be non-zero");
  if (!(p->T_liquidus > 0.0) || !(p->gas_const != 0.0)) return fail(PF_E_ARG, "T_liquidus and gas_const must be set");
  const pf_laser_t& L = p->laser;
  if (L.model < PF_T_BUFFER || L.model > PF_T_QUENCH) return fail(PF_E_ARG, "unknown temperature model");
  if (L.model == PF_T_BUFFER) {
    if (!T_dev) return fail(PF_E_ARG, "temperature model PF_T_BUFFER needs a T buffer");
    if (L.T_dtype != PF_F64 && L.T_dtype != PF_F32) return fail(PF_E_ARG, "T buffer dtype must be PF_F64 or PF_F32");
  }
  if (L.model == PF_T_ROSENTHAL && (!(L.kappa_T != 0.0) || !(L.alpha != 0.0)))
    return fail(PF_E_ARG, "Rosenthal model needs kappa_T and alpha");
  if (L.model == PF_T_GAUSSIAN && (!(L.sigma_xy > 0.0) || !(L.sigma_z > 0.0)))
    return fail(PF_E_ARG, "Gaussian model needs sigma_xy, sigma_z > 0");
  const bool aniso = p->anisotropy > 0.0;
  if (aniso && !axes_dev) return fail(PF_E_ARG, "anisotropy > 0 needs the grain-direction table");

  memset(&a, 0, sizeof(a));
  a.T = T_dev;
  a.axes = axes_dev;
  a.nz = p->nz;
  a.ny = p->ny;
  a.nx = p->nx;
  a.k = p->num_oris;
  a.zc = zc;
  a.tmodel = L.model;
  a.t_f64 = L.T_dtype == PF_F64;
  a.force_liquid = p->force_liquid != 0;
  a.hx = p->h[0];
  a.hy = p->h[1];
  a.hz = p->h[2];
  a.ihx2 = 1.0 / (p->h[0] * p->h[0]);
  a.ihy2 = 1.0 / (p->h[1] * p->h[1]);
  a.ihz2 = 1.0 / (p->h[2] * p->h[2]);
  a.m_over_ad = p->m_g / p->ad_hoc;
  a.kappa_ad = p->kappa_g * p->ad_hoc;
  a.gamma4 = 4.0 * p->gamma;
  a.B = 1.0 - 4.0 * p->gamma;
  a.ihx2f = (float)a.ihx2;
  a.ihy2f = (float)a.ihy2;
  a.ihz2f = (float)a.ihz2;
  a.Bf = (float)a.B;
  a.L0 = p->L0;
  a.Qg = p->Qg;
  a.gas_const = p->gas_const;
  a.T_l = p->T_liquidus;
  a.T_clip = p->T_clip;
  a.theta = p->zeta_sharpness;
  // |theta*(Tc/T_l - 1)| >= 20 (tanh == +-1 exactly in FP64) holds for |Tc - T_l| > 20*T_l/theta;
  // use twice that as the compare window.
  a.zeta_window = p->zeta_sharpness > 0.0 ? 40.0 * p->T_liquidus / p->zeta_sharpness : INFINITY;
  a.eps = aniso ? p->anisotropy : 0.0;
  a.liquid_T = p->T_liquidus + p->liquid_shift;
  a.stiff = fabs(a.m_over_ad) * (4.0 + 4.0 * fabs(p->gamma)) +
            4.0 * fabs(a.kappa_ad) * (1.0 + a.eps) * (a.ihx2 + a.ihy2 + a.ihz2);
  a.T_amb = L.T_ambient;
  a.p0x = L.pos0[0];
  a.p0y = L.pos0[1];
  a.p0z = L.pos0[2];
  a.dirx = L.dir[0];
  a.diry = L.dir[1];
  a.dirz = L.dir[2];
  a.vel = L.vel;
  if (L.model == PF_T_ROSENTHAL) {
    a.kappa_T = L.kappa_T;
    a.alpha = L.alpha;
    a.q_coef = rosenthal_q(L.Q, L.kappa_T);  // simple/example.py:35 Q / (2*pi*kappa)
    a.pe = rosenthal_pe(L.vel, L.alpha);     // vel / (2*alpha)
  }
  a.T_cap = L.T_cap;
  a.dT_peak = L.dT_peak;
  if (L.model == PF_T_GAUSSIAN) {
    a.g_sxy = 1.0 / (2.0 * L.sigma_xy * L.sigma_xy);
    a.g_sz = 1.0 / (2.0 * L.sigma_z * L.sigma_z);
  }
  a.T_ref = L.T_ref;
  a.grad_T = L.grad_T;
  a.cool_rate = L.cool_rate;
  return PF_OK;
}

// ------------------------------------------------------------------ dispatch tables
namespace {
struct Tables {
  std::vector<StepEntry> t[3];
  Tables() {
    using PartFn = const StepEntry* (*)(int*);
    const PartFn parts[kInstParts] = {&step_table_part_0, &step_table_part_1, &step_table_part_2,
                                      &step_table_part_3, &step_table_part_4, &step_table_part_5,
                                      &step_table_part_6, &step_table_part_7, &step_table_part_8,
                                      &step_table_part_9, &step_table_part_10, &step_table_part_11};
    for (PartFn f : parts) {
      int n = 0;
      const StepEntry* e = f(&n);
      for (int i = 0; i < n; ++i) t[e[i].dtype].push_back(e[i]);
    }
  }
};
const Tables& tables() {
  static Tables tb;
  return tb;
}
}  // namespace

int entry_count(int dtype, int k, bool aniso) {
  if (!valid_dtype(dtype)) return 0;
  const Tables& tb = tables();
  int c = 0;
  for (const StepEntry& e : tb.t[dtype])
    if (e.k == k && (e.aniso != 0) == aniso) ++c;
  return c;
}

const StepEntry* find_entry(int dtype, int k, bool aniso, int tile) {
  if (!valid_dtype(dtype)) return nullptr;
  const Tables& tb = tables();
  if (tile < 0) tile = 0;
  for (const StepEntry& e : tb.t[dtype])
    if (e.k == k && (e.aniso != 0) == aniso && e.tile == tile) return &e;
  return nullptr;
}

int default_zchunk(int nz, int ny, int nx, int tx, int ty) {
  // about 2048 workgroups (four rounds of 2 per CU on 256 CUs: shorter marches balance the
  // data-dependent work better), at least 4 planes per march
  const long long txy = (long long)((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
  long long chunks = (2048 + txy / 2) / txy;
  if (chunks < 1) chunks = 1;
  int zc = (int)((nz + chunks - 1) / chunks);
  if (zc < 4) zc = 4;
  if (zc > nz) zc = nz;
  return zc;
}

// The tiled kernels (the march) address a z-plane with 32-bit byte offsets and write it
// back through raw buffer stores whose out-of-range offsets (2^31 and above) are dropped.
int march_fits(const pf_params_t* p, int dtype) {
  if ((long long)p->ny * p->nx * p->num_oris * (long long)dtype_size(dtype) >= (1LL << 31))
    return fail(PF_E_ARG, "the tiled kernels need a z-plane of fewer than 2^31 bytes (tile -2 runs the generic kernel)");
  return PF_OK;
}

void set_tiles(StepDev& a, int tx, int ty, int zchunk) {
  const bool tail = (zchunk & PF_ZCHUNK_TAIL) != 0;
  int zc = zchunk & ~PF_ZCHUNK_TAIL;
  if (zc <= 0) zc = default_zchunk(a.nz, a.ny, a.nx, tx, ty);
  if (zc > a.nz) zc = a.nz;
  a.zc = zc;
  a.zs = zc;
  a.zsplit = 0;
  if (tail) {
    // the lowest quarter of the planes in whole chunks of ceil(zc / 2), at least one full chunk above
    const int zs = (zc + 1) / 2;
    int nbot = (a.nz / 4 + zs - 1) / zs * zs;
    if (nbot > a.nz - zc) nbot = (a.nz - zc) / zs * zs;
    if (zs < zc && nbot > 0) {
      a.zs = zs;
      a.zsplit = nbot;
    }
  }
  a.tiles_x = (a.nx + tx - 1) / tx;
  a.tiles_y = (a.ny + ty - 1) / ty;
  a.tiles_z = (a.zsplit > 0 ? (a.zsplit + a.zs - 1) / a.zs : 0) + (a.nz - a.zsplit + zc - 1) / zc;
}

StepLaunch generic_step(int dtype, bool aniso) {
  switch (dtype) {
    case PF_F64: return aniso ? &launch_generic<double, true> : &launch_generic<double, false>;
    case PF_F32: return aniso ? &launch_generic<float, true> : &launch_generic<float, false>;
    case PF_BF16: return aniso ? &launch_generic<bf16_t, true> : &launch_generic<bf16_t, false>;
  }
  return nullptr;
}

static StepLaunch generic_rhs(int dtype, bool aniso) {
  switch (dtype) {
    case PF_F64: return aniso ? &launch_generic<double, true, true> : &launch_generic<double, false, true>;
    case PF_F32: return aniso ? &launch_generic<float, true, true> : &launch_generic<float, false, true>;
    case PF_BF16: return aniso ? &launch_generic<bf16_t, true, true> : &launch_generic<bf16_t, false, true>;
  }
  return nullptr;
}

BatchLaunch generic_batch(int dtype, bool aniso) {
  switch (dtype) {
    case PF_F64: return aniso ? &launch_batch_generic<double, true> : &launch_batch_generic<double, false>;
    case PF_F32: return aniso ? &launch_batch_generic<float, true> : &launch_batch_generic<float, false>;
    case PF_BF16: return aniso ? &launch_batch_generic<bf16_t, true> : &launch_batch_generic<bf16_t, false>;
  }
  return nullptr;
}

// ------------------------------------------------------------------ auxiliary kernels
template <class TT>
__global__ __launch_bounds__(256) void temperature_kernel(const StepDev a, TT* __restrict__ out, double t) {
  const long long n = (long long)a.nz * a.ny * a.nx;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int ix, iy, iz;
  if (n <= 0xffffffffLL) {  // 32-bit divisions: a fraction of the cost of the 64-bit ones
    const unsigned u = (unsigned)c, r = u / (unsigned)a.nx;
    ix = (int)(u - r * (unsigned)a.nx);
    iz = (int)(r / (unsigned)a.ny);
    iy = (int)(r - (unsigned)iz * (unsigned)a.ny);
  } else {
    ix = (int)(c % a.nx);
    iy = (int)((c / a.nx) % a.ny);
    iz = (int)(c / ((long long)a.nx * a.ny));
  }
  out[c] = (TT)cell_temperature(a, ix, iy, iz, c, t);
}

template <class TS>
__global__ __launch_bounds__(256) void onehot_kernel(TS* __restrict__ eta, const int32_t* __restrict__ ori,
                                                    long long n_cells, int k) {
  using TC = typename Traits<TS>::TC;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_cells * k) return;
  const long long c = e / k;
  const int i = (int)(e - c * k);
  eta[e] = from_tc<TS>(ori[c] == i ? (TC)1 : (TC)0);
}

// nearest seed of each cell centroid (seeds staged through LDS in blocks of 256)
__global__ __launch_bounds__(256) void voronoi_kernel(const double* __restrict__ seeds, int ns,
                                                     const int32_t* __restrict__ gori, int nz, int ny, int nx,
                                                     double hx, double hy, double hz, int32_t* __restrict__ out) {
  __shared__ double sx[256], sy[256], sz[256];
  const long long n = (long long)nz * ny * nx;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = (int)(c % nx), iy = (int)((c / nx) % ny), iz = (int)(c / ((long long)nx * ny));
  const double cx = ((double)ix + 0.5) * hx, cy = ((double)iy + 0.5) * hy, cz = ((double)iz + 0.5) * hz;
  double best = INFINITY;
  int bi = 0;
  for (int s0 = 0; s0 < ns; s0 += 256) {
    __syncthreads();
    if (s0 + (int)threadIdx.x < ns) {
      sx[threadIdx.x] = seeds[3 * (s0 + threadIdx.x) + 0];
      sy[threadIdx.x] = seeds[3 * (s0 + threadIdx.x) + 1];
      sz[threadIdx.x] = seeds[3 * (s0 + threadIdx.x) + 2];
    }
    __syncthreads();
    const int m = min(256, ns - s0);
    for (int j = 0; j < m; ++j) {
      const double dx = cx - sx[j], dy = cy - sy[j], dz = cz - sz[j];
      const double d = dx * dx + dy * dy + dz * dz;
      if (d < best) {
        best = d;
        bi = s0 + j;
      }
    }
  }
  if (c < n) out[c] = gori ? gori[bi] : bi;
}

// numpy.argmax semantics: first maximum; a NaN is the maximum (first NaN wins)
template <class TS>
__device__ __forceinline__ int cell_argmax(const TS* __restrict__ row, int k) {
  using TC = typename Traits<TS>::TC;
  TC best = to_tc(row[0]);
  int bi = 0;
  if (best != best) return 0;
  for (int i = 1; i < k; ++i) {
    const TC v = to_tc(row[i]);
    if (v != v) return i;
    if (v > best) {
      best = v;
      bi = i;
    }
  }
  return bi;
}

template <class TS>
__global__ __launch_bounds__(256) void argmax_kernel(const TS* __restrict__ eta, long long n_cells, int k,
                                                    int32_t* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  out[c] = cell_argmax(eta + c * k, k);
}

// ---- deterministic FP64 diagnostics
constexpr int RB = 256;          // threads per reduction block
constexpr int RMAXK = 256;       // histogram capacity (K <= RMAXK)
constexpr int NF = 4;            // F_local, F_grad, max T, sum eta
constexpr int NI = 3;            // changed, liquid, nonfinite

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ long long wave_sum_i(long long v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

static int reduce_blocks(long long n) {
  long long b = (n + RB * 8 - 1) / (RB * 8);
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (int)b;
}
static const long long kColRows = 4096;

template <class TS>
__global__ __launch_bounds__(RB) void reduce_cells_kernel(const StepDev a, const TS* __restrict__ eta,
                                                         const int32_t* __restrict__ argmax0, double t, double m_g,
                                                         double gamma, double ad_hoc, double half_kappa_ad,
                                                         long long per_block, double* __restrict__ part_f,
                                                         long long* __restrict__ part_i) {
  using TC = typename Traits<TS>::TC;
  __shared__ int hist[RMAXK];
  __shared__ double sf[RB / 64][NF];
  __shared__ long long si[RB / 64][NI];
  const int K = a.k;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long long n = (long long)nz * ny * nx;
  for (int i = threadIdx.x; i < K; i += RB) hist[i] = 0;
  __syncthreads();
  const long long c0 = (long long)blockIdx.x * per_block;
  const long long c1 = min(n, c0 + per_block);
  double f_loc = 0.0, f_grad = 0.0, tmax = -INFINITY, esum = 0.0;
  long long changed = 0, liquid = 0, nonfinite = 0;
  const double ih2[3] = {a.ihx2, a.ihy2, a.ihz2};
  for (long long c = c0 + threadIdx.x; c < c1; c += RB) {
    const int ix = (int)(c % nx), iy = (int)((c / nx) % ny), iz = (int)(c / ((long long)nx * ny));
    const TS* row = eta + c * K;
    double S = 0.0, s4 = 0.0, q = 0.0;
    for (int i = 0; i < K; ++i) {
      const double v = (double)to_tc(row[i]);
      const double v2 = v * v;
      S += v2;
      s4 += v2 * v2;
      q += v2 * v2 / 4.0 - v2 / 2.0;
      esum += v;
      if (!isfinite(v)) ++nonfinite;
    }
    const double T = cell_temperature(a, ix, iy, iz, c, t);
    const double Tc = T > a.T_clip ? a.T_clip : T;
    const double x = a.theta * (Tc / a.T_l - 1.0);
    const double zeta = 0.5 * (1.0 - tanh(x));
    const double omz = 1.0 - zeta;
    // allen_cahn.py:45-48 with sum_ij (eta_i eta_j)^2 = S^2
    f_loc += m_g * (q + gamma * (S * S - s4) + omz * omz * S) / ad_hoc;
    const long long nb[3] = {ix + 1 < nx ? c + 1 : -1, iy + 1 < ny ? c + nx : -1,
                             iz + 1 < nz ? c + (long long)nx * ny : -1};
    for (int d = 0; d < 3; ++d) {
      if (nb[d] < 0) continue;
      const TS* r2 = eta + nb[d] * K;
      double g = 0.0;
      for (int i = 0; i < K; ++i) {
        const double df = (double)to_tc(r2[i]) - (double)to_tc(row[i]);
        g += df * df;
      }
      f_grad += g * ih2[d];
    }
    const int am = cell_argmax(row, K);
    atomicAdd(&hist[am], 1);
    if (argmax0 && argmax0[c] != am) ++changed;
    if (T >= a.T_l) ++liquid;
    tmax = fmax(tmax, T);
  }
  f_grad *= half_kappa_ad;
  // fixed-order block reduction: wave butterflies, then waves in index order
  f_loc = wave_sum(f_loc);
  f_grad = wave_sum(f_grad);
  tmax = wave_max(tmax);
  esum = wave_sum(esum);
  changed = wave_sum_i(changed);
  liquid = wave_sum_i(liquid);
  nonfinite = wave_sum_i(nonfinite);
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    sf[w][0] = f_loc;
    sf[w][1] = f_grad;
    sf[w][2] = tmax;
    sf[w][3] = esum;
    si[w][0] = changed;
    si[w][1] = liquid;
    si[w][2] = nonfinite;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[NF] = {0.0, 0.0, -INFINITY, 0.0};
    long long ri[NI] = {0, 0, 0};
    for (int j = 0; j < RB / 64; ++j) {
      r[0] += sf[j][0];
      r[1] += sf[j][1];
      r[2] = fmax(r[2], sf[j][2]);
      r[3] += sf[j][3];
      for (int q2 = 0; q2 < NI; ++q2) ri[q2] += si[j][q2];
    }
    for (int q2 = 0; q2 < NF; ++q2) part_f[(long long)blockIdx.x * NF + q2] = r[q2];
    for (int q2 = 0; q2 < NI; ++q2) part_i[(long long)blockIdx.x * (K + NI) + K + q2] = ri[q2];
  }
  for (int i = threadIdx.x; i < K; i += RB) part_i[(long long)blockIdx.x * (K + NI) + i] = hist[i];
}

// per-orientation column sums over fixed row chunks (sequential within a chunk)
template <class TS>
__global__ __launch_bounds__(256) void reduce_cols_kernel(const TS* __restrict__ eta, long long n, int K,
                                                         long long nchunks, double* __restrict__ part_v) {
  const long long tix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tix >= nchunks * K) return;
  const long long ch = tix / K;
  const int i = (int)(tix - ch * K);
  const long long r0 = ch * kColRows, r1 = min(n, r0 + kColRows);
  double s = 0.0;
  for (long long r = r0; r < r1; ++r) s += (double)to_tc(eta[r * K + i]);
  part_v[ch * K + i] = s;
}

__global__ __launch_bounds__(256) void reduce_final_kernel(int nb, long long nchunks, int K, double dV,
                                                          const double* __restrict__ part_f,
                                                          const long long* __restrict__ part_i,
                                                          const double* __restrict__ part_v,
                                                          double* __restrict__ out_f, long long* __restrict__ out_i) {
  // one thread per output value, sequential over partials (fixed order)
  const int tid = threadIdx.x;
  for (int q = tid; q < NF; q += blockDim.x) {
    double r = q == 2 ? -INFINITY : 0.0;
    for (int b = 0; b < nb; ++b) r = q == 2 ? fmax(r, part_f[(long long)b * NF + q]) : r + part_f[(long long)b * NF + q];
    out_f[q] = (q == 0 || q == 1) ? r * dV : r;
  }
  for (int i = tid; i < K; i += blockDim.x) {
    double s = 0.0;
    for (long long ch = 0; ch < nchunks; ++ch) s += part_v[ch * K + i];
    out_f[NF + i] = s * dV;
  }
  for (int q = tid; q < K + NI; q += blockDim.x) {
    long long s = 0;
    for (int b = 0; b < nb; ++b) s += part_i[(long long)b * (K + NI) + q];
    out_i[q] = s;
  }
}

// CFD coupling (coupled/example.py:31-42): T_pf[(z, y, x)] = r T_past[(x, y, z)] + (1 - r) T_future[(x, y, z)]
template <class TT>
__global__ __launch_bounds__(256) void cfd_temperature_kernel(const TT* __restrict__ past, const TT* __restrict__ fut,
                                                             int nz, int ny, int nx, double r, TT* __restrict__ out) {
  const long long n = (long long)nz * ny * nx;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int ix = (int)(c % nx), iy = (int)((c / nx) % ny), iz = (int)(c / ((long long)nx * ny));
  const long long src = ((long long)ix * ny + iy) * nz + iz;
  out[c] = (TT)(r * (double)past[src] + (1.0 - r) * (double)fut[src]);
}

static bool grid_ok(long long n, long long threads_per_item) {
  return (n + 255) / 256 * threads_per_item < (1LL << 31);
}

}  // namespace pf

using namespace pf;

// ====================================================================== C ABI
extern "C" {

int pf_version(void) { return PFHIP_ABI_VERSION; }

const char* pf_last_error(void) { return g_err.c_str(); }

int pf_abi_sizes(size_t* params, size_t* laser, size_t* batch_inst) {
  if (params) *params = sizeof(pf_params_t);
  if (laser) *laser = sizeof(pf_laser_t);
  if (batch_inst) *batch_inst = sizeof(pf_batch_inst_t);
  return PF_OK;
}

int pf_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int pf_tile_count(int32_t dtype, int32_t k) { return entry_count(dtype, k, false); }

int pf_tile_shape(int32_t dtype, int32_t k, int32_t tile, int32_t* tx, int32_t* ty) {
  const StepEntry* e = find_entry(dtype, k, false, tile);
  if (!e) return fail(PF_E_ARG, "no such tile shape for this (dtype, K)");
  if (tx) *tx = e->tx;
  if (ty) *ty = e->ty;
  return PF_OK;
}

int pf_tile_lanes(int32_t dtype, int32_t k, int32_t tile) {
  const StepEntry* e = find_entry(dtype, k, false, tile);
  if (!e) return fail(PF_E_ARG, "no such tile shape for this (dtype, K)");
  return e->h;
}

int pf_fast_path(int32_t dtype, int32_t k) { return entry_count(dtype, k, false) > 0 ? 1 : 0; }

int pf_step(const void* eta_in_dev, void* eta_out_dev, int32_t dtype, const void* T_dev, const pf_params_t* p,
            const void* axes_dev, double t_prev, double t_crt, int32_t tile, int32_t zchunk, void* stream) {
  if (!eta_in_dev || !eta_out_dev) return fail(PF_E_ARG, "eta buffers must be non-NULL");
  if (eta_in_dev == eta_out_dev) return fail(PF_E_ARG, "eta_in and eta_out must not alias");
  if (((uintptr_t)eta_in_dev | (uintptr_t)eta_out_dev) & 15)
    return fail(PF_E_ARG, "eta buffers must be 16-byte aligned");
  StepDev a;
  int rc = make_stepdev(p, dtype, T_dev, axes_dev, zchunk, a);
  if (rc) return rc;
  const StepDyn d = make_dyn(a, eta_in_dev, eta_out_dev, t_prev, t_crt - t_prev);  // allen_cahn.py:112
  const bool aniso = p->anisotropy > 0.0;
  if (tile < -2) return fail(PF_E_ARG, "tile must be >= -2");
  const StepEntry* e = tile == -2 ? nullptr : find_entry(dtype, p->num_oris, aniso, tile);
  hipError_t err;
  if (e) {
    if (zchunk < 0 || (zchunk & ~(PF_ZCHUNK_TAIL | 0xffff)) != 0)
      return fail(PF_E_ARG, "zchunk must be a plane count in [0, 65535], optionally | PF_ZCHUNK_TAIL");
    if (march_fits(p, dtype)) return PF_E_ARG;
    set_tiles(a, e->tx, e->ty, zchunk);
    const long long nt = (long long)a.tiles_x * a.tiles_y * a.tiles_z;
    if (nt >= (1LL << 31)) return fail(PF_E_ARG, "too many tiles");
    err = e->step(a, d, (hipStream_t)stream);
  } else {
    if (tile > 0 && entry_count(dtype, p->num_oris, aniso) > 0) return fail(PF_E_ARG, "no such tile shape");
    const long long n = (long long)p->nz * p->ny * p->nx;
    if (!grid_ok(n, 1)) return fail(PF_E_ARG, "grid too large for the generic kernel");
    err = generic_step(dtype, aniso)(a, d, (hipStream_t)stream);
  }
  if (err != hipSuccess) return hip_fail(err, "pf_step launch");
  return PF_OK;
}

int pf_rhs(const void* eta_in_dev, void* rhs_out_dev, int32_t dtype, const void* T_dev, const pf_params_t* p,
           const void* axes_dev, double t, void* stream) {
  if (!eta_in_dev || !rhs_out_dev) return fail(PF_E_ARG, "eta/rhs buffers must be non-NULL");
  if (eta_in_dev == rhs_out_dev) return fail(PF_E_ARG, "eta_in and rhs_out must not alias");
  StepDev a;
  int rc = make_stepdev(p, dtype, T_dev, axes_dev, 0, a);
  if (rc) return rc;
  const StepDyn d = make_dyn(a, eta_in_dev, rhs_out_dev, t, 1.0);  // c1, c2 carry L only: rhs = -L (grad + laplace)
  const long long n = (long long)p->nz * p->ny * p->nx;
  if (!grid_ok(n, 1)) return fail(PF_E_ARG, "grid too large");
  hipError_t err = generic_rhs(dtype, p->anisotropy > 0.0)(a, d, (hipStream_t)stream);
  return err == hipSuccess ? PF_OK : hip_fail(err, "pf_rhs launch");
}

int pf_temperature(void* T_dev, int32_t T_dtype, const pf_params_t* p, double t, void* stream) {
  if (!T_dev) return fail(PF_E_ARG, "T buffer is NULL");
  if (T_dtype != PF_F64 && T_dtype != PF_F32) return fail(PF_E_ARG, "T dtype must be PF_F64 or PF_F32");
  if (p && p->laser.model == PF_T_BUFFER) return fail(PF_E_ARG, "PF_T_BUFFER has no closed-form temperature");
  StepDev a;
  int rc = make_stepdev(p, PF_F64, nullptr, (const void*)1, 0, a);
  if (rc) return rc;
  const long long n = (long long)p->nz * p->ny * p->nx;
  if (!grid_ok(n, 1)) return fail(PF_E_ARG, "grid too large");
  const dim3 g((unsigned)((n + 255) / 256));
  if (T_dtype == PF_F64)
    temperature_kernel<double><<<g, 256, 0, (hipStream_t)stream>>>(a, (double*)T_dev, t);
  else
    temperature_kernel<float><<<g, 256, 0, (hipStream_t)stream>>>(a, (float*)T_dev, t);
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? PF_OK : hip_fail(err, "pf_temperature launch");
}

int pf_cfd_temperature(const void* T_past_dev, const void* T_future_dev, int32_t T_dtype, int32_t nz, int32_t ny,
                       int32_t nx, double t_pf, double t_cfd, double dt_cfd, void* T_out_dev, void* stream) {
  if (!T_past_dev || !T_future_dev || !T_out_dev) return fail(PF_E_ARG, "NULL buffer");
  if (T_dtype != PF_F64 && T_dtype != PF_F32) return fail(PF_E_ARG, "T dtype must be PF_F64 or PF_F32");
  if (nz < 1 || ny < 1 || nx < 1) return fail(PF_E_ARG, "grid dimensions must be >= 1");
  if (!(dt_cfd != 0.0)) return fail(PF_E_ARG, "dt_cfd must be non-zero");
  const long long n = (long long)nz * ny * nx;
  if (!grid_ok(n, 1)) return fail(PF_E_ARG, "grid too large");
  const double r = (t_cfd - t_pf) / dt_cfd;  // coupled/example.py:40
  const dim3 g((unsigned)((n + 255) / 256));
  hipStream_t s = (hipStream_t)stream;
  if (T_dtype == PF_F64)
    cfd_temperature_kernel<double><<<g, 256, 0, s>>>((const double*)T_past_dev, (const double*)T_future_dev, nz, ny,
                                                     nx, r, (double*)T_out_dev);
  else
    cfd_temperature_kernel<float><<<g, 256, 0, s>>>((const float*)T_past_dev, (const float*)T_future_dev, nz, ny, nx,
                                                    r, (float*)T_out_dev);
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? PF_OK : hip_fail(err, "pf_cfd_temperature launch");
}

int pf_init_onehot(void* eta_dev, int32_t dtype, const int32_t* ori_dev, int64_t n_cells, int32_t k, void* stream) {
  if (!eta_dev || !ori_dev) return fail(PF_E_ARG, "NULL buffer");
  if (!valid_dtype(dtype)) return fail(PF_E_ARG, "bad dtype");
  if (n_cells < 1 || k < 1) return fail(PF_E_ARG, "n_cells and k must be >= 1");
  const long long ne = (long long)n_cells * k;
  if (!grid_ok(ne, 1)) return fail(PF_E_ARG, "state too large");
  const dim3 g((unsigned)((ne + 255) / 256));
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case PF_F64: onehot_kernel<double><<<g, 256, 0, s>>>((double*)eta_dev, ori_dev, n_cells, k); break;
    case PF_F32: onehot_kernel<float><<<g, 256, 0, s>>>((float*)eta_dev, ori_dev, n_cells, k); break;
    case PF_BF16: onehot_kernel<bf16_t><<<g, 256, 0, s>>>((bf16_t*)eta_dev, ori_dev, n_cells, k); break;
  }
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? PF_OK : hip_fail(err, "pf_init_onehot launch");
}

int pf_voronoi(const double* seeds_dev, int32_t n_seeds, const int32_t* grain_ori_dev, const pf_params_t* p,
               int32_t* out_dev, void* stream) {
  if (!seeds_dev || !out_dev || !p) return fail(PF_E_ARG, "NULL argument");
  if (n_seeds < 1) return fail(PF_E_ARG, "n_seeds must be >= 1");
  if (p->nz < 1 || p->ny < 1 || p->nx < 1) return fail(PF_E_ARG, "grid dimensions must be >= 1");
  const long long n = (long long)p->nz * p->ny * p->nx;
  if (!grid_ok(n, 1)) return fail(PF_E_ARG, "grid too large");
  voronoi_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      seeds_dev, n_seeds, grain_ori_dev, p->nz, p->ny, p->nx, p->h[0], p->h[1], p->h[2], out_dev);
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? PF_OK : hip_fail(err, "pf_voronoi launch");
}

int pf_argmax(const void* eta_dev, int32_t dtype, int64_t n_cells, int32_t k, int32_t* out_dev, void* stream) {
  if (!eta_dev || !out_dev) return fail(PF_E_ARG, "NULL buffer");
  if (!valid_dtype(dtype)) return fail(PF_E_ARG, "bad dtype");
  if (n_cells < 1 || k < 1) return fail(PF_E_ARG, "n_cells and k must be >= 1");
  if (!grid_ok(n_cells, 1)) return fail(PF_E_ARG, "grid too large");
  const dim3 g((unsigned)((n_cells + 255) / 256));
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case PF_F64: argmax_kernel<double><<<g, 256, 0, s>>>((const double*)eta_dev, n_cells, k, out_dev); break;
    case PF_F32: argmax_kernel<float><<<g, 256, 0, s>>>((const float*)eta_dev, n_cells, k, out_dev); break;
    case PF_BF16: argmax_kernel<bf16_t><<<g, 256, 0, s>>>((const bf16_t*)eta_dev, n_cells, k, out_dev); break;
  }
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? PF_OK : hip_fail(err, "pf_argmax launch");
}

size_t pf_reduce_workspace_bytes(int64_t n_cells, int32_t k) {
  if (n_cells < 1 || k < 1) return 0;
  const long long nb = reduce_blocks(n_cells);
  const long long nch = (n_cells + kColRows - 1) / kColRows;
  return (size_t)(nb * NF * 8 + nb * (k + NI) * 8 + nch * k * 8);
}

}  // extern "C"

namespace {

// Argument checks shared by pf_reduce and pf_reduce_noalloc (host only, no HIP call):
// fills the instance description and the workspace size the reduction needs.
int reduce_check(const void* eta_dev, int32_t dtype, const void* T_dev, const pf_params_t* p,
                 const double* out_f64_dev, const int64_t* out_i64_dev, StepDev& a, size_t& need) {
  if (!eta_dev || !out_f64_dev || !out_i64_dev) return fail(PF_E_ARG, "NULL buffer");
  int rc = make_stepdev(p, dtype, T_dev, (const void*)1, 0, a);
  if (rc) return rc;
  if (p->num_oris > RMAXK) return fail(PF_E_UNSUPPORTED, "pf_reduce supports K <= 256");
  const long long n = (long long)p->nz * p->ny * p->nx;
  const long long nch = (n + kColRows - 1) / kColRows;
  if (!grid_ok(nch * p->num_oris, 1)) return fail(PF_E_ARG, "grid too large");
  need = pf_reduce_workspace_bytes(n, p->num_oris);
  return PF_OK;
}

// The three reduction launches over a workspace of at least pf_reduce_workspace_bytes bytes.
// Every partial result is written before it is read, so the workspace needs no clearing.
int reduce_launch(const StepDev& a, const void* eta_dev, int32_t dtype, const int32_t* argmax0_dev,
                  const pf_params_t* p, double t, double* out_f64_dev, int64_t* out_i64_dev, void* ws_dev,
                  void* stream) {
  const int K = p->num_oris;
  const long long n = (long long)p->nz * p->ny * p->nx;
  const int nb = reduce_blocks(n);
  const long long per_block = (n + nb - 1) / nb;
  const long long nch = (n + kColRows - 1) / kColRows;
  double* part_f = (double*)ws_dev;
  long long* part_i = (long long*)(part_f + (long long)nb * NF);
  double* part_v = (double*)(part_i + (long long)nb * (K + NI));
  hipStream_t s = (hipStream_t)stream;
  const double hk = 0.5 * p->kappa_g * p->ad_hoc;
  const dim3 gc((unsigned)((nch * K + 255) / 256));
  switch (dtype) {
    case PF_F64:
      reduce_cells_kernel<double><<<nb, RB, 0, s>>>(a, (const double*)eta_dev, argmax0_dev, t, p->m_g, p->gamma,
                                                     p->ad_hoc, hk, per_block, part_f, part_i);
      reduce_cols_kernel<double><<<gc, 256, 0, s>>>((const double*)eta_dev, n, K, nch, part_v);
      break;
    case PF_F32:
      reduce_cells_kernel<float><<<nb, RB, 0, s>>>(a, (const float*)eta_dev, argmax0_dev, t, p->m_g, p->gamma,
                                                    p->ad_hoc, hk, per_block, part_f, part_i);
      reduce_cols_kernel<float><<<gc, 256, 0, s>>>((const float*)eta_dev, n, K, nch, part_v);
      break;
    case PF_BF16:
      reduce_cells_kernel<bf16_t><<<nb, RB, 0, s>>>(a, (const bf16_t*)eta_dev, argmax0_dev, t, p->m_g, p->gamma,
                                                     p->ad_hoc, hk, per_block, part_f, part_i);
      reduce_cols_kernel<bf16_t><<<gc, 256, 0, s>>>((const bf16_t*)eta_dev, n, K, nch, part_v);
      break;
  }
  const double dV = p->h[0] * p->h[1] * p->h[2];
  reduce_final_kernel<<<1, 256, 0, s>>>(nb, nch, K, dV, part_f, part_i, part_v, out_f64_dev, (long long*)out_i64_dev);
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? PF_OK : hip_fail(err, "pf_reduce launch");
}

}  // namespace

// ---- allocation-free diagnostics (pf_reduce_noalloc)
// The reduction's partial results live in this fixed array instead of a caller workspace.
// Each GPU holds its own copy from code-object load until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: kernels get the device address
// that hipGetSymbolAddress returns for the current device (scratch_dev below).
static __device__ __align__(256) unsigned char scratch[PF_REDUCE_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices];  // device address of `scratch` per device, looked up once

// Device address of `scratch` on the current device (set up on the first call per device).
int scratch_dev(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(PF_E_UNSUPPORTED, "pf_reduce_noalloc supports devices 0..63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(PF_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[dev] = ptr;
  }
  *out = g_scratch_dev[dev];
  return PF_OK;
}

}  // namespace

extern "C" {

int pf_reduce(const void* eta_dev, int32_t dtype, const void* T_dev, const int32_t* argmax0_dev,
              const pf_params_t* p, double t, double* out_f64_dev, int64_t* out_i64_dev, void* ws_dev,
              size_t ws_bytes, void* stream) {
  StepDev a;
  size_t need = 0;
  int rc = reduce_check(eta_dev, dtype, T_dev, p, out_f64_dev, out_i64_dev, a, need);
  if (rc) return rc;
  if (!ws_dev) return fail(PF_E_ARG, "NULL workspace");
  if (ws_bytes < need) return fail(PF_E_ARG, "workspace too small");
  return reduce_launch(a, eta_dev, dtype, argmax0_dev, p, t, out_f64_dev, out_i64_dev, ws_dev, stream);
}

int pf_reduce_noalloc(const void* eta_dev, int32_t dtype, const void* T_dev, const int32_t* argmax0_dev,
                      const pf_params_t* p, double t, double* out_f64_dev, int64_t* out_i64_dev, void* stream) {
  StepDev a;
  size_t need = 0;
  int rc = reduce_check(eta_dev, dtype, T_dev, p, out_f64_dev, out_i64_dev, a, need);
  if (rc) return rc;
  if (need > (size_t)PF_REDUCE_SCRATCH_BYTES)
    return fail(PF_E_ARG, "pf_reduce_noalloc: the reduction needs " + std::to_string(need) +
                              " bytes of scratch, more than its fixed " + std::to_string(PF_REDUCE_SCRATCH_BYTES) +
                              " (use pf_reduce with a workspace)");
  void* ws = nullptr;
  rc = scratch_dev(&ws);
  if (rc) return rc;
  return reduce_launch(a, eta_dev, dtype, argmax0_dev, p, t, out_f64_dev, out_i64_dev, ws, stream);
}

}  // extern "C"
