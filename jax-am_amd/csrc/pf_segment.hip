// Grain segmentation on the GPU (include/pfhip.h pf_segment, pf_segment_moments):
// connected components of face-adjacent non-liquid cells with equal orientation index, the
// graph search of jax_am/phase_field/utils.py:217-265 done as label propagation.
//
// Every non-liquid cell starts with its own index as label.  A pass replaces each label by
// the minimum over the cell's label, the labels of its same-orientation non-liquid face
// neighbours, and the label of the cell its label names (pointer jumping).  A label is always
// the index of a cell of the same component and never increases, so passes may read labels
// being updated by other threads; the host repeats passes until one changes nothing, when every
// cell holds the smallest index of its component.  Roots (label == own index) then get compact
// grain ids, and the per-grain moments are exact int64 sums (independent of summation order).
#include "pf_host.h"

namespace pf {

__global__ __launch_bounds__(256) void seg_init_kernel(const StepDev a, double t, int32_t* __restrict__ label) {
  const long long n = (long long)a.nz * a.ny * a.nx;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int ix = (int)(c % a.nx), iy = (int)((c / a.nx) % a.ny), iz = (int)(c / ((long long)a.nx * a.ny));
  const double T = cell_temperature(a, ix, iy, iz, c, t);
  label[c] = T < a.T_l ? (int32_t)c : -1;  // utils.py:193: non-liquid = T < T_liquidus
}

__global__ __launch_bounds__(256) void seg_pass_kernel(int nz, int ny, int nx, const int32_t* __restrict__ ori,
                                                       int32_t* label, int32_t* __restrict__ changed) {
  const long long n = (long long)nz * ny * nx;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int32_t l0 = label[c];
  if (l0 < 0) return;
  const int ix = (int)(c % nx), iy = (int)((c / nx) % ny), iz = (int)(c / ((long long)nx * ny));
  const int32_t o = ori[c];
  int32_t m = l0;
  const long long sxy = (long long)nx * ny;
  const long long nb[6] = {ix > 0 ? c - 1 : -1,       ix < nx - 1 ? c + 1 : -1, iy > 0 ? c - nx : -1,
                           iy < ny - 1 ? c + nx : -1, iz > 0 ? c - sxy : -1,    iz < nz - 1 ? c + sxy : -1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (nb[k] < 0) continue;
    const int32_t ln = label[nb[k]];
    if (ln >= 0 && ori[nb[k]] == o && ln < m) m = ln;
  }
  const int32_t lm = label[m];  // pointer jump (m is a cell of the same component: lm >= 0)
  if (lm >= 0 && lm < m) m = lm;
  if (m < l0) {
    label[c] = m;
    *changed = 1;
  }
}

__global__ __launch_bounds__(256) void seg_roots_kernel(long long n, const int32_t* __restrict__ label,
                                                        int32_t* __restrict__ cid, int32_t* __restrict__ root,
                                                        int32_t* __restrict__ counter) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (label[c] == (int32_t)c) {
    const int32_t g = atomicAdd(counter, 1);
    cid[c] = g;
    root[g] = (int32_t)c;
  }
}

// One thread per run of up to RUN cells along x: sums are accumulated locally while consecutive
// cells belong to the same grain and flushed with int64 atomics when the grain changes.
constexpr int RUN = 16;
constexpr int NMOM = 10;

__device__ __forceinline__ void seg_flush(unsigned long long* __restrict__ mom, int32_t g, const long long (&s)[NMOM]) {
  unsigned long long* p = mom + (long long)g * NMOM;
#pragma unroll
  for (int k = 0; k < NMOM; ++k)
    if (s[k] != 0) atomicAdd(p + k, (unsigned long long)s[k]);
}

__global__ __launch_bounds__(256) void seg_moments_kernel(int nz, int ny, int nx, const int32_t* __restrict__ label,
                                                          const int32_t* __restrict__ cid,
                                                          unsigned long long* __restrict__ mom) {
  const int runs_x = (nx + RUN - 1) / RUN;
  const long long nrun = (long long)nz * ny * runs_x;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrun) return;
  const int rx = (int)(r % runs_x);
  const long long row = r / runs_x;  // (iz, iy)
  const int iy = (int)(row % ny), iz = (int)(row / ny);
  const int x0 = rx * RUN, x1 = min(nx, x0 + RUN);
  long long s[NMOM] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int32_t cur = -1;
  for (int ix = x0; ix < x1; ++ix) {
    const long long c = row * nx + ix;
    const int32_t l = label[c];
    if (l < 0) continue;
    const int32_t g = cid[l];
    if (g != cur) {
      if (cur >= 0) seg_flush(mom, cur, s);
#pragma unroll
      for (int k = 0; k < NMOM; ++k) s[k] = 0;
      cur = g;
    }
    s[0] += 1;
    s[1] += ix;
    s[2] += iy;
    s[3] += iz;
    s[4] += (long long)ix * ix;
    s[5] += (long long)iy * iy;
    s[6] += (long long)iz * iz;
    s[7] += (long long)ix * iy;
    s[8] += (long long)ix * iz;
    s[9] += (long long)iy * iz;
  }
  if (cur >= 0) seg_flush(mom, cur, s);
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_segment(const int32_t* ori_dev, const void* T_dev, const pf_params_t* p, double t, int32_t* labels_dev,
               int32_t* cid_dev, int32_t* root_dev, int32_t* ws_dev, int64_t max_passes, int32_t* n_grains,
               void* stream) {
  if (!ori_dev || !labels_dev || !cid_dev || !root_dev || !ws_dev || !n_grains) return fail(PF_E_ARG, "NULL argument");
  StepDev a;
  int rc = make_stepdev(p, PF_F64, T_dev, (const void*)1, 0, a);
  if (rc) return rc;
  const long long n = (long long)p->nz * p->ny * p->nx;
  if (n >= (1LL << 31) - 1) return fail(PF_E_ARG, "pf_segment supports fewer than 2^31 cells");
  if (max_passes < 1) return fail(PF_E_ARG, "max_passes must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  const dim3 g((unsigned)((n + 255) / 256));
  seg_init_kernel<<<g, 256, 0, s>>>(a, t, labels_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "pf_segment init");
  // passes in groups of four between convergence checks
  int32_t changed = 1;
  long long passes = 0;
  while (changed) {
    if (passes >= max_passes) return fail(PF_E_STATE, "pf_segment: label propagation did not converge");
    e = hipMemsetAsync(ws_dev, 0, sizeof(int32_t), s);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) {
      seg_pass_kernel<<<g, 256, 0, s>>>(p->nz, p->ny, p->nx, ori_dev, labels_dev, ws_dev);
      e = hipGetLastError();
    }
    passes += 4;
    if (e == hipSuccess) e = hipMemcpyAsync(&changed, ws_dev, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "pf_segment propagation");
  }
  e = hipMemsetAsync(ws_dev + 1, 0, sizeof(int32_t), s);
  if (e == hipSuccess) {
    seg_roots_kernel<<<g, 256, 0, s>>>(n, labels_dev, cid_dev, root_dev, ws_dev + 1);
    e = hipGetLastError();
  }
  int32_t count = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&count, ws_dev + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "pf_segment roots");
  *n_grains = count;
  return PF_OK;
}

int pf_segment_moments(const int32_t* labels_dev, const int32_t* cid_dev, const pf_params_t* p, int32_t n_grains,
                       int64_t* moments_dev, void* stream) {
  if (!labels_dev || !cid_dev || !p || (!moments_dev && n_grains > 0)) return fail(PF_E_ARG, "NULL argument");
  if (n_grains < 0) return fail(PF_E_ARG, "n_grains must be >= 0");
  if (p->nz < 1 || p->ny < 1 || p->nx < 1) return fail(PF_E_ARG, "grid dimensions must be >= 1");
  if (n_grains == 0) return PF_OK;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipMemsetAsync(moments_dev, 0, sizeof(int64_t) * NMOM * (size_t)n_grains, s);
  const long long nrun = (long long)p->nz * p->ny * ((p->nx + RUN - 1) / RUN);
  if (e == hipSuccess) {
    seg_moments_kernel<<<(unsigned)((nrun + 255) / 256), 256, 0, s>>>(p->nz, p->ny, p->nx, labels_dev, cid_dev,
                                                                      (unsigned long long*)moments_dev);
    e = hipGetLastError();
  }
  return e == hipSuccess ? PF_OK : hip_fail(e, "pf_segment_moments");
}

}  // extern "C"
