// Kernel entry points of the fused step (single instance and batched), their
// launchers, and the dispatch-table entry type.  Instantiated per storage dtype in
// pf_inst_*.hip.
#pragma once

#include "pf_device.h"

namespace pf {

// Per-slot device state of the batched engine.  The step index n gives
// t_prev = t0 + n*dt and t_crt = t0 + (n+1)*dt (the reference's np.arange time grid,
// applications/phase_field/simple/example.py:52), so h = t_crt - t_prev per step.
struct BatchSlot {
  int steps_left;
  int parity;  // state lives in eta[parity]
  int n;
  int active;
  double t0, dt;
  void* eta[2];
  double tnow;         // time of the current state
  const double* path;  // per-step rows (PF_PATH_ROW values, include/pfhip.h) of a source table, or null
};

// A source-table row's position, direction, strength and speed into an instance description:
// the fields make_stepdev derives from pf_laser_t pos0, dir, Q (Rosenthal) or dT_peak
// (Gaussian) and vel, with the same expressions, so a row gives bitwise the description that
// pf_step builds from a descriptor holding those values.
__host__ __device__ inline void apply_path_row(StepDev& a, const double* r) {
  a.p0x = r[2];
  a.p0y = r[3];
  a.p0z = r[4];
  a.dirx = r[5];
  a.diry = r[6];
  a.dirz = r[7];
  a.vel = r[9];
  if (a.tmodel == PF_T_ROSENTHAL) {
    a.q_coef = rosenthal_q(r[8], a.kappa_T);
    a.pe = rosenthal_pe(r[9], a.alpha);
  } else {
    a.dT_peak = r[8];
  }
}

struct WorkItem {
  int slot;
  int tile;
};

// Step arguments of slot s at its current step (allen_cahn.py:112  h = t_crt - t_prev).
__host__ __device__ inline StepDyn slot_dyn(const StepDev& a, const BatchSlot& s) {
  const double tp = s.t0 + (double)s.n * s.dt;
  const double tc = s.t0 + (double)(s.n + 1) * s.dt;
  return make_dyn(a, s.eta[s.parity], s.eta[s.parity ^ 1], tp, tc - tp);
}

template <int K, class TS, bool ANISO, int TX, int TY, int H>
__global__ __launch_bounds__(TX* TY* H) __attribute__((amdgpu_waves_per_eu(H + 1))) void step_kernel(const StepDev a,
                                                                                               const StepDyn d) {
  using TC = typename Traits<TS>::TC;
  using G = Tile<TS, K, TX, TY, H>;
  __shared__ __attribute__((aligned(16))) TC smem[G::NBUF * G::PSTR];
  __shared__ __attribute__((aligned(16))) TS stage[G::STAGE];
  __shared__ __attribute__((aligned(16))) TC ax[ANISO ? G::AXN : 4];
  __shared__ ZTab zt;
  int tx, ty, tz;
  tile_coords(blockIdx.x, a.tiles_x, a.tiles_y, a.tiles_z, tx, ty, tz);
  march_tile<K, TS, ANISO, TX, TY, H>(a, d, tx, ty, tz, smem, stage, ax, &zt);
}

template <class TS, bool ANISO, bool RHS = false>
__global__ __launch_bounds__(256) void generic_kernel(const StepDev a, const StepDyn d) {
  const long long n = (long long)a.nz * a.ny * a.nx;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell < n) generic_cell<TS, ANISO, RHS>(a, d, cell);
}

// Persistent batched step: a fixed grid walks the device-resident work list (slot,
// tile) of one (dtype, K, anisotropy, tile shape) class.  Every workgroup exits after
// the list is exhausted; slots with no steps left are skipped.  descs[slot] and
// dyns[slot] are read in place (scalar loads), never copied into registers wholesale.
template <int K, class TS, bool ANISO, int TX, int TY, int H>
__global__ __launch_bounds__(TX* TY* H) __attribute__((amdgpu_waves_per_eu(2))) void batch_kernel(
    const StepDev* __restrict__ descs, const StepDyn* __restrict__ dyns, const BatchSlot* __restrict__ slots,
    const WorkItem* __restrict__ work_all, const int2* __restrict__ cinfo, int cls) {
  using TC = typename Traits<TS>::TC;
  using G = Tile<TS, K, TX, TY, H>;
  __shared__ __attribute__((aligned(16))) TC smem[G::NBUF * G::PSTR];
  __shared__ __attribute__((aligned(16))) TS stage[G::STAGE];
  __shared__ __attribute__((aligned(16))) TC ax[ANISO ? G::AXN : 4];
  __shared__ ZTab zt;
  const int2 ci = cinfo[cls];  // (offset, count) of this class's work list
  const WorkItem* __restrict__ work = work_all + ci.x;
  for (int w = blockIdx.x; w < ci.y; w += gridDim.x) {
    const WorkItem it = work[w];
    const BatchSlot& s = slots[it.slot];
    if (s.steps_left <= 0 || !s.active) continue;
    const StepDev& a = descs[it.slot];
    int tx, ty, tz;
    tile_coords(it.tile, a.tiles_x, a.tiles_y, a.tiles_z, tx, ty, tz);
    march_tile<K, TS, ANISO, TX, TY, H>(a, dyns[it.slot], tx, ty, tz, smem, stage, ax, &zt);
    lds_barrier();  // LDS is reused by the next tile
  }
}

// Batched generic-K step: work items are (slot, block of 256 cells).
template <class TS, bool ANISO>
__global__ __launch_bounds__(256) void batch_generic_kernel(const StepDev* __restrict__ descs,
                                                           const StepDyn* __restrict__ dyns,
                                                           const BatchSlot* __restrict__ slots,
                                                           const WorkItem* __restrict__ work_all,
                                                           const int2* __restrict__ cinfo, int cls) {
  const int2 ci = cinfo[cls];
  const WorkItem* __restrict__ work = work_all + ci.x;
  for (int w = blockIdx.x; w < ci.y; w += gridDim.x) {
    const WorkItem it = work[w];
    const BatchSlot& s = slots[it.slot];
    if (s.steps_left <= 0 || !s.active) continue;
    const StepDev& a = descs[it.slot];
    const long long n = (long long)a.nz * a.ny * a.nx;
    const long long cell = (long long)it.tile * 256 + threadIdx.x;
    if (cell < n) generic_cell<TS, ANISO>(a, dyns[it.slot], cell);
  }
}

using StepLaunch = hipError_t (*)(const StepDev&, const StepDyn&, hipStream_t);
using BatchLaunch = hipError_t (*)(const StepDev*, const StepDyn*, const BatchSlot*, const WorkItem*, const int2*,
                                   int cls, int grid, hipStream_t);

struct StepEntry {
  int dtype, k, aniso, tile, tx, ty, h;
  StepLaunch step;
  BatchLaunch batch;
  size_t lds;
};

template <int K, class TS, bool ANISO, int TX, int TY, int H>
hipError_t launch_step(const StepDev& a, const StepDyn& d, hipStream_t s) {
  const int n = a.tiles_x * a.tiles_y * a.tiles_z;
  step_kernel<K, TS, ANISO, TX, TY, H><<<dim3(n), dim3(TX * TY * H), 0, s>>>(a, d);
  return hipGetLastError();
}

template <int K, class TS, bool ANISO, int TX, int TY, int H>
hipError_t launch_batch(const StepDev* a, const StepDyn* d, const BatchSlot* sl, const WorkItem* w, const int2* ci,
                        int cls, int grid, hipStream_t s) {
  batch_kernel<K, TS, ANISO, TX, TY, H><<<dim3(grid), dim3(TX * TY * H), 0, s>>>(a, d, sl, w, ci, cls);
  return hipGetLastError();
}

template <class TS, bool ANISO, bool RHS = false>
hipError_t launch_generic(const StepDev& a, const StepDyn& d, hipStream_t s) {
  const long long n = (long long)a.nz * a.ny * a.nx;
  generic_kernel<TS, ANISO, RHS><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(a, d);
  return hipGetLastError();
}

template <class TS, bool ANISO>
hipError_t launch_batch_generic(const StepDev* a, const StepDyn* d, const BatchSlot* sl, const WorkItem* w,
                                const int2* ci, int cls, int grid, hipStream_t s) {
  batch_generic_kernel<TS, ANISO><<<dim3(grid), dim3(256), 0, s>>>(a, d, sl, w, ci, cls);
  return hipGetLastError();
}

#define PF_ENTRY(DT, TS, K, A, T, TX, TY, H)                                      \
  {DT,                                                                           \
   K,                                                                            \
   A,                                                                            \
   T,                                                                            \
   TX,                                                                           \
   TY,                                                                           \
   H,                                                                            \
   &launch_step<K, TS, (A != 0), TX, TY, H>,                                     \
   &launch_batch<K, TS, (A != 0), TX, TY, H>,                                    \
   Tile<TS, K, TX, TY, H>::LDS_BYTES}

// instantiation tables, one per compiled part of pf_inst.hip (see Makefile):
// parts 0-3 F32, 4-7 BF16, 8-11 F64, each over a K range
constexpr int kInstParts = 12;
const StepEntry* step_table_part_0(int* n);
const StepEntry* step_table_part_1(int* n);
const StepEntry* step_table_part_2(int* n);
const StepEntry* step_table_part_3(int* n);
const StepEntry* step_table_part_4(int* n);
const StepEntry* step_table_part_5(int* n);
const StepEntry* step_table_part_6(int* n);
const StepEntry* step_table_part_7(int* n);
const StepEntry* step_table_part_8(int* n);
const StepEntry* step_table_part_9(int* n);
const StepEntry* step_table_part_10(int* n);
const StepEntry* step_table_part_11(int* n);
// generic-K launchers by dtype
StepLaunch generic_step(int dtype, bool aniso);
BatchLaunch generic_batch(int dtype, bool aniso);

}  // namespace pf
