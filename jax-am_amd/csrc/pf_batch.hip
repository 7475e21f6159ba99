// Batched engine of libpfhip.so: many independent phase-field instances advanced
// together.  Design:
//   - Slot table (StepDev descriptors, BatchSlot step counters and each slot's StepDyn
//     step arguments) lives in device memory; the device advances each slot's step
//     index and ping-pong parity and rebuilds its StepDyn, so a captured graph stays
//     valid while instances come and go.
//   - Instances are grouped into classes (dtype, K, anisotropy, tile shape).  Each
//     class has a device-resident work list of (slot, tile) items walked by one
//     persistent launch per step with a fixed grid.
//   - `chunk` steps (launches of the classes that have work + advance) are captured
//     into one HIP graph and replayed.  Graphs are cached by (chunk, set of classes with
//     work): admitting or removing an instance only rewrites table entries and work
//     lists, and a new capture happens only for a class set not seen before.
// Every tile of an instance runs exactly the per-cell arithmetic of pf_step with the
// same tile shape, so an instance's trajectory does not depend on what else is in
// the batch.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "pf_host.h"

namespace pf {

// After each step: advance every active slot's counters and build its next step's arguments;
// a source-table slot also takes its next row's times and source into its description.
__global__ __launch_bounds__(64) void advance_kernel(StepDev* __restrict__ descs, BatchSlot* __restrict__ slots,
                                                    StepDyn* __restrict__ dyns, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  BatchSlot b = slots[s];
  if (b.active && b.steps_left > 0) {
    const int done = b.n;
    b.steps_left -= 1;
    b.parity ^= 1;
    b.n += 1;
    if (b.path) {
      b.tnow = b.path[PF_PATH_ROW * done + 1];
      slots[s] = b;
      if (b.steps_left > 0) {
        StepDev a = descs[s];
        const double* r = b.path + PF_PATH_ROW * b.n;
        apply_path_row(a, r);
        descs[s] = a;
        dyns[s] = make_dyn(a, b.eta[b.parity], b.eta[b.parity ^ 1], r[0], r[1] - r[0]);
      }
    } else {
      b.tnow = b.t0 + (double)b.n * b.dt;
      slots[s] = b;
      dyns[s] = slot_dyn(descs[s], b);
    }
  }
}

struct BatchClass {
  int dtype, k, aniso, tile;
  BatchLaunch fn;
  int grid;
  int off;    // offset of this class's work list in d_work
  int count;  // host copy of the work-item count
  int users;  // slots in this class; a class without users is retired and its index reused
};

struct SlotHost {
  bool used = false;
  int cls = -1;
  int items = 0;  // work items of this slot
};

}  // namespace pf

using namespace pf;

struct pf_batch {
  int max_inst = 0, max_tiles = 0;
  static constexpr int kMaxClasses = 64;
  hipStream_t stream = nullptr;
  StepDev* d_desc = nullptr;
  StepDyn* d_dyn = nullptr;
  BatchSlot* d_slots = nullptr;
  WorkItem* d_work = nullptr;
  int2* d_info = nullptr;  // per class (work-list offset, count)
  std::vector<SlotHost> slots;
  std::vector<BatchClass> classes;
  bool work_dirty = false;
  struct Graph {
    int chunk;
    unsigned long long mask;
    hipGraph_t def;
    hipGraphExec_t exec;
  };
  static constexpr int kMaxGraphs = 16;
  std::vector<Graph> graphs;  // most recently used last
  std::mutex mu;
};

namespace {

int rebuild_work(pf_batch* b) {
  // pack each class's work items contiguously, in slot order
  std::vector<std::vector<WorkItem>> per(b->classes.size());
  for (int s = 0; s < b->max_inst; ++s) {
    const SlotHost& sh = b->slots[s];
    if (!sh.used) continue;
    for (int t = 0; t < sh.items; ++t) per[sh.cls].push_back(WorkItem{s, t});
  }
  std::vector<WorkItem> all;
  std::vector<int2> info(pf_batch::kMaxClasses, make_int2(0, 0));
  for (size_t c = 0; c < b->classes.size(); ++c) {
    b->classes[c].off = (int)all.size();
    b->classes[c].count = (int)per[c].size();
    info[c] = make_int2((int)all.size(), (int)per[c].size());
    all.insert(all.end(), per[c].begin(), per[c].end());
  }
  if ((int)all.size() > b->max_tiles) return fail(PF_E_ARG, "batch work list exceeds max_tiles");
  hipError_t e = hipSuccess;
  if (!all.empty())
    e = hipMemcpyAsync(b->d_work, all.data(), all.size() * sizeof(WorkItem), hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_info, info.data(), info.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream);
  // the host vectors above are pageable and local: finish the copies before they go away
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "pf_batch work-list upload");
  b->work_dirty = false;
  return PF_OK;
}

unsigned long long active_mask(const pf_batch* b) {
  unsigned long long m = 0;
  for (size_t c = 0; c < b->classes.size(); ++c)
    if (b->classes[c].count > 0) m |= 1ULL << c;
  return m;
}

int launch_step_all(pf_batch* b) {
  for (size_t c = 0; c < b->classes.size(); ++c) {
    const BatchClass& k = b->classes[c];
    if (k.count == 0) continue;
    hipError_t e = k.fn(b->d_desc, b->d_dyn, b->d_slots, b->d_work, b->d_info, (int)c, k.grid, b->stream);
    if (e != hipSuccess) return hip_fail(e, "pf_batch step launch");
  }
  advance_kernel<<<(b->max_inst + 63) / 64, 64, 0, b->stream>>>(b->d_desc, b->d_slots, b->d_dyn, b->max_inst);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? PF_OK : hip_fail(e, "pf_batch advance launch");
}

void drop_graphs(pf_batch* b) {
  for (pf_batch::Graph& g : b->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.def);
  }
  b->graphs.clear();
}

// Destroy the cached graphs that launch class c (its index is about to be retired).  The
// caller has synchronized the stream, so none of them is in flight.
void drop_graphs_of_class(pf_batch* b, int c) {
  const unsigned long long bit = 1ULL << c;
  for (size_t i = 0; i < b->graphs.size();) {
    if (b->graphs[i].mask & bit) {
      (void)hipGraphExecDestroy(b->graphs[i].exec);
      (void)hipGraphDestroy(b->graphs[i].def);
      b->graphs.erase(b->graphs.begin() + i);
    } else {
      ++i;
    }
  }
}

// The graph of `chunk` steps for the current class set (captured on first use).
int get_graph(pf_batch* b, int chunk, hipGraphExec_t* out) {
  const unsigned long long mask = active_mask(b);
  for (size_t i = 0; i < b->graphs.size(); ++i) {
    if (b->graphs[i].chunk == chunk && b->graphs[i].mask == mask) {
      pf_batch::Graph g = b->graphs[i];
      b->graphs.erase(b->graphs.begin() + i);
      b->graphs.push_back(g);
      *out = g.exec;
      return PF_OK;
    }
  }
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
  int rc = PF_OK;
  for (int i = 0; i < chunk && rc == PF_OK; ++i) rc = launch_step_all(b);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(e, "hipGraphInstantiate");
  }
  if ((int)b->graphs.size() >= pf_batch::kMaxGraphs) {
    // the oldest graph may still be in flight: finish the stream before destroying it
    (void)hipStreamSynchronize(b->stream);
    (void)hipGraphExecDestroy(b->graphs.front().exec);
    (void)hipGraphDestroy(b->graphs.front().def);
    b->graphs.erase(b->graphs.begin());
  }
  b->graphs.push_back(pf_batch::Graph{chunk, mask, g, ex});
  *out = ex;
  return PF_OK;
}

}  // namespace

extern "C" {

pf_batch_t* pf_batch_create(int32_t max_instances, int32_t max_tiles, void* stream) {
  if (max_instances < 1 || max_tiles < 1) {
    set_error("max_instances and max_tiles must be >= 1");
    return nullptr;
  }
  pf_batch* b = new pf_batch();
  b->max_inst = max_instances;
  b->max_tiles = max_tiles;
  b->stream = (hipStream_t)stream;
  b->slots.resize(max_instances);
  hipError_t e = hipMalloc((void**)&b->d_desc, sizeof(StepDev) * max_instances);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_dyn, sizeof(StepDyn) * max_instances);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_slots, sizeof(BatchSlot) * max_instances);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_work, sizeof(WorkItem) * max_tiles);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_info, sizeof(int2) * pf_batch::kMaxClasses);
  if (e == hipSuccess) {
    std::vector<BatchSlot> z(max_instances);
    memset(z.data(), 0, z.size() * sizeof(BatchSlot));
    std::vector<int2> zc(pf_batch::kMaxClasses, make_int2(0, 0));
    e = hipMemcpyAsync(b->d_slots, z.data(), z.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->d_info, zc.data(), zc.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  }
  if (e != hipSuccess) {
    hip_fail(e, "pf_batch_create");
    pf_batch_destroy(b);
    return nullptr;
  }
  return b;
}

void pf_batch_destroy(pf_batch_t* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_dyn) (void)hipFree(b->d_dyn);
  if (b->d_slots) (void)hipFree(b->d_slots);
  if (b->d_work) (void)hipFree(b->d_work);
  if (b->d_info) (void)hipFree(b->d_info);
  delete b;
}

int pf_batch_add(pf_batch_t* b, const pf_batch_inst_t* inst) {
  if (!b || !inst) return fail(PF_E_ARG, "NULL batch or instance");
  std::lock_guard<std::mutex> lk(b->mu);
  if (!inst->eta_dev[0] || !inst->eta_dev[1] || inst->eta_dev[0] == inst->eta_dev[1])
    return fail(PF_E_ARG, "an instance needs two distinct state buffers");
  if (((uintptr_t)inst->eta_dev[0] | (uintptr_t)inst->eta_dev[1]) & 15)
    return fail(PF_E_ARG, "eta buffers must be 16-byte aligned");
  if (inst->steps < 0) return fail(PF_E_ARG, "steps must be >= 0");
  if (inst->zchunk < 0 || (inst->zchunk & ~(PF_ZCHUNK_TAIL | 0xffff)) != 0)
    return fail(PF_E_ARG, "zchunk must be a plane count in [0, 65535], optionally | PF_ZCHUNK_TAIL");
  if (inst->path_dev && inst->p.laser.model != PF_T_ROSENTHAL && inst->p.laser.model != PF_T_GAUSSIAN)
    return fail(PF_E_ARG, "a source table needs a Rosenthal or Gaussian source");
  int slot = -1;
  for (int s = 0; s < b->max_inst; ++s)
    if (!b->slots[s].used) {
      slot = s;
      break;
    }
  if (slot < 0) return fail(PF_E_STATE, "no free batch slot");
  StepDev a;
  int rc = make_stepdev(&inst->p, inst->dtype, inst->T_dev, inst->axes_dev, inst->zchunk, a);
  if (rc) return rc;
  const bool aniso = inst->p.anisotropy > 0.0;
  const StepEntry* e = find_entry(inst->dtype, inst->p.num_oris, aniso, inst->tile);
  int items, tile = -1, grid;
  BatchLaunch fn;
  if (e) {
    if (march_fits(&inst->p, inst->dtype)) return PF_E_ARG;
    set_tiles(a, e->tx, e->ty, inst->zchunk);
    items = a.tiles_x * a.tiles_y * a.tiles_z;
    tile = e->tile;
    fn = e->batch;
    grid = 1024;
  } else {
    if (inst->tile > 0 && entry_count(inst->dtype, inst->p.num_oris, aniso) > 0)
      return fail(PF_E_ARG, "no such tile shape");
    const long long n = (long long)inst->p.nz * inst->p.ny * inst->p.nx;
    if ((n + 255) / 256 >= (1LL << 31)) return fail(PF_E_ARG, "instance too large");
    items = (int)((n + 255) / 256);
    fn = generic_batch(inst->dtype, aniso);
    grid = 2048;
  }
  int total = 0;
  for (const SlotHost& sh : b->slots)
    if (sh.used) total += sh.items;
  if (total + items > b->max_tiles) return fail(PF_E_STATE, "batch work list full (max_tiles)");
  int cls = -1, free_cls = -1;
  for (size_t c = 0; c < b->classes.size(); ++c) {
    const BatchClass& k = b->classes[c];
    if (k.users == 0) {
      if (free_cls < 0) free_cls = (int)c;
    } else if (k.dtype == inst->dtype && k.k == inst->p.num_oris && k.aniso == (int)aniso && k.tile == tile) {
      cls = (int)c;
    }
  }
  if (cls < 0) {
    // a new class takes a retired index (its graphs were dropped when it retired) or a new one
    const BatchClass nc{inst->dtype, inst->p.num_oris, (int)aniso, tile, fn, grid, 0, 0, 0};
    if (free_cls >= 0) {
      b->classes[free_cls] = nc;
      cls = free_cls;
    } else {
      if ((int)b->classes.size() >= pf_batch::kMaxClasses)
        return fail(PF_E_STATE, "too many instance classes at once (64)");
      b->classes.push_back(nc);
      cls = (int)b->classes.size() - 1;
    }
  }
  BatchSlot st;
  memset(&st, 0, sizeof(st));
  st.steps_left = inst->steps;
  st.parity = 0;
  st.n = 0;
  st.active = 1;
  st.t0 = inst->t0;
  st.dt = inst->dt;
  st.eta[0] = inst->eta_dev[0];
  st.eta[1] = inst->eta_dev[1];
  st.tnow = inst->t0;
  st.path = inst->path_dev;
  StepDyn dy = slot_dyn(a, st);
  hipError_t he = hipSuccess;
  if (inst->path_dev && inst->steps > 0) {  // the first row's times and source
    double r0[PF_PATH_ROW];
    he = hipMemcpyAsync(r0, inst->path_dev, sizeof(r0), hipMemcpyDeviceToHost, b->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(b->stream);
    if (he != hipSuccess) return hip_fail(he, "pf_batch_add path");
    apply_path_row(a, r0);
    dy = make_dyn(a, st.eta[0], st.eta[1], r0[0], r0[1] - r0[0]);
  }
  he = hipMemcpyAsync(b->d_desc + slot, &a, sizeof(StepDev), hipMemcpyHostToDevice, b->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(b->d_dyn + slot, &dy, sizeof(StepDyn), hipMemcpyHostToDevice, b->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(b->d_slots + slot, &st, sizeof(BatchSlot), hipMemcpyHostToDevice, b->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(b->stream);  // a, dy and st are locals
  if (he != hipSuccess) return hip_fail(he, "pf_batch_add upload");
  b->slots[slot].used = true;
  b->slots[slot].cls = cls;
  b->slots[slot].items = items;
  b->classes[cls].users += 1;
  b->work_dirty = true;
  return slot;
}

int pf_batch_remove(pf_batch_t* b, int32_t slot) {
  if (!b) return fail(PF_E_ARG, "NULL batch");
  std::lock_guard<std::mutex> lk(b->mu);
  if (slot < 0 || slot >= b->max_inst || !b->slots[slot].used) return fail(PF_E_ARG, "slot not in use");
  BatchSlot st;
  memset(&st, 0, sizeof(st));
  hipError_t he = hipMemcpyAsync(b->d_slots + slot, &st, sizeof(BatchSlot), hipMemcpyHostToDevice, b->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(b->stream);  // st is a local
  if (he != hipSuccess) return hip_fail(he, "pf_batch_remove");
  const int cls = b->slots[slot].cls;
  b->slots[slot] = SlotHost();
  if (--b->classes[cls].users == 0) drop_graphs_of_class(b, cls);  // retired: its index is free again
  b->work_dirty = true;
  return PF_OK;
}

int pf_batch_run(pf_batch_t* b, int32_t steps, int32_t chunk) {
  if (!b) return fail(PF_E_ARG, "NULL batch");
  if (steps < 0) return fail(PF_E_ARG, "steps must be >= 0");
  std::lock_guard<std::mutex> lk(b->mu);
  if (b->work_dirty) {
    int rc = rebuild_work(b);
    if (rc) return rc;
  }
  int done = 0;
  if (chunk > 0 && steps >= chunk) {
    hipGraphExec_t g = nullptr;
    int rc = get_graph(b, chunk, &g);
    if (rc) return rc;
    for (; done + chunk <= steps; done += chunk) {
      hipError_t e = hipGraphLaunch(g, b->stream);
      if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    }
  }
  for (; done < steps; ++done) {
    int rc = launch_step_all(b);
    if (rc) return rc;
  }
  return PF_OK;
}

int pf_batch_status(pf_batch_t* b, int32_t* steps_left, double* t, int32_t* parity, int32_t n) {
  if (!b) return fail(PF_E_ARG, "NULL batch");
  std::lock_guard<std::mutex> lk(b->mu);
  if (n > b->max_inst) n = b->max_inst;
  std::vector<BatchSlot> h(b->max_inst);
  hipError_t e = hipMemcpyAsync(h.data(), b->d_slots, sizeof(BatchSlot) * b->max_inst, hipMemcpyDeviceToHost,
                                b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "pf_batch_status");
  for (int s = 0; s < n; ++s) {
    if (steps_left) steps_left[s] = b->slots[s].used ? h[s].steps_left : -1;
    if (t) t[s] = h[s].tnow;
    if (parity) parity[s] = h[s].parity;
  }
  return PF_OK;
}

int pf_batch_active(const pf_batch_t* b) {
  if (!b) return 0;
  int c = 0;
  for (const SlotHost& s : b->slots) c += s.used ? 1 : 0;
  return c;
}

}  // extern "C"
