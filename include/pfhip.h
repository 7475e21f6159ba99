/*
 * pfhip.h - C ABI of libpfhip.so, the MI355X (gfx950) implementation of the
 * multi-order-parameter Allen-Cahn explicit step of JAX-AM's phase-field module.
 *
 * State layout (identical to the reference):
 *   eta[N, K] row-major, orientation channel innermost, cells flattened in
 *   C order (z, y, x) with x fastest (jax_am/phase_field/allen_cahn.py:60,150;
 *   jax_am/phase_field/utils.py:43-45).  T[N] (reference: (N, 1)).
 *
 * Conventions:
 *   - Every pointer named *_dev is a device pointer owned by the caller
 *     (PyTorch tensors in the Python host layer).  No entry point of the step
 *     path allocates device memory; pf_reduce takes a caller workspace.
 *   - Every call is stream-ordered on the hipStream_t passed as `stream`
 *     (NULL = the legacy default stream) and returns immediately.
 *   - Return value 0 = success; a negative PF_E* code otherwise, with a
 *     message available from pf_last_error() (thread-local).
 *   - No C++ exception crosses this boundary.
 */
#ifndef PFHIP_H
#define PFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFHIP_ABI_VERSION 2

/* values per row of a batch instance's source table (pf_batch_inst_t.path_dev) */
#define PF_PATH_ROW 10

/* Flag of a zchunk argument (pf_step, pf_batch_inst_t): the lowest quarter of the z-planes (rounded
 * up to whole chunks, leaving at least one full-length chunk) is marched in chunks of half the
 * length.  Marches are dispatched from the top of the domain down, so the workgroups that start last
 * are the short ones and the launch's tail is shorter.  Results do not depend on the chunking. */
#define PF_ZCHUNK_TAIL 0x10000

/* storage dtypes of eta (and of the optional temperature buffer) */
enum pf_dtype { PF_F64 = 0, PF_F32 = 1, PF_BF16 = 2 };

/* temperature models evaluated in-kernel at the cell centroids */
enum pf_tmodel {
  PF_T_BUFFER = 0,    /* T read from a caller buffer T[N] (reference-general: coupled CFD, any app) */
  PF_T_ROSENTHAL = 1, /* moving point source, applications/phase_field/simple/example.py:21-37 */
  PF_T_GAUSSIAN = 2,  /* moving Gaussian spot */
  PF_T_QUENCH = 3     /* linear quench, applications/phase_field/solidification/example.py:21-34 */
};

/* error codes */
enum pf_err {
  PF_OK = 0,
  PF_E_ARG = -1,      /* bad shape, dtype, pointer or parameter */
  PF_E_UNSUPPORTED = -2,
  PF_E_HIP = -3,      /* HIP runtime error (launch or API) */
  PF_E_STATE = -4     /* object used in a wrong state */
};

/*
 * Laser / temperature descriptor.  Source position at time t is
 * pos0 + vel * t * dir (dir is a unit vector).  X, Y, Z are centroid minus
 * source position.
 *   Rosenthal: T = T_ambient + Q/(2*pi*kappa_T) / R * exp(-vel/(2*alpha) * (R + proj)),
 *              R = |(X,Y,Z)|, proj = (X,Y,Z).dir            (simple/example.py:31-35,
 *              multi_scan/example.py:41-46); T_cap > 0 caps T (multi_scan/example.py:47).
 *   Gaussian:  T = T_ambient + dT_peak * exp(-(X^2+Y^2)/(2 sigma_xy^2) - Z^2/(2 sigma_z^2)).
 *   Quench:    T = T_ref + grad_T * z - cool_rate * t         (solidification/example.py:27-32).
 */
typedef struct pf_laser {
  int32_t model;      /* enum pf_tmodel */
  int32_t T_dtype;    /* PF_T_BUFFER only: PF_F64 or PF_F32 */
  double T_ambient;
  double pos0[3];
  double dir[3];
  double vel;
  double Q, kappa_T, alpha, T_cap;
  double dT_peak, sigma_xy, sigma_z;
  double T_ref, grad_T, cool_rate;
} pf_laser_t;

/*
 * Step parameters: the reference's yaml keys (applications/phase_field/simple/
 * pf_params.yaml) plus the constants the reference hard-codes in
 * allen_cahn.py, made explicit.
 */
typedef struct pf_params {
  int32_t nz, ny, nx, num_oris;  /* grid (Nz, Ny, Nx) and K */
  double h[3];                   /* mesh_h_xyz = (h_x, h_y, h_z)       utils.py:55-57 */
  double m_g, gamma, kappa_g, ad_hoc;     /* gamma: allen_cahn.py:43 hard-codes 1 */
  double L0, Qg, gas_const, T_liquidus;
  double T_clip;                 /* 2000, allen_cahn.py:56 */
  double zeta_sharpness;         /* 1e10, allen_cahn.py:57 */
  double anisotropy;             /* epsilon; > 0 enables the cubic term, allen_cahn.py:68-72 */
  int32_t force_liquid;          /* eta <- 0 where T > T_liquidus + liquid_shift, allen_cahn.py:117-128 */
  int32_t reserved0;
  double liquid_shift;           /* 1.0, allen_cahn.py:124 */
  pf_laser_t laser;
} pf_params_t;

/* Library version (PFHIP_ABI_VERSION) and the last error message of this thread. */
int pf_version(void);
const char* pf_last_error(void);
/* sizeof(pf_params_t), sizeof(pf_laser_t), sizeof(pf_batch_inst_t): lets bindings check their layouts. */
int pf_abi_sizes(size_t* params, size_t* laser, size_t* batch_inst);

/* Number of HIP devices visible (0 when none; never an error). */
int pf_device_count(void);

/*
 * One explicit Euler step, fused:  eta_out = force_eta(eta_in + (t_crt - t_prev) * state_rhs(eta_in, t_prev, [T])).
 * Replaces explicit_euler (allen_cahn.py:106-114) applied to state_rhs
 * (allen_cahn.py:53-85), plus the optional liquid reset get_force_eta_fn
 * (allen_cahn.py:117-128), as called by PFSolver.stepper (allen_cahn.py:139-143).
 *   eta_in_dev/eta_out_dev: [N, K] in `dtype` (must not alias).
 *   T_dev: [N] temperature in p->laser.T_dtype when p->laser.model == PF_T_BUFFER, else ignored (may be NULL).
 *   axes_dev: unique_grain_directions [3, K, 3] (utils.py:104-109), F64 when dtype == PF_F64 and F32
 *             otherwise; required when p->anisotropy > 0, else may be NULL.
 *   tile: index of the tile shape (pf_tile_count/pf_tile_shape), -1 for the default, or -2 to
 *         force the generic-K kernel (same arithmetic, one thread per cell, no LDS tiling).
 *   zchunk: z-planes per workgroup march (0 = default), optionally ORed with PF_ZCHUNK_TAIL.
 */
int pf_step(const void* eta_in_dev, void* eta_out_dev, int32_t dtype, const void* T_dev,
            const pf_params_t* p, const void* axes_dev, double t_prev, double t_crt,
            int32_t tile, int32_t zchunk, void* stream);

/*
 * Right-hand side only: rhs_out = state_rhs(eta_in, t, [T]) (allen_cahn.py:53-85), no Euler
 * update and no liquid reset.  Used by integrators other than explicit Euler (rk4,
 * allen_cahn.py:90-103).  Same arguments as pf_step; rhs_out has eta's dtype and shape.
 */
int pf_rhs(const void* eta_in_dev, void* rhs_out_dev, int32_t dtype, const void* T_dev,
           const pf_params_t* p, const void* axes_dev, double t, void* stream);

/* Tile shapes available for (dtype, K): count, (TX, TY) cells of shape `tile`, and the
 * threads per cell of shape `tile` (1, or 2 when a lane pair splits the K channels). */
int pf_tile_count(int32_t dtype, int32_t k);
int pf_tile_shape(int32_t dtype, int32_t k, int32_t tile, int32_t* tx, int32_t* ty);
int pf_tile_lanes(int32_t dtype, int32_t k, int32_t tile);
/* 1 if (dtype, K) runs on the register-tiled fused kernel, 0 if on the generic-K kernel. */
int pf_fast_path(int32_t dtype, int32_t k);

/*
 * Temperature field at time t into T_dev[N] (F64 or F32).  Replaces the app-side
 * get_T_laser / get_T_quench (simple/example.py:21-37, solidification/example.py:21-34),
 * used for inspect_sol / write_sols.  Model PF_T_BUFFER is rejected.
 */
int pf_temperature(void* T_dev, int32_t T_dtype, const pf_params_t* p, double t, void* stream);

/*
 * One-hot initial condition: eta[c, :] = 0, eta[c, cell_ori_inds[c]] = 1.
 * Replaces PFSolver.ini_cond (allen_cahn.py:145-153).  cell_ori_inds_dev: int32[N].
 */
/* Temperature for one-way CFD coupling (applications/phase_field/coupled/example.py:31-42):
 * T_out[(z, y, x)] = r * T_past[(x, y, z)] + (1 - r) * T_future[(x, y, z)], r = (t_cfd - t_pf) / dt_cfd,
 * i.e. the reference's interpolate_T followed by convert_temperature (CFD arrays are (Nx, Ny, Nz),
 * x slowest; the result is in cell order, x fastest, ready for PF_T_BUFFER). */
int pf_cfd_temperature(const void* T_past_dev, const void* T_future_dev, int32_t T_dtype, int32_t nz, int32_t ny,
                       int32_t nx, double t_pf, double t_cfd, double dt_cfd, void* T_out_dev, void* stream);
int pf_init_onehot(void* eta_dev, int32_t dtype, const int32_t* cell_ori_inds_dev,
                   int64_t n_cells, int32_t k, void* stream);

/*
 * Voronoi grain map on the grid: cell c (centroid ((i+1/2)h_x, (j+1/2)h_y, (k+1/2)h_z)) takes
 * the nearest of n_seeds seeds (seeds_dev: F64 [n_seeds][3]; ties go to the lower index) and
 * out_dev[c] = grain_ori_dev[seed] (or the seed index when grain_ori_dev is NULL).  Replaces the
 * Neper tessellation + hex meshing of jax_am/phase_field/neper.py:5-22 and the grain -> orientation
 * lookup of utils.py:39-40 for request-time initial states.  Uses p->nz/ny/nx and p->h only.
 */
int pf_voronoi(const double* seeds_dev, int32_t n_seeds, const int32_t* grain_ori_dev,
               const pf_params_t* p, int32_t* out_dev, void* stream);

/*
 * argmax over orientations (first maximum wins, NaN counts as maximum, like numpy):
 * out_dev[c] = argmax_i eta[c, i].  Replaces the argmax of inspect_sol / write_sols
 * (allen_cahn.py:171-172,189).
 */
int pf_argmax(const void* eta_dev, int32_t dtype, int64_t n_cells, int32_t k,
              int32_t* out_dev, void* stream);

/*
 * FP64 diagnostics (deterministic fixed-order reduction; result depends only on the
 * inputs and the shape).  Covers inspect_sol (allen_cahn.py:165-180) and adds the
 * free energy and per-orientation volumes:
 *   out_f64_dev[0] = F_local  = dV * sum_c f_loc(eta_c, zeta_c) / ad_hoc
 *   out_f64_dev[1] = F_grad   = dV * kappa_g*ad_hoc/2 * sum_faces sum_i (d eta_i / h)^2
 *   out_f64_dev[2] = max T
 *   out_f64_dev[3] = sum of all eta (checksum)
 *   out_f64_dev[4 + i] = dV * sum_c eta_ci                                        (i < K)
 *   out_i64_dev[i] = #cells with argmax == i                                      (i < K)
 *   out_i64_dev[K + 0] = #cells whose argmax differs from argmax0_dev (0 if argmax0_dev NULL)
 *   out_i64_dev[K + 1] = #cells with T >= T_liquidus
 *   out_i64_dev[K + 2] = #non-finite eta values
 * T is taken from p->laser at time t (or T_dev for PF_T_BUFFER).
 * ws_dev must hold pf_reduce_workspace_bytes(n_cells, k) bytes.
 */
size_t pf_reduce_workspace_bytes(int64_t n_cells, int32_t k);
int pf_reduce(const void* eta_dev, int32_t dtype, const void* T_dev, const int32_t* argmax0_dev,
              const pf_params_t* p, double t, double* out_f64_dev, int64_t* out_i64_dev,
              void* ws_dev, size_t ws_bytes, void* stream);

/*
 * pf_reduce without a workspace: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as pf_reduce minus ws_dev/ws_bytes; the
 * partial results live in a fixed device array of PF_REDUCE_SCRATCH_BYTES bytes inside the
 * library, one per GPU, held until the process ends (its device address is looked up on the
 * first call on each device).  Calls no allocator.  Returns PF_E_ARG when the reduction needs
 * more than PF_REDUCE_SCRATCH_BYTES (pf_reduce_workspace_bytes(n_cells, K); e.g. K = 20 fits
 * up to ~2e8 cells).  Every call on one GPU uses that GPU's array: calls that overlap on one
 * GPU (on different streams, or from several host threads) race; calls ordered on one stream
 * do not.  Nothing in the array needs to start at zero.
 */
#define PF_REDUCE_SCRATCH_BYTES (8u << 20)
int pf_reduce_noalloc(const void* eta_dev, int32_t dtype, const void* T_dev, const int32_t* argmax0_dev,
                      const pf_params_t* p, double t, double* out_f64_dev, int64_t* out_i64_dev, void* stream);

/*
 * Batched engine: many independent instances (uneven grids, K, dtype) advanced
 * together by persistent launches over a device-resident tile work list, captured
 * into a HIP graph of `chunk` steps.  No counterpart in the reference (which runs
 * one solve per process); each instance follows the same per-step arithmetic as
 * pf_step with the same tile shape, so its result is bitwise the same alone and in
 * the mix.  The engine allocates its tables once at creation; eta buffers are the
 * caller's.  See pf_batch_* below.
 */
typedef struct pf_batch pf_batch_t;

typedef struct pf_batch_inst {
  void* eta_dev[2];           /* ping-pong state buffers [N, K]; the state starts in eta_dev[0] */
  const void* T_dev;          /* PF_T_BUFFER only */
  const void* axes_dev;       /* [3, K, 3] like pf_step */
  int32_t dtype;
  int32_t tile;               /* tile shape index (-1 = default) */
  int32_t zchunk;             /* 0 = default; optionally | PF_ZCHUNK_TAIL */
  int32_t steps;              /* number of Euler steps to run */
  double t0, dt;              /* start time and fixed step */
  pf_params_t p;
  /* Optional per-step source table, NULL for none: device array [steps][PF_PATH_ROW] of FP64
   * rows (t_prev, t_crt, pos0 x, y, z, dir x, y, z, strength, vel), strength = Q (Rosenthal) or
   * dT_peak (Gaussian).  Step i then runs from t_prev to t_crt of row i with p.laser's pos0, dir,
   * strength and vel replaced by the row's, exactly as pf_step called with that row's times and
   * source: toolpaths (multi_scan/example.py:52-105) and sources whose power and speed change
   * with time (simple/example_grf.py:21-46).  t0/dt are then only the reported start.  Needs a
   * Rosenthal or Gaussian source.  The table is the caller's and must outlive the slot. */
  const double* path_dev;
} pf_batch_inst_t;

pf_batch_t* pf_batch_create(int32_t max_instances, int32_t max_tiles, void* stream);
void pf_batch_destroy(pf_batch_t* b);
/* Put an instance into a free slot; returns the slot index (>= 0) or a PF_E* code. */
int pf_batch_add(pf_batch_t* b, const pf_batch_inst_t* inst);
/* Free a slot (its steps may be unfinished; the caller owns the buffers). */
int pf_batch_remove(pf_batch_t* b, int32_t slot);
/* Run `steps` steps of every active slot (one graph replay per `chunk` steps;
   chunk <= 0 launches without a graph).  Stream-ordered. */
int pf_batch_run(pf_batch_t* b, int32_t steps, int32_t chunk);
/* Copy the per-slot steps-left counters and times to host arrays (synchronizes the stream). */
int pf_batch_status(pf_batch_t* b, int32_t* steps_left, double* t, int32_t* parity, int32_t n);
int pf_batch_active(const pf_batch_t* b);

/*
 * Grain segmentation (jax_am/phase_field/utils.py:181-265, process_eta / BFS): connected
 * components of face-adjacent non-liquid cells (T < T_liquidus at time t, temperature from
 * p->laser or T_dev as in pf_step) with equal orientation index ori_dev[N] (argmax eta).
 * On return labels_dev[N] holds, per cell, the smallest cell index of its grain (-1 for liquid
 * cells); *n_grains grains exist; grain g (0 <= g < *n_grains, in no particular order) has root
 * cell root_dev[g] (capacity N) and cid_dev[root_dev[g]] == g (cid_dev: N entries).
 * ws_dev: 2 int32 of device scratch.  Label propagation passes run until one changes nothing
 * (at most max_passes, else PF_E_STATE).  Synchronizes the stream.
 */
int pf_segment(const int32_t* ori_dev, const void* T_dev, const pf_params_t* p, double t, int32_t* labels_dev,
               int32_t* cid_dev, int32_t* root_dev, int32_t* ws_dev, int64_t max_passes, int32_t* n_grains,
               void* stream);
/* Exact integer moments of each grain of pf_segment into moments_dev[n_grains][10]: number of
 * cells and the sums of ix, iy, iz, ix^2, iy^2, iz^2, ix*iy, ix*iz, iy*iz over its cells (cell
 * indices; centroid = (index + 1/2) * h).  utils.py:268-310 derives volumes and aspect ratios
 * from these. */
int pf_segment_moments(const int32_t* labels_dev, const int32_t* cid_dev, const pf_params_t* p, int32_t n_grains,
                       int64_t* moments_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PFHIP_H */
