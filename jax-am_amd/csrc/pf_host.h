// Host-side helpers shared by the C-ABI translation units: error reporting,
// parameter validation, StepDev construction and kernel selection.
#pragma once

#include <string>

#include "pf_kernels.h"

namespace pf {

// thread-local last error (pf_last_error)
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

size_t dtype_size(int dtype);
bool valid_dtype(int dtype);

// Validates p and fills the instance-constant part of StepDev (everything except
// in/out/t_prev/hstep).  Returns PF_OK or a PF_E* code with the error set.
int make_stepdev(const pf_params_t* p, int dtype, const void* T_dev, const void* axes_dev, int tile_zc,
                 StepDev& a);

// Kernel selection for (dtype, K, anisotropy, tile).  Returns nullptr when (dtype, K) has no
// register-tiled instantiation; tile < 0 picks the default shape.
const StepEntry* find_entry(int dtype, int k, bool aniso, int tile);
int entry_count(int dtype, int k, bool aniso);
// default z-chunk for a grid and tile shape
int default_zchunk(int nz, int ny, int nx, int tx, int ty);
// fills the tile grid (tiles_x/y/z, zc) for a tile shape
int march_fits(const pf_params_t* p, int dtype);
void set_tiles(StepDev& a, int tx, int ty, int zc);

}  // namespace pf
