// Launch interface between the C-ABI layer (chem_capi.cpp) and the HIP
// kernels (chem_bdf.hip).  Internal to the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "chem_params.h"

extern "C" {
int tgr_step_occupancy(int neq);
int tgr_step_occupancy_w(int neq, int waves);
int tgr_step_waves_supported(int neq, int waves);
int tgr_step_block_waves(void);
int tgr_launch_step(const tgr::StepParams* p, int neq, int grid, hipStream_t s);
int tgr_launch_prodrates(const tgr::DevMech* m, int64_t n, const double* C, const double* T, const double* P,
                         double* cdot, hipStream_t s);
int tgr_launch_rhs(const tgr::DevMech* m, int64_t n, double p_gp, const double* y, double* f, hipStream_t s);
int tgr_launch_temperature(const tgr::DevMech* m, int64_t n, const double* Y, const double* hmix, const double* T0,
                           double* T, int* iters, hipStream_t s);
int tgr_launch_props(const tgr::PropsParams* p, int neq, int lane_map, hipStream_t s);
int tgr_launch_rhs_lanes(int neq, int64_t n, double p_gp, const double* y, double* f, hipStream_t s);
int tgr_surr_occupancy(int neq, int adaptive);
// waves per workgroup of the guided-solve kernel (its workspace share per block)
int tgr_surr_block_waves(void);
int tgr_launch_surr_features(const tgr::StepParams* p, int neq, int64_t rows, int in_pad, const float* shift,
                             const float* scale, float log_dt, uint16_t* X, hipStream_t s);
int tgr_launch_surr_set_column(uint16_t* X, int64_t rows, int in_pad, int col, float value, hipStream_t s);
int tgr_launch_surr(const tgr::SurrParams* P, int neq, int grid, hipStream_t s);
int tgr_launch_diffusion(const tgr::DiffParams* p, int neq, hipStream_t s);
// the thread-per-cell step kernel (chem_lane.hip): supported sizes, workspace
// doubles per workgroup, workgroups per CU, launch
int tgr_step_lane_supported(int neq);
int64_t tgr_step_lane_ws(int neq);
int tgr_step_lane_occupancy(int neq);
int tgr_launch_step_lane(const tgr::StepParams* p, int neq, int grid, hipStream_t s);
// device address and size of the non-allocating entry point's scratch (noalloc.hip)
int tgr_scratch_resolve(int device, void** ptr, size_t* bytes);
}
