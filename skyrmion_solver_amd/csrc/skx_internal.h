// Internal declarations shared by the kernel and runtime translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/skx.h"
#include "skx_physics.h"

namespace skx {

// Device-resident solver status.  Written only by the last block of a
// reducing kernel (thread 0, ordinary vector stores) or by the host between
// runs; read by the kernels of the next launch.
struct DevStatus {
  double energy;       // last energy reduced (sum of e*vol)
  double baryon;       // last baryon number reduced
  double e_prev;       // energy the next arrest decision compares against
  double err_signed;   // signed max of the last stage-4 gradient (reference "err")
  double err_abs;      // max |g| of the last stage-4 gradient
  double keep;         // 1.0: use V as stored; 0.0: V is arrested (treated as 0)
  double tol;          // convergence threshold on err_abs
  int32_t arrest_pending;  // V was produced by a step whose arrest is not decided yet
  int32_t done;            // run finished (converged, step limit, or non-finite)
  int32_t steps;           // steps completed in the current run
  int32_t max_steps;       // step limit of the current run
  int32_t nonfinite;       // a non-finite energy or gradient was seen
  int32_t killkinen;       // arrest enabled (reference p_i[8])
  uint32_t cnt_sum;        // last-block counters (reset by the last block)
  uint32_t cnt_max;
  int32_t hist_cap;        // capacity of the optional per-step history
  int32_t pad_;
};

// Per-step record kept on the device when a history buffer is attached.
struct StepRecord {
  double energy;      // energy of the field entering the step
  double err_signed;  // stage-4 signed max
  double err_abs;     // stage-4 max |g|
  double baryon;      // baryon number of the field entering the step
};

enum StageFlags : int {
  FUSE_ENERGY = 1,   // stage 1 / 0: reduce the energy of the stage field
  FUSE_BARYON = 2,   // stage 1 / 0: reduce the baryon number as well
  CHECK_DONE = 4,    // exit early when st->done is set
  DECIDE_ARREST = 8, // stage 1 / 0: make the pending arrest decision from the reduced energy
  RUN_MODE = 16,     // stage 4: count steps / test convergence in st
};

template <typename T>
struct StageArgs {
  const T* F;    // field (stage field base)
  const T* V;    // velocity
  const T* la;   // stage 3: l1, stage 4: l2
  const T* lb;   // stage 4: l1
  const T* lc;   // stage 4: l3
  T* lout;       // stages 1-3: l_s
  T* Fout;       // stage 4: new field
  T* Vout;       // stage 4: new velocity
  T* aux;        // rho: per-site D^a_j (9 components) then R^a_ij (9), written by the curl pass
  double* eg_out;   // optional: projected gradient (double) of this stage
  double* dens_e;   // optional (stage 0): energy density per site
  double* dens_b;   // optional (stage 0): baryon density per site
  DevStatus* st;
  double* partial;  // per-block reduction partials, 2 per block
  StepRecord* hist; // optional per-step history
  int X, Y, Z, halo;
  int xchunk;
  int nbz, nby, nbx;
  long long N;      // sites per component
  T ilsx, ilsy, ilsz;
  PhysParams P;
  int unit_length;
  int flags;
};

// One lattice of a batched launch.  args[s-1][parity] are the arguments of
// RK4 stage s for ping-pong parity 0 / 1 (F, V read from buffer `parity`,
// written to the other); the lattice owns blocks
// [first_block, first_block + nblocks) of every batched stage launch.
template <typename T>
// Block map 0 serves the stages on the batch's base tile, map 1 the stages on
// the 32 x 16 tile (FP32 per-stage defaults); a lattice holds a run of blocks
// in each map it uses.
struct BatchSlot {
  StageArgs<T> args[4][2];
  int first_block[2];
  int nblocks[2];
  int active;
  int pad_[3];
};

// The two-lanes-per-site stage kernels (skx_stage_pair.h) address a lane's
// second component as an offset of N elements past the first one's buffer
// descriptor; they apply where that offset, plus one plane, stays below the
// 2^31-byte range of a buffer offset.
inline bool lane_pair_fits(long long N, long long YZ, size_t elem) {
  return (unsigned long long)(N + YZ) * elem < (1ull << 31);
}

// Kernel launchers (skx_kernels.hip).  Return hipSuccess or the launch error.
struct LaunchShape {
  int tz, ty, xchunk;
  int lanes = 1;  // threads per site: 2 for the FP64 pion stage kernels of skx_stage_pair.h
};

hipError_t launch_stage_f64(int theory, int stage, const StageArgs<double>& a, const LaunchShape& s, hipStream_t st);
hipError_t launch_stage_f32(int theory, int stage, const StageArgs<float>& a, const LaunchShape& s, hipStream_t st);
int stage_blocks(int X, int Y, int Z, const LaunchShape& s);
bool shape_supported(const LaunchShape& s);
// resident blocks per CU of the stage kernel (hipOccupancyMaxActiveBlocksPerMultiprocessor)
hipError_t stage_blocks_per_cu(int prec_fp32, int theory, int stage, const LaunchShape& s, int* blocks);

hipError_t launch_stage_batch_f64(int theory, int stage, const LaunchShape& s, const BatchSlot<double>* slots,
                                  const int* blk_inst, int map, int parity, int grid, hipStream_t st);
hipError_t launch_stage_batch_f32(int theory, int stage, const LaunchShape& s, const BatchSlot<float>* slots,
                                  const int* blk_inst, int map, int parity, int grid, hipStream_t st);

// Per-(precision, theory) launchers, instantiated in skx_stage_<theory>_<precision>.hip.
template <typename T, int TH>
hipError_t stage_launch(int stage, const StageArgs<T>& a, const LaunchShape& s, hipStream_t st);
template <typename T, int TH>
hipError_t stage_launch_batch(int stage, const LaunchShape& s, const BatchSlot<T>* slots, const int* blk_inst,
                              int map, int parity, int grid, hipStream_t st);
template <typename T, int TH>
hipError_t stage_occupancy(int stage, const LaunchShape& s, int* blocks);

// Device initial data (skx_init.hip).
struct InitArgs {
  skx_init_desc d;
  int X, Y, Z, nf, halo;
  double lsx, lsy, lsz;
};
hipError_t launch_init_field_f64(double* F, double* V, const InitArgs& a, hipStream_t st);
hipError_t launch_init_field_f32(float* F, float* V, const InitArgs& a, hipStream_t st);
hipError_t launch_set_component_f64(double* Fc, const double* dens, long long N, double scale, hipStream_t st);
hipError_t launch_set_component_f32(float* Fc, const double* dens, long long N, double scale, hipStream_t st);

// Post-convergence observables (skx_observables.hip): per-block FP64 partial
// sums, `nv` values per block, blocks from observable_blocks(N).
struct ObsArgs {
  int kind, nv;
  int X, Y, Z, halo;
  double lsx, lsy, lsz;
  double com[3];
  PhysParams P;
};
constexpr int kObsMaxBlocks = 2048;
int observable_blocks(long long N);
hipError_t launch_observable_f64(const double* F, double* partial, const ObsArgs& a, int blocks, hipStream_t st);
hipError_t launch_observable_f32(const float* F, double* partial, const ObsArgs& a, int blocks, hipStream_t st);
// Per-device address and size of the observables' static partials array (the
// allocation-free entry point's scratch); `device` must be current.
hipError_t observable_scratch(int device, double** ptr, size_t* bytes);

hipError_t launch_fill(void* p, size_t bytes, hipStream_t st);  // zero fill kernel (graph safe)
hipError_t launch_apply_arrest_f64(double* V, long long n, const DevStatus* st, hipStream_t s);
hipError_t launch_apply_arrest_f32(float* V, long long n, const DevStatus* st, hipStream_t s);

}  // namespace skx
