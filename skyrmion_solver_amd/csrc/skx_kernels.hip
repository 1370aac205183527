// Small kernels of the runtime (zero fill, arrest) and the dispatch from
// runtime theory / precision / stage / tile shape to the stage kernels, which
// are instantiated one (precision, theory) per translation unit
// (skx_stage_<theory>_<precision>.hip).
#include "skx_device.h"

namespace skx {

// ------------------------------------------------------------------ small kernels
__global__ void k_fill_zero(uint32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}

template <typename T>
__global__ void k_apply_arrest(T* V, long long n, const DevStatus* st) {
  if (st->keep != 0.0) return;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) V[i] = T(0);
}

hipError_t launch_fill(void* p, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 4;
  if (n == 0) return hipSuccess;
  const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, st, static_cast<uint32_t*>(p), n);
  return hipGetLastError();
}

hipError_t launch_apply_arrest_f64(double* V, long long n, const DevStatus* s, hipStream_t st) {
  const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_apply_arrest<double>, dim3(blocks), dim3(256), 0, st, V, n, s);
  return hipGetLastError();
}
hipError_t launch_apply_arrest_f32(float* V, long long n, const DevStatus* s, hipStream_t st) {
  const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_apply_arrest<float>, dim3(blocks), dim3(256), 0, st, V, n, s);
  return hipGetLastError();
}

// ------------------------------------------------------------------ dispatch
int stage_blocks(int X, int Y, int Z, const LaunchShape& s) {
  const int nbz = (Z + s.tz - 1) / s.tz, nby = (Y + s.ty - 1) / s.ty, nbx = (X + s.xchunk - 1) / s.xchunk;
  return nbz * nby * nbx;
}

bool shape_supported(const LaunchShape& s) {
  return (s.tz == 64 && s.ty == 4) || (s.tz == 32 && s.ty == 8) || (s.tz == 16 && s.ty == 16);
}

hipError_t launch_stage_f64(int theory, int stage, const StageArgs<double>& a, const LaunchShape& s, hipStream_t st) {
  switch (theory) {
    case PION: return stage_launch<double, PION>(stage, a, s, st);
    case OMEGA: return stage_launch<double, OMEGA>(stage, a, s, st);
    case COULOMB: return stage_launch<double, COULOMB>(stage, a, s, st);
    case RHO: return stage_launch<double, RHO>(stage, a, s, st);
    case BERGER: return stage_launch<double, BERGER>(stage, a, s, st);
    case LIGHTLY: return stage_launch<double, LIGHTLY>(stage, a, s, st);
  }
  return hipErrorInvalidValue;
}
hipError_t launch_stage_f32(int theory, int stage, const StageArgs<float>& a, const LaunchShape& s, hipStream_t st) {
  switch (theory) {
    case PION: return stage_launch<float, PION>(stage, a, s, st);
    case OMEGA: return stage_launch<float, OMEGA>(stage, a, s, st);
    case COULOMB: return stage_launch<float, COULOMB>(stage, a, s, st);
    case RHO: return stage_launch<float, RHO>(stage, a, s, st);
    case BERGER: return stage_launch<float, BERGER>(stage, a, s, st);
    case LIGHTLY: return stage_launch<float, LIGHTLY>(stage, a, s, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_stage_batch_f64(int theory, int stage, const LaunchShape& s, const BatchSlot<double>* slots,
                                  const int* blk_inst, int map, int parity, int grid, hipStream_t st) {
  switch (theory) {
    case PION: return stage_launch_batch<double, PION>(stage, s, slots, blk_inst, map, parity, grid, st);
    case OMEGA: return stage_launch_batch<double, OMEGA>(stage, s, slots, blk_inst, map, parity, grid, st);
    case COULOMB: return stage_launch_batch<double, COULOMB>(stage, s, slots, blk_inst, map, parity, grid, st);
    case RHO: return stage_launch_batch<double, RHO>(stage, s, slots, blk_inst, map, parity, grid, st);
    case BERGER: return stage_launch_batch<double, BERGER>(stage, s, slots, blk_inst, map, parity, grid, st);
    case LIGHTLY: return stage_launch_batch<double, LIGHTLY>(stage, s, slots, blk_inst, map, parity, grid, st);
  }
  return hipErrorInvalidValue;
}
hipError_t launch_stage_batch_f32(int theory, int stage, const LaunchShape& s, const BatchSlot<float>* slots,
                                  const int* blk_inst, int map, int parity, int grid, hipStream_t st) {
  switch (theory) {
    case PION: return stage_launch_batch<float, PION>(stage, s, slots, blk_inst, map, parity, grid, st);
    case OMEGA: return stage_launch_batch<float, OMEGA>(stage, s, slots, blk_inst, map, parity, grid, st);
    case COULOMB: return stage_launch_batch<float, COULOMB>(stage, s, slots, blk_inst, map, parity, grid, st);
    case RHO: return stage_launch_batch<float, RHO>(stage, s, slots, blk_inst, map, parity, grid, st);
    case BERGER: return stage_launch_batch<float, BERGER>(stage, s, slots, blk_inst, map, parity, grid, st);
    case LIGHTLY: return stage_launch_batch<float, LIGHTLY>(stage, s, slots, blk_inst, map, parity, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t stage_blocks_per_cu(int prec_fp32, int theory, int stage, const LaunchShape& s, int* blocks) {
  switch (theory) {
    case PION: return prec_fp32 ? stage_occupancy<float, PION>(stage, s, blocks) : stage_occupancy<double, PION>(stage, s, blocks);
    case OMEGA: return prec_fp32 ? stage_occupancy<float, OMEGA>(stage, s, blocks) : stage_occupancy<double, OMEGA>(stage, s, blocks);
    case COULOMB: return prec_fp32 ? stage_occupancy<float, COULOMB>(stage, s, blocks) : stage_occupancy<double, COULOMB>(stage, s, blocks);
    case RHO: return prec_fp32 ? stage_occupancy<float, RHO>(stage, s, blocks) : stage_occupancy<double, RHO>(stage, s, blocks);
    case BERGER: return prec_fp32 ? stage_occupancy<float, BERGER>(stage, s, blocks) : stage_occupancy<double, BERGER>(stage, s, blocks);
    case LIGHTLY: return prec_fp32 ? stage_occupancy<float, LIGHTLY>(stage, s, blocks) : stage_occupancy<double, LIGHTLY>(stage, s, blocks);
  }
  return hipErrorInvalidValue;
}

}  // namespace skx
