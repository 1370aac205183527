// Device helpers shared by the stage kernels: reductions, stencil
// coefficients and buffer addressing.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "skx_internal.h"

namespace skx {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double load_partial(const double* p) {
  unsigned long long u =
      __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(static_cast<long long>(u));
}

// Two-value block reduction, then the last block to finish folds every
// block's partials in a fixed order (deterministic, no memset needed).
// IS_MAX selects max instead of (Neumaier-compensated) sum.  Returns true in
// thread 0 of the last block, with the grid-wide results in r0 / r1.
//
// Only device-scope atomics carry data between blocks, so no memory fence is
// needed: a block publishes its partials with atomic exchanges and increments
// the counter only after both exchanges have returned (performed at the
// device's coherence point); the block that takes the counter to nblk then
// reads every partial with device-scope atomic loads.  A device-scope
// release / acquire fence would instead write back and invalidate the whole
// L2 of the block's XCD, which cost about 7 us per launch at 256 blocks on
// the 128^3 stage kernels.  The status fields the last block writes are read
// by later kernels only.
template <int NT, bool IS_MAX>
__device__ bool grid_reduce2(double& r0, double& r1, double* partial, uint32_t* counter, double* red, int blk,
                             int nblk) {
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double a = IS_MAX ? wave_max(r0) : wave_sum(r0);
  double b = IS_MAX ? wave_max(r1) : wave_sum(r1);
  if (lane == 0) {
    red[w] = a;
    red[NW + w] = b;
  }
  __syncthreads();
  __shared__ bool is_last;
  if (tid == 0) {
    double s0 = red[0], s1 = red[NW];
    for (int i = 1; i < NW; ++i) {
      s0 = IS_MAX ? fmax(s0, red[i]) : s0 + red[i];
      s1 = IS_MAX ? fmax(s1, red[NW + i]) : s1 + red[NW + i];
    }
    auto* pu = reinterpret_cast<unsigned long long*>(partial) + 2 * blk;
    const unsigned long long u0 = __hip_atomic_exchange(pu, (unsigned long long)__double_as_longlong(s0),
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long u1 = __hip_atomic_exchange(pu + 1, (unsigned long long)__double_as_longlong(s1),
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the increment waits for both exchanges' returned values (and no memory access moves across it)
    asm volatile("" ::"v"(u0), "v"(u1) : "memory");
    const uint32_t prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (prev == (uint32_t)nblk - 1u);
  }
  __syncthreads();
  if (!is_last) return false;
  const int nb = nblk;
  double s0 = IS_MAX ? -1.0e300 : 0.0, c0 = 0.0;
  double s1 = IS_MAX ? -1.0e300 : 0.0, c1 = 0.0;
  for (int i = tid; i < nb; i += NT) {
    const double v0 = load_partial(partial + 2 * i);
    const double v1 = load_partial(partial + 2 * i + 1);
    if (IS_MAX) {
      s0 = fmax(s0, v0);
      s1 = fmax(s1, v1);
    } else {  // Neumaier compensated summation
      double t = s0 + v0;
      c0 += (fabs(s0) >= fabs(v0)) ? ((s0 - t) + v0) : ((v0 - t) + s0);
      s0 = t;
      t = s1 + v1;
      c1 += (fabs(s1) >= fabs(v1)) ? ((s1 - t) + v1) : ((v1 - t) + s1);
      s1 = t;
    }
  }
  if (!IS_MAX) {
    s0 += c0;
    s1 += c1;
  }
  __syncthreads();
  a = IS_MAX ? wave_max(s0) : wave_sum(s0);
  b = IS_MAX ? wave_max(s1) : wave_sum(s1);
  if (lane == 0) {
    red[w] = a;
    red[NW + w] = b;
  }
  __syncthreads();
  if (tid != 0) return false;
  s0 = red[0];
  s1 = red[NW];
  for (int i = 1; i < NW; ++i) {
    s0 = IS_MAX ? fmax(s0, red[i]) : s0 + red[i];
    s1 = IS_MAX ? fmax(s1, red[NW + i]) : s1 + red[NW + i];
  }
  *counter = 0u;  // ready for the next launch
  r0 = s0;
  r1 = s1;
  return true;
}

// The end of a stage launch, after every block's march: stages 0 / 1 reduce
// the energy and baryon sums of the incoming field and take the pending
// arrest decision; stage 4 reduces the signed max and max|g| (+inf marks a
// non-finite gradient) and counts the step.  Thread 0 of the last block to
// finish updates the lattice's DevStatus.
template <int NT, int STAGE, typename T>
__device__ __forceinline__ void stage_epilogue(const StageArgs<T>& A, double acc0, double acc1, bool bad,
                                               double* red, int blk, int nblk) {
  if constexpr (STAGE <= 1) {
    if (STAGE == 1 && !(A.flags & (FUSE_ENERGY | FUSE_BARYON))) return;
    double r0 = acc0, r1 = acc1;
    if (grid_reduce2<NT, false>(r0, r1, A.partial, &A.st->cnt_sum, red, blk, nblk)) {
      DevStatus* st = A.st;
      if (A.flags & FUSE_ENERGY) st->energy = r0;
      if (A.flags & FUSE_BARYON) st->baryon = r1;
      if (!isfinite(r0) || !isfinite(r1)) {
        st->nonfinite = 1;
        st->done = 1;
      }
      if (A.flags & DECIDE_ARREST) {
        double kp = 1.0;
        if (st->arrest_pending && st->killkinen && (r0 > st->e_prev)) kp = 0.0;
        st->keep = kp;
        st->arrest_pending = 0;
        st->e_prev = r0;
      }
      if (A.hist && (A.flags & DECIDE_ARREST) && st->steps < st->hist_cap) {
        A.hist[st->steps].energy = r0;
        A.hist[st->steps].baryon = r1;
      }
    }
  } else if constexpr (STAGE == 4) {
    // +inf in the max|g| slot marks a non-finite gradient
    double r0 = acc0, r1 = bad ? __longlong_as_double(0x7ff0000000000000LL) : acc1;
    if (grid_reduce2<NT, true>(r0, r1, A.partial, &A.st->cnt_max, red, blk, nblk)) {
      DevStatus* st = A.st;
      st->err_signed = r0;
      st->err_abs = r1;
      if (!isfinite(r1)) {
        st->nonfinite = 1;
        st->done = 1;
      }
      if (A.flags & RUN_MODE) {
        if (A.hist && st->steps < st->hist_cap) {
          A.hist[st->steps].err_signed = r0;
          A.hist[st->steps].err_abs = r1;
        }
        st->steps += 1;
        st->arrest_pending = 1;
        if (r1 < st->tol || st->steps >= st->max_steps) st->done = 1;
      }
    }
  }
}

template <int K>
using IC = std::integral_constant<int, K>;

// Stencil coefficients of one axis with the inverse spacing folded in.
template <typename T>
struct AxisCoef {
  T d1a, d1b;       // D1 = d1a (f[-2] - f[+2]) + d1b (f[+1] - f[-1])
  T d2a, d2b, d2c;  // D2 = d2a (f[-1] + f[+1]) - d2b (f[-2] + f[+2]) - d2c f[0]
  __device__ void set(T ils) {
    d1a = T(1.0 / 12.0) * ils;
    d1b = T(2.0 / 3.0) * ils;
    const T i2 = ils * ils;
    d2a = T(4.0 / 3.0) * i2;
    d2b = T(1.0 / 12.0) * i2;
    d2c = T(2.5) * i2;
  }
};

template <typename T>
__device__ __forceinline__ T cd1(T fm2, T fm1, T fp1, T fp2, T a, T b) {
  return a * (fm2 - fp2) + b * (fp1 - fm1);
}
template <typename T>
__device__ __forceinline__ T cd2(T fm2, T fm1, T f0, T fp1, T fp2, T a, T b, T c) {
  return a * (fm1 + fp1) - b * (fm2 + fp2) - c * f0;
}

// XCD-aware block order: the tile that block `blk` of a lattice's `nblk`
// blocks takes.  Workgroups are dispatched round-robin over the 8 XCDs, so
// the blocks whose index within the lattice's run of blocks agrees mod 8
// share an XCD and its L2 (also in a batch, whose run starts anywhere).
// Block w = 8q + j takes tile j Q + min(j, r) + q of the nblk = 8Q + r tiles:
// each XCD gets a contiguous range of tiles, so neighbouring tiles read each
// other's y and z halo lines from the same L2.  Without it (or with only
// nblk % 8 == 0 remapped, at 160^3 where the launches have 700 and 250
// blocks) the pion FP32 stage kernels fetched 2.5-4.3x their algorithmic bytes.
__device__ __forceinline__ int xcd_tile_order(int blk, int nblk) {
  const int j = blk & 7, q = blk >> 3;
  return j * (nblk >> 3) + min(j, nblk & 7) + q;
}

// Position of one site of a stage kernel's tile: the block's (y, z) tile and
// x-chunk in the XCD-aware order, and the site `s` of the TZ x TY tile.
struct TileSite {
  int z0, y0;  // tile origin
  int xs, xe;  // planes [xs, xe) of the block's x-chunk
  int ty, tz;  // site within the tile
  int y, z;    // lattice coordinates of the site
  bool own;    // the site lies in the lattice
};
template <int TZ, int TY, typename T>
__device__ __forceinline__ TileSite tile_site(const StageArgs<T>& A, int blk, int nblk, int s) {
  TileSite t;
  t.tz = s % TZ;
  t.ty = s / TZ;
  int lin = xcd_tile_order(blk, nblk);
  const int bz = lin % A.nbz;
  lin /= A.nbz;
  const int by = lin % A.nby;
  const int bx = lin / A.nby;
  t.z0 = bz * TZ;
  t.y0 = by * TY;
  t.xs = bx * A.xchunk;
  t.xe = min(t.xs + A.xchunk, A.X);
  t.y = t.y0 + t.ty;
  t.z = t.z0 + t.tz;
  t.own = (t.y < A.Y) && (t.z < A.Z);
  return t;
}

// Halo cell h of a tile with TR rows of TW cells (2-cell halo): rows 0, 1,
// TR-2, TR-1 in full, then columns 0, 1, TW-2, TW-1 of the rows between.
template <int TW, int TR>
__device__ __forceinline__ void halo_cell(int h, int& r, int& cc) {
  if (h < 2 * TW) {
    r = h / TW;
    cc = h - r * TW;
  } else if (h < 4 * TW) {
    const int m = h - 2 * TW;
    r = TR - 2 + m / TW;
    cc = m % TW;
  } else {
    const int m = h - 4 * TW, q = m & 3;
    r = 2 + m / 4;
    cc = q < 2 ? q : TW - 4 + q;
  }
}

// Stencil coefficients of one site with the guards of core/derivatives.py:54,
// 106,121 folded in along y and z.  The first differences along y and z are
// formed with their guards (halo <= i < n - halo) in the coefficients, so
// they are 0 off the band at once; the mixed derivatives built from them carry
// the stricter guards (halo < i < n - halo), so an unguarded difference is
// never needed.  The x guards change per plane and are applied there.
template <typename T>
struct SiteCoef {
  AxisCoef<T> cx, cy, cz;
  T y1a, y1b, z1a, z1b;           // first differences along y, z
  T y2a, y2b, y2c, z2a, z2b, z2c; // pure second differences along y, z
  T mxy_a, mxy_b;                 // x-part of d_x d_y, y guard folded
  T mxz_a, mxz_b;                 // x-part of d_x d_z, z guard folded
  T myz_a, myz_b;                 // y-part of d_y d_z, y and z guards folded
  __device__ __forceinline__ SiteCoef(const StageArgs<T>& A, int y, int z) {
    const int Y = A.Y, Z = A.Z, halo = A.halo;
    const bool gy1 = (y >= halo) && (y < Y - halo), gz1 = (z >= halo) && (z < Z - halo);
    const bool gy2 = (y > halo) && (y < Y - halo), gz2 = (z > halo) && (z < Z - halo);
    cx.set(A.ilsx);
    cy.set(A.ilsy);
    cz.set(A.ilsz);
    y1a = gy1 ? cy.d1a : T(0);
    y1b = gy1 ? cy.d1b : T(0);
    z1a = gz1 ? cz.d1a : T(0);
    z1b = gz1 ? cz.d1b : T(0);
    y2a = gy1 ? cy.d2a : T(0);
    y2b = gy1 ? cy.d2b : T(0);
    y2c = gy1 ? cy.d2c : T(0);
    z2a = gz1 ? cz.d2a : T(0);
    z2b = gz1 ? cz.d2b : T(0);
    z2c = gz1 ? cz.d2c : T(0);
    mxy_a = gy2 ? cx.d1a : T(0);
    mxy_b = gy2 ? cx.d1b : T(0);
    mxz_a = gz2 ? cx.d1a : T(0);
    mxz_b = gz2 ? cx.d1b : T(0);
    myz_a = (gy2 && gz2) ? cy.d1a : T(0);
    myz_b = (gy2 && gz2) ? cy.d1b : T(0);
  }
};

// RK4 stage field at one site, formed on load from the field f, the velocity
// v (ignored while the arrest holds, keep == 0) and the previous stage's
// l = -dt g (core/integrator.py:382-396: Temp = F + f_s k_{s-1},
// k_s = dt (V + f_s l_{s-1}), f = 0, 1/2, 1/2, 1).
template <int STAGE, typename T>
__device__ __forceinline__ T stage_value(T f, T v, T l, T dt, T keep) {
  const T vk = keep != T(0) ? v : T(0);
  if constexpr (STAGE == 2) return f + T(0.5) * (dt * vk);
  else if constexpr (STAGE == 3) return f + T(0.5) * (dt * (vk + T(0.5) * l));
  else if constexpr (STAGE == 4) return f + dt * (vk + T(0.5) * l);
  else return f;
}

// The stage field of one stage with its coefficients formed once per launch.
// FP64 form: stage_value, the reference's operand order.  FP32 form:
//   stage 2  f + (dt/2 keep) v,  stage 3  f + (dt/2 keep) v + (dt/4) l,  stage 4  f + (dt keep) v + (dt/2) l
// as two fused multiply-adds (the same sums, rounded once per term).
template <typename T, int STAGE>
struct StageCoef {
  T dt, keep, cv, cl;
  __device__ StageCoef(T dt_, T keep_) : dt(dt_), keep(keep_) {
    const T fv = STAGE == 4 ? T(1) : T(0.5), fl = STAGE == 4 ? T(0.5) : T(0.25);
    cv = keep_ != T(0) ? fv * dt_ : T(0);
    cl = fl * dt_;
  }
  __device__ __forceinline__ T operator()(T f, T v, T l) const {
    if constexpr (sizeof(T) == 8 || STAGE <= 1) {
      return stage_value<STAGE>(f, v, l, dt, keep);
    } else if constexpr (STAGE == 2) {
      return fma(cv, v, f);
    } else {
      return fma(cl, l, fma(cv, v, f));
    }
  }
};

// Buffer addressing: one resource per (array, component, plane) with the
// plane's byte size as the range, so every load / store takes a 32-bit lane
// offset and the scalar base; offsets at or beyond the range (OOB) read 0 and
// drop stores in hardware, which is how halo cells outside the lattice are
// zeroed.
constexpr uint32_t kOOB = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
}
template <typename T>
__device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, off, 0, 0);
  } else {
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, 0, 0);
  }
}


}  // namespace skx
