// Stage kernels of the arrested-Newton-flow step for gfx950.
//
// One kernel template covers the energy pass (stage 0) and the four RK4
// gradient stages of one step.  Each stage is one pass over the lattice:
//
//   stage field   Temp_s = F + f_s * dt * (V + f'_s * l_{s-1})   (formed on load)
//   stage 1       l1 = -dt g(Temp_1)            [+ energy / baryon of F fused]
//   stage 2       l2 = -dt g(Temp_2)
//   stage 3       l3 = -dt g(Temp_3)
//   stage 4       l4 = -dt g(Temp_4); F', V' = RK4 combine; S^3 normalise and
//                 tangent projection; signed-max and max|g| reductions fused
//
// which is the data flow of core/integrator.py:382-409 (k_s and Temp are
// never stored; they are recomputed from F, V and the l_s on the fly).
//
// Tiling: lanes run along the unit-stride axis (reference z), a block covers
// a TZ x TY tile of the (y, z) plane and marches along the slowest axis
// (reference x) over a chunk of planes.  Each incoming plane of the stage
// field (with a 2-site y/z halo) is staged in a 4-slot LDS ring, so the
// centre plane (two behind the lead) is still resident for its y/z second
// differences; one barrier per plane.  Per-site values along x (the field
// and its raw y / z first differences) live in 5-deep register queues, and
// raw z-differences of every tile row go to a 3-slot LDS ring for the yz
// mixed derivative.  Mixed derivatives are first differences of raw first
// differences, which is the reference's 25-tap tensor stencil
// (core/derivatives.py:121-173) regrouped.
#include "skx_internal.h"

namespace skx {

template <int TH>
struct Traits;
template <>
struct Traits<PION> {
  static constexpr int NF = 4;
  static constexpr int NSK = 4;
  static constexpr unsigned MIXED = 0xFu;  // components needing mixed second derivatives
};
template <>
struct Traits<OMEGA> {
  static constexpr int NF = 5;
  static constexpr int NSK = 4;
  static constexpr unsigned MIXED = 0x0u;
};
template <>
struct Traits<COULOMB> {
  static constexpr int NF = 5;
  static constexpr int NSK = 4;
  static constexpr unsigned MIXED = 0xFu;
};

// (1/12 f[-2] - 2/3 f[-1] + 2/3 f[+1] - 1/12 f[+2]) / ls   (core/derivatives.py:55)
template <typename T>
__device__ __forceinline__ T st_d1(T fm2, T fm1, T fp1, T fp2, T ils) {
  return (T(1.0 / 12.0) * (fm2 - fp2) + T(2.0 / 3.0) * (fp1 - fm1)) * ils;
}
// (-1/12 f[-2] + 4/3 f[-1] - 5/2 f[0] + 4/3 f[+1] - 1/12 f[+2]) / ls^2   (core/derivatives.py:107)
template <typename T>
__device__ __forceinline__ T st_d2(T fm2, T fm1, T f0, T fp1, T fp2, T ils2) {
  return (T(4.0 / 3.0) * (fm1 + fp1) - T(1.0 / 12.0) * (fm2 + fp2) - T(2.5) * f0) * ils2;
}

template <int TH, typename T, int LD>
__device__ __forceinline__ void point_gradient(const T* phi, const T (*F)[LD], const T (*S)[LD], const PhysParams& P,
                                               T* g) {
  if constexpr (TH == PION) gradient_pion<T>(phi, F, S, P, g);
  if constexpr (TH == OMEGA) gradient_omega<T>(phi, F, S, P, g);
  if constexpr (TH == COULOMB) gradient_coulomb<T>(phi, F, S, P, g);
}

template <int TH, typename T, int LD>
__device__ __forceinline__ double point_energy(const T* phi, const T (*F)[LD], const PhysParams& P) {
  if constexpr (TH == PION) return energy_pion<T, LD>(phi, F, P);
  if constexpr (TH == OMEGA) return energy_omega<T, LD>(phi, F, P);
  if constexpr (TH == COULOMB) return energy_coulomb<T, LD>(phi, F, P);
  return 0.0;
}

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
template <int NT, bool IS_MAX>
__device__ bool grid_reduce2(double& r0, double& r1, double* partial, uint32_t* counter, double* red) {
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
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
    __threadfence();
    const uint32_t prev = atomicAdd(counter, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  const int nb = gridDim.x;
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

template <typename T, int TH, int STAGE, int TZ, int TY>
__global__ void __launch_bounds__(TZ* TY) k_stage(const StageArgs<T> A) {
  using Tr = Traits<TH>;
  constexpr int NF = Tr::NF;
  constexpr int NSK = Tr::NSK;
  constexpr int NT = TZ * TY;
  constexpr int TW = TZ + 4;  // tile row length (z with halo)
  constexpr int TR = TY + 4;  // tile rows (y with halo)
  constexpr int TE = TW * TR;
  constexpr int LPT = (TE + NT - 1) / NT;
  constexpr int NA = (STAGE <= 1) ? 1 : (STAGE == 2 ? 2 : 3);
  constexpr bool D2 = STAGE >= 1;
  constexpr bool ANY_MIXED = D2 && (Tr::MIXED != 0u);
  constexpr int FSLOTS = 4;  // stage-field planes: lead p, p-1, centre p-2, and the slot being refilled
  constexpr int ZSLOTS = 3;  // raw z-differences for the yz mixed derivative: planes p, p-1, p-2
  static_assert(NT % 64 == 0, "block must be whole wavefronts");
  static_assert(TY >= 4, "halo rows of the z-difference ring are filled by the first four rows of threads");

  __shared__ T fring[FSLOTS][NF][TE];
  __shared__ T zring[ANY_MIXED ? ZSLOTS : 1][ANY_MIXED ? NF : 1][ANY_MIXED ? TR * TZ : 1];
  __shared__ double red[2 * (NT / 64)];

  if ((A.flags & CHECK_DONE) && A.st->done) return;  // uniform per launch

  const int tid = threadIdx.x;
  const int tz = tid % TZ, ty = tid / TZ;
  int lin = blockIdx.x;
  const int bz = lin % A.nbz;
  lin /= A.nbz;
  const int by = lin % A.nby;
  const int bx = lin / A.nby;
  const int z0 = bz * TZ, y0 = by * TY;
  const int xs = bx * A.xchunk;
  const int xe = min(xs + A.xchunk, A.X);
  const int y = y0 + ty, z = z0 + tz;
  const int Y = A.Y, Z = A.Z, X = A.X, halo = A.halo;
  const bool own = (y < Y) && (z < Z);
  const long long N = A.N;
  const long long YZ = (long long)Y * Z;

  const bool gy1 = (y >= halo) && (y < Y - halo), gz1 = (z >= halo) && (z < Z - halo);
  const bool gy2 = (y > halo) && (y < Y - halo), gz2 = (z > halo) && (z < Z - halo);

  const T dt = T(A.P.dt);
  T keep = T(1);
  if constexpr (STAGE >= 2) keep = T(A.st->keep);
  const T ilsx = A.ilsx, ilsy = A.ilsy, ilsz = A.ilsz;
  const T ilsx2 = ilsx * ilsx, ilsy2 = ilsy * ilsy, ilsz2 = ilsz * ilsz;

  // tile elements this thread stages
  int eoff[LPT];
  bool evalid[LPT];
#pragma unroll
  for (int j = 0; j < LPT; ++j) {
    const int e = tid + j * NT;
    const int r = e / TW, cc = e - r * TW;
    const int gy = y0 - 2 + r, gz = z0 - 2 + cc;
    evalid[j] = (e < TE) && gy >= 0 && gy < Y && gz >= 0 && gz < Z;
    eoff[j] = gy * Z + gz;
  }
  // halo row of the z-difference ring this thread fills (rows 0,1 and TY+2,TY+3)
  const int hrow = (ty < 2) ? ty : (TY + ty);

  const T* src[3] = {A.F, A.V, A.la};
  T pre[NA][NF][LPT];
  auto issue = [&](int p) {
    const bool pin = (p >= 0) && (p < X);
    const long long pb = (long long)p * YZ;
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
      for (int a = 0; a < NF; ++a)
#pragma unroll
        for (int j = 0; j < LPT; ++j)
          pre[k][a][j] = (pin && evalid[j]) ? src[k][a * N + pb + eoff[j]] : T(0);
  };

  // own-column x-queues: index 0 = plane c-2 ... 4 = plane c+2 (lead)
  T qf[NF][5], qy[NF][5], qz[NF][5];
#pragma unroll
  for (int a = 0; a < NF; ++a)
#pragma unroll
    for (int k = 0; k < 5; ++k) qf[a][k] = qy[a][k] = qz[a][k] = T(0);

  double acc0 = (STAGE == 4) ? -1.0e300 : 0.0;
  double acc1 = 0.0;
  bool bad = false;

  issue(xs - 2);
  const int C = (ty + 2) * TW + (tz + 2);
  for (int p = xs - 2; p <= xe + 1; ++p) {
    const int fs = (p + 8) & (FSLOTS - 1);  // p >= -2
    // stage the stage-field plane p in LDS
#pragma unroll
    for (int a = 0; a < NF; ++a)
#pragma unroll
      for (int j = 0; j < LPT; ++j) {
        const int e = tid + j * NT;
        if (e < TE) {
          T v = pre[0][a][j];
          if constexpr (STAGE == 2) v = v + T(0.5) * (dt * (keep != T(0) ? pre[1][a][j] : T(0)));
          if constexpr (STAGE == 3)
            v = v + T(0.5) * (dt * ((keep != T(0) ? pre[1][a][j] : T(0)) + T(0.5) * pre[2][a][j]));
          if constexpr (STAGE == 4) v = v + dt * ((keep != T(0) ? pre[1][a][j] : T(0)) + T(0.5) * pre[2][a][j]);
          fring[fs][a][e] = v;
        }
      }
    __syncthreads();
    if (p + 1 <= xe + 1) issue(p + 1);

    // lead plane p: own-column value and raw first differences along y and z
    const int zs = (p + 9) % ZSLOTS;
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const T* t = fring[fs][a];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        qf[a][k] = qf[a][k + 1];
        qy[a][k] = qy[a][k + 1];
        qz[a][k] = qz[a][k + 1];
      }
      qf[a][4] = t[C];
      qy[a][4] = st_d1(t[C - 2 * TW], t[C - TW], t[C + TW], t[C + 2 * TW], ilsy);
      const T dz = st_d1(t[C - 2], t[C - 1], t[C + 1], t[C + 2], ilsz);
      qz[a][4] = dz;
      if constexpr (ANY_MIXED) {
        if ((Tr::MIXED >> a) & 1u) {
          zring[zs][a][(ty + 2) * TZ + tz] = dz;
          if (ty < 4) {
            const int h = hrow * TW + tz + 2;
            zring[zs][a][hrow * TZ + tz] = st_d1(t[h - 2], t[h - 1], t[h + 1], t[h + 2], ilsz);
          }
        }
      }
    }

    const int c = p - 2;
    if (c >= xs && own) {
      const bool gx1 = (c >= halo) && (c < X - halo), gx2 = (c > halo) && (c < X - halo);
      const int cfs = (c + 8) & (FSLOTS - 1);
      const int czs = (c + 9) % ZSLOTS;
      T phi[NF], F[3][NF], S[6][NF];
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        phi[a] = qf[a][2];
        F[0][a] = gx1 ? st_d1(qf[a][0], qf[a][1], qf[a][3], qf[a][4], ilsx) : T(0);
        F[1][a] = gy1 ? qy[a][2] : T(0);
        F[2][a] = gz1 ? qz[a][2] : T(0);
        if constexpr (D2) {
          const T* t = fring[cfs][a];
          const T f0 = phi[a];
          S[0][a] = gx1 ? st_d2(qf[a][0], qf[a][1], f0, qf[a][3], qf[a][4], ilsx2) : T(0);
          S[1][a] = gy1 ? st_d2(t[C - 2 * TW], t[C - TW], f0, t[C + TW], t[C + 2 * TW], ilsy2) : T(0);
          S[2][a] = gz1 ? st_d2(t[C - 2], t[C - 1], f0, t[C + 1], t[C + 2], ilsz2) : T(0);
          if ((Tr::MIXED >> a) & 1u) {
            S[3][a] = (gx2 && gy2) ? st_d1(qy[a][0], qy[a][1], qy[a][3], qy[a][4], ilsx) : T(0);
            S[4][a] = (gx2 && gz2) ? st_d1(qz[a][0], qz[a][1], qz[a][3], qz[a][4], ilsx) : T(0);
            if constexpr (ANY_MIXED) {
              const T* zr = zring[czs][a];
              S[5][a] = (gy2 && gz2) ? st_d1(zr[ty * TZ + tz], zr[(ty + 1) * TZ + tz], zr[(ty + 3) * TZ + tz],
                                             zr[(ty + 4) * TZ + tz], ilsy)
                                     : T(0);
            }
          } else {
            S[3][a] = S[4][a] = S[5][a] = T(0);
          }
        }
      }
      const long long so = (long long)c * YZ + (long long)y * Z + z;

      if constexpr (STAGE <= 1) {
        if (STAGE == 0 || (A.flags & FUSE_ENERGY)) {
          const double e = point_energy<TH, T, NF>(phi, F, A.P);
          acc0 += e;
          if (STAGE == 0 && A.dens_e) A.dens_e[so] = e;
        }
        if (A.flags & FUSE_BARYON) {
          const double b = baryon_density<T, NF>(phi, F, A.P);
          acc1 += b;
          if (STAGE == 0 && A.dens_b) A.dens_b[so] = b;
        }
      }
      if constexpr (D2) {
        T g[NF];
        point_gradient<TH, T, NF>(phi, F, S, A.P, g);
        if (A.eg_out) {
#pragma unroll
          for (int a = 0; a < NF; ++a) A.eg_out[a * N + so] = double(g[a]);
        }
        if constexpr (STAGE <= 3) {
#pragma unroll
          for (int a = 0; a < NF; ++a) A.lout[a * N + so] = -dt * g[a];
        } else {
          // RK4 combine (core/integrator.py:170-175), recomputing k1..k4 from V and l1..l3
          using R = double;
          R Fn[NF], Vn[NF];
          const R rdt = R(A.P.dt);
#pragma unroll
          for (int a = 0; a < NF; ++a) {
            const long long o = a * N + so;
            const R Fv = R(A.F[o]);
            const R Vk = (keep != T(0)) ? R(A.V[o]) : R(0);
            const R l1 = R(A.lb[o]), l2 = R(A.la[o]), l3 = R(A.lc[o]);
            const R l4 = R(-dt * g[a]);
            const R k1 = rdt * Vk;
            const R k2 = rdt * (Vk + 0.5 * l1);
            const R k3 = rdt * (Vk + 0.5 * l2);
            const R k4 = rdt * (Vk + l3);
            Vn[a] = Vk + (1.0 / 6.0) * (l1 + 2.0 * l2 + 2.0 * l3 + l4);
            Fn[a] = Fv + (1.0 / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
            const double gv = double(g[a]);
            acc0 = fmax(acc0, gv);
            acc1 = fmax(acc1, fabs(gv));
            bad |= !isfinite(gv);
          }
          if (A.unit_length) {
            R s = 0.0;
#pragma unroll
            for (int a = 0; a < NSK; ++a) s += Fn[a] * Fn[a];
            s = sqrt(s);
            if (s != 0.0) {
#pragma unroll
              for (int a = 0; a < NSK; ++a) Fn[a] /= s;
            }
            R lm = 0.0;
#pragma unroll
            for (int a = 0; a < NSK; ++a) lm += Vn[a] * Fn[a];
#pragma unroll
            for (int a = 0; a < NSK; ++a) Vn[a] -= lm * Fn[a];
          }
#pragma unroll
          for (int a = 0; a < NF; ++a) {
            A.Fout[a * N + so] = T(Fn[a]);
            A.Vout[a * N + so] = T(Vn[a]);
          }
        }
      }
    }
  }

  // ---- fused reductions
  if constexpr (STAGE <= 1) {
    if (STAGE == 1 && !(A.flags & (FUSE_ENERGY | FUSE_BARYON))) return;
    double r0 = acc0, r1 = acc1;
    if (grid_reduce2<NT, false>(r0, r1, A.partial, &A.st->cnt_sum, red)) {
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
    if (grid_reduce2<NT, true>(r0, r1, A.partial, &A.st->cnt_max, red)) {
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

template <typename T, int TH, int STAGE, int TZ, int TY>
static hipError_t launch_one(const StageArgs<T>& a, hipStream_t st) {
  StageArgs<T> b = a;
  b.nbz = (a.Z + TZ - 1) / TZ;
  b.nby = (a.Y + TY - 1) / TY;
  b.nbx = (a.X + a.xchunk - 1) / a.xchunk;
  const int blocks = b.nbz * b.nby * b.nbx;
  hipLaunchKernelGGL((k_stage<T, TH, STAGE, TZ, TY>), dim3(blocks), dim3(TZ * TY), 0, st, b);
  return hipGetLastError();
}

template <typename T, int TH, int TZ, int TY>
static hipError_t launch_by_stage(int stage, const StageArgs<T>& a, hipStream_t st) {
  switch (stage) {
    case 0: return launch_one<T, TH, 0, TZ, TY>(a, st);
    case 1: return launch_one<T, TH, 1, TZ, TY>(a, st);
    case 2: return launch_one<T, TH, 2, TZ, TY>(a, st);
    case 3: return launch_one<T, TH, 3, TZ, TY>(a, st);
    case 4: return launch_one<T, TH, 4, TZ, TY>(a, st);
  }
  return hipErrorInvalidValue;
}

template <typename T, int TZ, int TY>
static hipError_t launch_by_theory(int theory, int stage, const StageArgs<T>& a, hipStream_t st) {
  switch (theory) {
    case PION: return launch_by_stage<T, PION, TZ, TY>(stage, a, st);
    case OMEGA: return launch_by_stage<T, OMEGA, TZ, TY>(stage, a, st);
    case COULOMB: return launch_by_stage<T, COULOMB, TZ, TY>(stage, a, st);
  }
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t launch_stage_t(int theory, int stage, const StageArgs<T>& a, const LaunchShape& s, hipStream_t st) {
  StageArgs<T> b = a;
  b.xchunk = s.xchunk;
  if (s.tz == 64 && s.ty == 4) return launch_by_theory<T, 64, 4>(theory, stage, b, st);
  if (s.tz == 32 && s.ty == 8) return launch_by_theory<T, 32, 8>(theory, stage, b, st);
  if (s.tz == 16 && s.ty == 16) return launch_by_theory<T, 16, 16>(theory, stage, b, st);
  return hipErrorInvalidValue;
}

hipError_t launch_stage_f64(int theory, int stage, const StageArgs<double>& a, const LaunchShape& s, hipStream_t st) {
  return launch_stage_t<double>(theory, stage, a, s, st);
}
hipError_t launch_stage_f32(int theory, int stage, const StageArgs<float>& a, const LaunchShape& s, hipStream_t st) {
  return launch_stage_t<float>(theory, stage, a, s, st);
}

}  // namespace skx
