// Post-convergence observables: per-site integrands of the reference's
// observable kernels (theories/nuclear_skyrme/kernels.py:485-1033; centre of
// mass and RMS radius also omega / Coulomb / rho), evaluated in FP64 from the
// stored field and summed per block; the host folds the block partials in a
// fixed, compensated order.
//
// Kinds and the values summed (grid x = lsx * i, create_grid_kernel :50-56;
// b = baryon density times vol; r = x - com):
//   COM   4: b, x b, y b, z b                              (:678-718)
//   RMS   1: r^2 |b|                                       (:720-759)
//   QUAD  9: (3 r_i r_j - d_ij r^2) b / 2                  (:761-812)
//   U     9: 2 U0_ij vol                                   (:814-851)
//   V     9: 2 eps_ilm eps_jnp r_l r_n A_mp vol            (:853-905)
//   W     9: 2 eps_jlm r_l B_im vol                        (:907-975)
//   D     1: (r_i r_j T_ij - r^2 tr T / 3) vol             (:977-1033)
#include <hip/hip_runtime.h>

#include <mutex>

#include "skx_internal.h"

namespace skx {

// Block partials of the allocation-free observable entry point
// (skx_observable_noalloc): kObsMaxBlocks blocks x 9 values, 147456 bytes on
// every device the library runs on, held until the process ends.  Host code
// never uses this name as an address: it takes the per-device address from
// observable_scratch() below.
static __device__ __align__(256) double scratch[kObsMaxBlocks * 9];

namespace {

constexpr int kObsThreads = 256;

// first derivative along one axis at index i of n, stride s (4th order,
// zero within `halo` of either end: core/derivatives.py:17-67)
template <typename T>
__device__ inline double d1_at(const T* f, long long s, int i, int n, int halo, double ils) {
  if (i < halo || i >= n - halo) return 0.0;
  const double fm2 = double(f[-2 * s]), fm1 = double(f[-s]), fp1 = double(f[s]), fp2 = double(f[2 * s]);
  return ((fm2 - fp2) * (1.0 / 12.0) + (fp1 - fm1) * (2.0 / 3.0)) * ils;
}

__device__ inline double e3(int a, int b, int c) { return eps3(a, b, c); }

// Per-site products of the left currents L_k^a = sigma d_k pi_a - pi_a d_k sigma
// + (pi x d_k pi)_a and the isospin currents T_i^a = pi_i pi_a - d_ia |pi|^2
// + sigma e_iab pi_b (_build_left_and_iso_currents, kernels.py:485-547):
// LL_ij = L_i.L_j, TT_ij = T_i.T_j, LT_ij = T_i.L_j.
__device__ inline void current_products(const double phi[4], const double (*Fd)[4], double LL[3][3], double TT[3][3],
                                        double LT[3][3], bool want_T) {
  const double sg = phi[0], p[3] = {phi[1], phi[2], phi[3]};
  double L[3][3];
  for (int k = 0; k < 3; ++k) {
    const double ds = Fd[k][0], dp[3] = {Fd[k][1], Fd[k][2], Fd[k][3]};
    const double cr[3] = {p[1] * dp[2] - p[2] * dp[1], p[2] * dp[0] - p[0] * dp[2], p[0] * dp[1] - p[1] * dp[0]};
    for (int a = 0; a < 3; ++a) L[k][a] = sg * dp[a] - p[a] * ds + cr[a];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) LL[i][j] = L[i][0] * L[j][0] + L[i][1] * L[j][1] + L[i][2] * L[j][2];
  if (!want_T) return;
  const double psq = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
  double Tc[3][3];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 3; ++a) {
      double et = 0.0;
      for (int b = 0; b < 3; ++b) et += e3(i, a, b) * p[b];
      Tc[i][a] = p[i] * p[a] - (i == a ? psq : 0.0) + sg * et;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      TT[i][j] = Tc[i][0] * Tc[j][0] + Tc[i][1] * Tc[j][1] + Tc[i][2] * Tc[j][2];
      LT[i][j] = Tc[i][0] * L[j][0] + Tc[i][1] * L[j][1] + Tc[i][2] * L[j][2];
    }
}

template <typename T, int KIND>
__device__ void site_values(const ObsArgs& A, const T* F, long long s, int x, int y, int z, double* v) {
  const long long N = (long long)A.X * A.Y * A.Z;
  const long long YZ = (long long)A.Y * A.Z;
  double phi[4], Fd[3][4];
  for (int a = 0; a < 4; ++a) {
    const T* fa = F + a * N + s;
    phi[a] = double(*fa);
    Fd[0][a] = d1_at(fa, YZ, x, A.X, A.halo, 1.0 / A.lsx);
    Fd[1][a] = d1_at(fa, (long long)A.Z, y, A.Y, A.halo, 1.0 / A.lsy);
    Fd[2][a] = d1_at(fa, 1, z, A.Z, A.halo, 1.0 / A.lsz);
  }
  const double r[3] = {A.lsx * x - A.com[0], A.lsy * y - A.com[1], A.lsz * z - A.com[2]};
  const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  const double k2 = A.P.k2, vol = A.P.vol;
  if constexpr (KIND == SKX_OBS_COM || KIND == SKX_OBS_RMS || KIND == SKX_OBS_QUADRUPOLE) {
    const double b = baryon_density<double, 4>(phi, Fd, A.P);
    if constexpr (KIND == SKX_OBS_COM) {
      v[0] = b;
      v[1] = (A.lsx * x) * b;
      v[2] = (A.lsy * y) * b;
      v[3] = (A.lsz * z) * b;
    } else if constexpr (KIND == SKX_OBS_RMS) {
      v[0] = r2 * fabs(b);
    } else {
      const double rho = 0.5 * b;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v[3 * i + j] = (3.0 * r[i] * r[j] - (i == j ? r2 : 0.0)) * rho;
    }
  } else if constexpr (KIND == SKX_OBS_U || KIND == SKX_OBS_V) {
    // U0 and A of _compute_U0_and_A (kernels.py:549-613)
    double LL[3][3], TT[3][3], LT[3][3];
    current_products(phi, Fd, LL, TT, LT, KIND == SKX_OBS_U);
    double M[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double q = 0.0;
        if constexpr (KIND == SKX_OBS_U) {
          for (int k = 0; k < 3; ++k) q += TT[i][j] * LL[k][k] - LT[i][k] * LT[j][k];
          M[i][j] = TT[i][j] + k2 * q;
        } else {
          for (int k = 0; k < 3; ++k) q += LL[i][j] * LL[k][k] - LL[i][k] * LL[j][k];
          M[i][j] = LL[i][j] + k2 * q;
        }
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        if constexpr (KIND == SKX_OBS_U) {
          v[3 * i + j] = 2.0 * M[i][j] * vol;
        } else {
          double val = 0.0;
          for (int l = 0; l < 3; ++l)
            for (int n = 0; n < 3; ++n)
              for (int m = 0; m < 3; ++m)
                for (int q = 0; q < 3; ++q) {
                  const double e = e3(i, l, m) * e3(j, n, q);
                  if (e != 0.0) val += e * r[l] * r[n] * M[m][q];
                }
          v[3 * i + j] = 2.0 * val * vol;
        }
      }
  } else if constexpr (KIND == SKX_OBS_W) {
    double LL[3][3], TT[3][3], LT[3][3];
    current_products(phi, Fd, LL, TT, LT, true);
    double Bm[3][3];
    for (int i = 0; i < 3; ++i)
      for (int m = 0; m < 3; ++m) {
        double val = LT[i][m];
        for (int k = 0; k < 3; ++k) val += k2 * (LT[i][m] * LL[k][k] - LT[i][k] * LL[m][k]);
        Bm[i][m] = val;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double val = 0.0;
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) {
            const double e = e3(j, l, m);
            if (e != 0.0) val += e * r[l] * Bm[i][m];
          }
        v[3 * i + j] = 2.0 * val * vol;
      }
  } else {
    // SKX_OBS_D: _compute_stress_tensor (kernels.py:615-676) and the D-term integrand
    double LL[3][3], TT[3][3], LT[3][3];
    current_products(phi, Fd, LL, TT, LT, false);
    const double trD = LL[0][0] + LL[1][1] + LL[2][2];
    double quartic = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = i + 1; j < 3; ++j) quartic += LL[i][i] * LL[j][j] - LL[i][j] * LL[i][j];
    const double e = 0.5 * trD + 0.5 * k2 * quartic + A.P.m2 * (1.0 - phi[0]);
    double Ts[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double D2 = LL[i][0] * LL[0][j] + LL[i][1] * LL[1][j] + LL[i][2] * LL[2][j];
        Ts[i][j] = LL[i][j] + k2 * (trD * LL[i][j] - D2) - (i == j ? e : 0.0);
      }
    double sc = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) sc += r[i] * r[j] * Ts[i][j];
    sc -= r2 * (Ts[0][0] + Ts[1][1] + Ts[2][2]) / 3.0;
    v[0] = sc * vol;
  }
}

template <int KIND>
constexpr int obs_count() {
  return KIND == SKX_OBS_COM ? 4 : (KIND == SKX_OBS_RMS || KIND == SKX_OBS_D) ? 1 : 9;
}

template <typename T, int KIND>
__global__ void __launch_bounds__(kObsThreads) k_observable(const T* __restrict__ F, double* __restrict__ partial,
                                                           ObsArgs A) {
  constexpr int MAXV = obs_count<KIND>();
  __shared__ double red[kObsThreads / 64][MAXV];
  double acc[MAXV];
  for (int q = 0; q < MAXV; ++q) acc[q] = 0.0;
  const long long N = (long long)A.X * A.Y * A.Z;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < N; s += stride) {
    const int z = (int)(s % A.Z);
    const int y = (int)((s / A.Z) % A.Y);
    const int x = (int)(s / ((long long)A.Z * A.Y));
    double v[MAXV];
    for (int q = 0; q < MAXV; ++q) v[q] = 0.0;
    site_values<T, KIND>(A, F, s, x, y, z, v);
    for (int q = 0; q < MAXV; ++q) acc[q] += v[q];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < MAXV; ++q) {
    double t = acc[q];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (lane == 0) red[w][q] = t;
  }
  __syncthreads();
  if (threadIdx.x < MAXV) {
    double t = 0.0;
    for (int k = 0; k < kObsThreads / 64; ++k) t += red[k][threadIdx.x];
    partial[(size_t)blockIdx.x * MAXV + threadIdx.x] = t;
  }
}

}  // namespace

int observable_blocks(long long N) {
  const long long b = (N + kObsThreads - 1) / kObsThreads;
  return (int)(b < kObsMaxBlocks ? b : kObsMaxBlocks);
}

namespace {

template <typename T>
hipError_t launch_observable(const T* F, double* partial, const ObsArgs& a, int blocks, hipStream_t st) {
  const dim3 g(blocks), b(kObsThreads);
  switch (a.kind) {
    case SKX_OBS_COM: hipLaunchKernelGGL((k_observable<T, SKX_OBS_COM>), g, b, 0, st, F, partial, a); break;
    case SKX_OBS_RMS: hipLaunchKernelGGL((k_observable<T, SKX_OBS_RMS>), g, b, 0, st, F, partial, a); break;
    case SKX_OBS_QUADRUPOLE:
      hipLaunchKernelGGL((k_observable<T, SKX_OBS_QUADRUPOLE>), g, b, 0, st, F, partial, a);
      break;
    case SKX_OBS_U: hipLaunchKernelGGL((k_observable<T, SKX_OBS_U>), g, b, 0, st, F, partial, a); break;
    case SKX_OBS_V: hipLaunchKernelGGL((k_observable<T, SKX_OBS_V>), g, b, 0, st, F, partial, a); break;
    case SKX_OBS_W: hipLaunchKernelGGL((k_observable<T, SKX_OBS_W>), g, b, 0, st, F, partial, a); break;
    case SKX_OBS_D: hipLaunchKernelGGL((k_observable<T, SKX_OBS_D>), g, b, 0, st, F, partial, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_observable_f64(const double* F, double* partial, const ObsArgs& a, int blocks, hipStream_t st) {
  return launch_observable<double>(F, partial, a, blocks, st);
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on `device`, which the caller has made current.
// Looked up with hipGetSymbolAddress on the first call per device and kept:
// each device holds its own copy of the array at its own address.
hipError_t observable_scratch(int device, double** ptr, size_t* bytes) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess) return r;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *ptr = g_scratch_ptr[device];
  *bytes = sizeof(scratch);
  return hipSuccess;
}
hipError_t launch_observable_f32(const float* F, double* partial, const ObsArgs& a, int blocks, hipStream_t st) {
  return launch_observable<float>(F, partial, a, blocks, st);
}

}  // namespace skx
