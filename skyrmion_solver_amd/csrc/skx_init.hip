// Device-side initial data: hedgehog and rational-map Skyrmions and their
// quaternion products, evaluated per site in FP64 and stored in the lattice's
// precision.  The algorithm is the host builder's
// (skyrmion_solver_amd/initial_data.py), which restates
// theories/nuclear_skyrme/initial_config.py:284-627 (maps, profile fits,
// quaternion product) and :701-1126 (hedgehog, rational map, product ansatz).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skx_internal.h"

namespace skx {

namespace {

// profilefun fits (initial_config.py:380-438): massless, then massive pion,
// for B = 1..9; coefficients a..e of r (a + r (b + r (c + r (d + r e)))).
__constant__ double kFitMassless[9][5] = {
    {0.630647, 0.0218873, -0.0245431, 0.00300978, -0.00010542},
    {0.162388, 0.297889, -0.0878554, 0.00931881, -0.000317398},
    {-0.000352877, 0.324957, -0.0792377, 0.00740042, -0.000232486},
    {-0.0469393, 0.294323, -0.0599714, 0.00457823, -0.000110399},
    {-0.0575367, 0.187057, -0.00254209, -0.00580857, 0.000517839},
    {-0.0450119, 0.104984, 0.0347624, -0.0119037, 0.000864464},
    {-0.0383968, 0.0703267, 0.0464174, -0.0130402, 0.000887014},
    {-0.013271, -0.00611001, 0.0751757, -0.0170388, 0.0010835},
    {0.00853506, -0.0576792, 0.0896121, -0.0181863, 0.00109042}};
__constant__ double kFitMassive[9][5] = {
    {0.839091, -0.148868, 0.222428, -0.0900171, 0.0107864},
    {0.219378, 0.485652, -0.154501, 0.0197873, -0.000774393},
    {0.0354001, 0.502948, -0.13181, 0.0142808, -0.000509656},
    {-0.0127361, 0.4427, -0.0941955, 0.00808476, -0.00024485},
    {-0.0568381, 0.337281, -0.0167685, -0.010309, 0.00119672},
    {-0.05592, 0.229393, 0.0498254, -0.0257167, 0.00248365},
    {-0.0575166, 0.196602, 0.056918, -0.0244016, 0.00219296},
    {-0.0335398, 0.0655383, 0.135378, -0.0430527, 0.00381564},
    {-0.0101412, -0.0285699, 0.181351, -0.0517179, 0.00441035}};

struct cplx {
  double re, im;
};
__device__ inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline cplx cscale(double s, cplx a) { return {s * a.re, s * a.im}; }
// complex division with 1e-18 added to |den|^2 (initial_config.py:249-250)
__device__ inline cplx cdiv(cplx n, cplx d) {
  const double q = d.re * d.re + d.im * d.im + 1.0e-18;
  return {(n.re * d.re + n.im * d.im) / q, (n.im * d.re - n.re * d.im) / q};
}

// R(z) for B = 1..8 (initial_config.py:446-549), z^B otherwise; 1e-9 added to
// the denominators the reference regularises.
__device__ cplx rational_map(cplx z, int b) {
  const double s3 = 1.7320508075688772;
  const double eps = 1.0e-9;
  const cplx one{1.0, 0.0};
  const cplx z2 = cmul(z, z);
  if (b == 1) return z;
  if (b == 2) return z2;
  if (b == 3) {
    const cplx num = csub(cmul(cplx{0.0, s3}, z2), one);
    cplx den = cmul(z, csub(z2, cplx{0.0, s3}));
    den.re += eps;
    return cdiv(num, den);
  }
  const cplx z4 = cmul(z2, z2);
  if (b == 4) {
    const cplx t = cmul(cplx{0.0, 2.0 * s3}, z2);
    return cdiv(cadd(cadd(z4, t), one), cadd(csub(z4, t), one));
  }
  if (b == 5) {
    const double a = 3.0660, bb = 3.9334;
    const cplx num = cmul(z, cadd(cadd(z4, cscale(bb, z2)), cplx{a, 0.0}));
    const cplx den = cadd(csub(cscale(a, z4), cscale(bb, z2)), one);
    return cdiv(num, den);
  }
  if (b == 6) {
    const double a = 0.1585;
    cplx den = cmul(z2, cadd(cmul(cplx{0.0, a}, z4), one));
    den.re += eps;
    return cdiv(cadd(z4, cplx{0.0, a}), den);
  }
  if (b == 7) {
    const cplx z5 = cmul(z4, z), z7 = cmul(z5, z2);
    const cplx num = csub(csub(csub(z7, cscale(7.0, z5)), cscale(7.0, z2)), one);
    const cplx den = cadd(csub(cadd(z7, cscale(7.0, z5)), cscale(7.0, z2)), one);
    return cdiv(num, den);
  }
  if (b == 8) {
    const double a = 0.1352;
    const cplx z6 = cmul(z4, z2);
    cplx den = cmul(z2, cadd(cscale(a, z6), one));
    den.re += eps;
    return cdiv(csub(z6, cplx{a, 0.0}), den);
  }
  cplx p = one;
  for (int k = 0; k < b; ++k) p = cmul(p, z);
  return p;
}

// pi (1 - tanh |r (a + r (b + ...))|) of the fit for B = b; `outside` for b beyond 1..9
__device__ double fit_profile(const double (*c)[5], double r, int b, double outside) {
  if (b < 1 || b > 9) return outside;
  const double* k = c[b - 1];
  const double poly = r * (k[0] + r * (k[1] + r * (k[2] + r * (k[3] + r * k[4]))));
  return kPi * (1.0 - tanh(fabs(poly)));
}

// profilefun (initial_config.py:343-442).  Beyond B = 9 the reference's
// coefficients stay 0: the massive fit is then pi (1 - tanh 0) = pi and the
// massless term of the m0 < 1 branch is never evaluated (it stays 0).
__device__ double profilefun(double r, int b, double m0) {
  double f = (m0 < 1.0) ? (1.0 - m0) * fit_profile(kFitMassless, r, b, 0.0) + m0 * fit_profile(kFitMassive, r, b, kPi)
                        : fit_profile(kFitMassive, r, b, kPi);
  if (r > 15.0) f = 0.0;
  if (r < 1.0e-6) f = kPi;
  return f;
}

__device__ double profile(const skx_lump& l, double r, int b) {
  if (l.profile == SKX_PROFILE_EXP) return kPi * exp(-r / l.profile_param);
  if (l.profile == SKX_PROFILE_GAUSS) {
    const double u = r / l.profile_param;
    return kPi * exp(-(u * u));
  }
  return profilefun(r, b, l.profile_param);
}

__device__ inline void mat3(const double* M, double& x, double& y, double& z) {
  const double a = M[0] * x + M[1] * y + M[2] * z;
  const double b = M[3] * x + M[4] * y + M[5] * z;
  const double c = M[6] * x + M[7] * y + M[8] * z;
  x = a;
  y = b;
  z = c;
}

// One lump at centred position (rx, ry, rz): (cos f, sin f n).
__device__ void eval_lump(const skx_lump& l, double rx, double ry, double rz, double q[4]) {
  rx -= l.center[0];
  ry -= l.center[1];
  rz -= l.center[2];
  if (l.kind == SKX_LUMP_HEDGEHOG) {  // initial_config.py:701-758
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    if (r < 1.0e-12) {
      q[0] = -1.0;
      q[1] = q[2] = q[3] = 0.0;
      return;
    }
    const double f = profile(l, r, 1);
    double nx = rx / r, ny = ry / r, nz = rz / r;
    mat3(l.iso_rot, nx, ny, nz);
    const double sf = sin(f);
    q[0] = cos(f);
    q[1] = sf * nx;
    q[2] = sf * ny;
    q[3] = sf * nz;
    return;
  }
  // rational map (initial_config.py:941-953, 553-588)
  const int b = l.baryon_number < 1 ? 1 : l.baryon_number;
  mat3(l.space_rot, rx, ry, rz);
  const double r = sqrt(rx * rx + ry * ry + rz * rz);
  cplx z{0.0, 0.0};
  if (r >= 1.0e-6) {
    const double c = fmin(fmax(rz / r, -1.0), 1.0);
    const double theta = acos(c), phi = atan2(ry, rx);
    const double t = tan(0.5 * theta);
    z = {cos(phi) * t, sin(phi) * t};
  }
  const double f = profile(l, r, b);
  const cplx R = rational_map(z, b);
  const double zsq = R.re * R.re + R.im * R.im;
  const double den = 1.0 + zsq;
  double nx = 2.0 * R.re / den, ny = 2.0 * R.im / den, nz = (1.0 - zsq) / den;
  mat3(l.iso_rot, nx, ny, nz);
  const double sf = sin(f);
  q[0] = cos(f);
  q[1] = sf * nx;
  q[2] = sf * ny;
  q[3] = sf * nz;
}

__device__ inline void normalise4(double q[4]) {
  const double s = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (s != 0.0)
    for (int a = 0; a < 4; ++a) q[a] /= s;
}

// counter-based N(0, 1) sample for (seed, index): splitmix64 hashes, Box-Muller
__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ inline double gauss(uint64_t seed, uint64_t idx) {
  const uint64_t h1 = mix64(seed ^ mix64(2 * idx)), h2 = mix64(seed ^ mix64(2 * idx + 1));
  const double u1 = ((h1 >> 11) + 1) * (1.0 / 9007199254740993.0);  // (0, 1]
  const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);        // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

template <typename T>
__global__ void k_init_field(T* __restrict__ F, T* __restrict__ V, InitArgs A) {
  const long long N = (long long)A.X * A.Y * A.Z;
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; s < N; s += stride) {
    const int z = (int)(s % A.Z);
    const int y = (int)((s / A.Z) % A.Y);
    const int x = (int)(s / ((long long)A.Z * A.Y));
    // centred coordinates: site i at ls * i, centre halfway between the first and last site
    const double rx = A.lsx * x - 0.5 * (A.lsx * (A.X - 1));
    const double ry = A.lsy * y - 0.5 * (A.lsy * (A.Y - 1));
    const double rz = A.lsz * z - 0.5 * (A.lsz * (A.Z - 1));
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    for (int l = 0; l < A.d.n_lumps; ++l) {
      double p[4];
      eval_lump(A.d.lumps[l], rx, ry, rz, p);
      if (l == 0) {
        for (int a = 0; a < 4; ++a) q[a] = p[a];
      } else {  // quaternion product (initial_config.py:591-627), renormalised
        const double r0 = q[0] * p[0] - q[1] * p[1] - q[2] * p[2] - q[3] * p[3];
        const double r1 = q[0] * p[1] + p[0] * q[1] + q[2] * p[3] - q[3] * p[2];
        const double r2 = q[0] * p[2] + p[0] * q[2] + q[3] * p[1] - q[1] * p[3];
        const double r3 = q[0] * p[3] + p[0] * q[3] + q[1] * p[2] - q[2] * p[1];
        q[0] = r0;
        q[1] = r1;
        q[2] = r2;
        q[3] = r3;
        normalise4(q);
      }
    }
    if (A.d.noise > 0.0)
      for (int a = 0; a < 4; ++a) q[a] += A.d.noise * gauss(A.d.noise_seed, (uint64_t)a * (uint64_t)N + (uint64_t)s);
    normalise4(q);
    const int h = A.halo;
    const bool band = A.d.vacuum_band && (x < h || x >= A.X - h || y < h || y >= A.Y - h || z < h || z >= A.Z - h);
    if (band) {
      q[0] = 1.0;
      q[1] = q[2] = q[3] = 0.0;
    }
    for (int a = 0; a < 4; ++a) F[a * N + s] = T(q[a]);
    for (int a = 4; a < A.nf; ++a) F[a * N + s] = T(0);
    for (int a = 0; a < A.nf; ++a) V[a * N + s] = T(0);
  }
}

// component `comp` = scale * dens (omega = -c CS from the baryon density CS * vol)
template <typename T>
__global__ void k_set_component(T* __restrict__ F, const double* __restrict__ dens, long long N, double scale) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; s < N; s += stride) F[s] = T(scale * dens[s]);
}

int grid_for(long long n) { return (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192); }

}  // namespace

hipError_t launch_init_field_f64(double* F, double* V, const InitArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_init_field<double>, dim3(grid_for((long long)a.X * a.Y * a.Z)), dim3(256), 0, st, F, V, a);
  return hipGetLastError();
}
hipError_t launch_init_field_f32(float* F, float* V, const InitArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_init_field<float>, dim3(grid_for((long long)a.X * a.Y * a.Z)), dim3(256), 0, st, F, V, a);
  return hipGetLastError();
}
hipError_t launch_set_component_f64(double* Fc, const double* dens, long long N, double scale, hipStream_t st) {
  hipLaunchKernelGGL(k_set_component<double>, dim3(grid_for(N)), dim3(256), 0, st, Fc, dens, N, scale);
  return hipGetLastError();
}
hipError_t launch_set_component_f32(float* Fc, const double* dens, long long N, double scale, hipStream_t st) {
  hipLaunchKernelGGL(k_set_component<float>, dim3(grid_for(N)), dim3(256), 0, st, Fc, dens, N, scale);
  return hipGetLastError();
}

}  // namespace skx
