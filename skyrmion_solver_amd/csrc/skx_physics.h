// Per-site physics of the engine: point gradients, energy and baryon
// densities for the pion (nuclear), omega, Coulomb and rho Skyrme theories.
//
// Every function is __host__ __device__ and works on one lattice site whose
// derivatives are already in registers:
//   phi[a]      field value of component a (the stage field "Temp")
//   F[i][a]     first derivative along axis i (0 = x, 1 = y, 2 = z)
//   S[k][a]     second derivatives, k = 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz
// The formulas are the Euler-Lagrange gradients the reference codes per site
// (file:line citations at each function); they are regrouped so that shared
// contractions are formed once, which changes rounding only (the engine's
// FP64 form matches the reference arithmetic to ~1e-14 relative).
#pragma once

#include <hip/hip_runtime.h>

#define SKX_HD __host__ __device__ __forceinline__

namespace skx {

constexpr double kPi = 3.14159265358979323846;

enum Theory : int { PION = 0, OMEGA = 1, COULOMB = 2, RHO = 3, BERGER = 4, LIGHTLY = 5 };

// Index of S for the pair (i, j) in the packed 6-entry symmetric layout.
SKX_HD constexpr int sidx(int i, int j) {
  return i == j ? i : (i + j == 1 ? 3 : (i + j == 2 ? 4 : 5));
}

// Theory constants handed to the kernels (filled from the packed p_f array).
struct PhysParams {
  double dt;     // p_f[7]
  double vol;    // p_f[6]
  double m2;     // squared mass coefficient (pion: mpi^2, omega/Coulomb: mass^2, rho: mpi^2)
  double k2;     // pion: kappa^2
  double coup;   // omega: c_omega; Coulomb: kappa; rho: mrho^2
  double calpha; // rho: c_alpha (p_f[10])
};

// 4-D "cross product": X_mu = det[e_mu, A, B, C] (columns e_mu, A, B, C).
// Equals the cofactor the reference forms with _det4_cols(e_mu, A, B, C)
// (theories/omega_skyrme/kernels.py:717-722).
template <typename T>
SKX_HD void cross4(const T A[4], const T B[4], const T C[4], T X[4]) {
  const T m01 = B[0] * C[1] - B[1] * C[0];
  const T m02 = B[0] * C[2] - B[2] * C[0];
  const T m03 = B[0] * C[3] - B[3] * C[0];
  const T m12 = B[1] * C[2] - B[2] * C[1];
  const T m13 = B[1] * C[3] - B[3] * C[1];
  const T m23 = B[2] * C[3] - B[3] * C[2];
  X[0] = A[1] * m23 - A[2] * m13 + A[3] * m12;
  X[1] = -(A[0] * m23 - A[2] * m03 + A[3] * m02);
  X[2] = A[0] * m13 - A[1] * m03 + A[3] * m01;
  X[3] = -(A[0] * m12 - A[1] * m02 + A[2] * m01);
}

// det[phi, F_x, F_y, F_z] over components 0..3
// (theories/nuclear_skyrme/kernels.py:441-446, _det4_cols :155-194).
template <typename T>
SKX_HD T det_phi_F(const T* phi, const T (*F)[4]) {
  T X[4];
  cross4<T>(F[0], F[1], F[2], X);
  return phi[0] * X[0] + phi[1] * X[1] + phi[2] * X[2] + phi[3] * X[3];
}

// Skyrme quartic + Dirichlet gradient on the 4 sigma-model components
// (theories/nuclear_skyrme/kernels.py:1109-1140; coulomb :651-680 with k2 = 1):
//   g_a = -sum_i S_ii,a - k2 * sum_{i != j} [ 2 (F_j.S_ij) F_ia + n_j S_ii,a
//                                             - (S_ii.F_j + F_i.S_ij) F_ja - (F_i.F_j) S_ij,a ]
// regrouped as
//   g_a = -sum_i (1 + k2 (n - n_i)) S_ii,a - k2 sum_k C_k F_ka + 2 k2 sum_{i<j} d_ij S_ij,a
// with n_i = F_i.F_i, n = sum n_i, d_ij = F_i.F_j and
//   C_k = sum_{i != k} (F_i.S_ik - S_ii.F_k).
template <typename T, int LD>
SKX_HD void skyrme_quartic_gradient_q(const T (*F)[LD], const T (*S)[LD], T q, T k2, T g[4]) {
  T n[3], d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) n[i] = F[i][0] * F[i][0] + F[i][1] * F[i][1] + F[i][2] * F[i][2] + F[i][3] * F[i][3];
  d[0] = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2] + F[0][3] * F[1][3];  // xy
  d[1] = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2] + F[0][3] * F[2][3];  // xz
  d[2] = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2] + F[1][3] * F[2][3];  // yz
  const T ntot = n[0] + n[1] + n[2];
  T C[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T c = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i == k) continue;
      const T* Sik = S[sidx(i, k)];
      const T* Sii = S[i];
      const T P = F[i][0] * Sik[0] + F[i][1] * Sik[1] + F[i][2] * Sik[2] + F[i][3] * Sik[3];
      const T Q = Sii[0] * F[k][0] + Sii[1] * F[k][1] + Sii[2] * F[k][2] + Sii[3] * F[k][3];
      c += P - Q;
    }
    C[k] = c;
  }
  const T w0 = q + k2 * (ntot - n[0]);
  const T w1 = q + k2 * (ntot - n[1]);
  const T w2 = q + k2 * (ntot - n[2]);
  const T e01 = T(2) * k2 * d[0], e02 = T(2) * k2 * d[1], e12 = T(2) * k2 * d[2];
  const T c0 = k2 * C[0], c1 = k2 * C[1], c2 = k2 * C[2];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    g[a] = -(w0 * S[0][a] + w1 * S[1][a] + w2 * S[2][a]) - (c0 * F[0][a] + c1 * F[1][a] + c2 * F[2][a]) +
           (e01 * S[3][a] + e02 * S[4][a] + e12 * S[5][a]);
  }
}

// The same with the sigma-model Laplacian at unit weight (nuclear, Coulomb, Berger).
template <typename T, int LD>
SKX_HD void skyrme_quartic_gradient(const T (*F)[LD], const T (*S)[LD], T k2, T g[4]) {
  skyrme_quartic_gradient_q<T, LD>(F, S, T(1), k2, g);
}

// g <- g - (g.phi) phi over the first NSK components
// (theories/nuclear_skyrme/kernels.py:267-272).
template <typename T, int NSK>
SKX_HD void project_tangent(T* g, const T* phi) {
  T lm = T(0);
#pragma unroll
  for (int a = 0; a < NSK; ++a) lm += g[a] * phi[a];
#pragma unroll
  for (int a = 0; a < NSK; ++a) g[a] -= lm * phi[a];
}

// Sum over i of d_i w * 2 X(phi, F_{i+1}, F_{i+2}) = sum_{ijk} eps_ijk d_i w det[e_mu, phi, F_j, F_k].
template <typename T, int LD>
SKX_HD void eps_coupling(const T* phi, const T (*F)[LD], const T dw[3], T out[4]) {
  T X[4];
  out[0] = out[1] = out[2] = out[3] = T(0);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    T Fj[4] = {F[j][0], F[j][1], F[j][2], F[j][3]};
    T Fk[4] = {F[k][0], F[k][1], F[k][2], F[k][3]};
    T ph[4] = {phi[0], phi[1], phi[2], phi[3]};
    cross4<T>(ph, Fj, Fk, X);
    const T s = T(2) * dw[i];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) out[mu] += s * X[mu];
  }
}

template <typename T, int LD>
SKX_HD T chern_simons(const T* phi, const T (*F)[LD]) {
  T Fx[4] = {F[0][0], F[0][1], F[0][2], F[0][3]};
  T Fy[4] = {F[1][0], F[1][1], F[1][2], F[1][3]};
  T Fz[4] = {F[2][0], F[2][1], F[2][2], F[2][3]};
  T X[4];
  cross4<T>(Fx, Fy, Fz, X);
  const T det = phi[0] * X[0] + phi[1] * X[1] + phi[2] * X[2] + phi[3] * X[3];
  return -det * T(1.0 / (2.0 * kPi * kPi));
}

// ---------------------------------------------------------------- gradients
// Each returns the projected gradient g (what the reference stores in
// EnergyGradient); the caller forms l = -dt g.

// theories/nuclear_skyrme/kernels.py:1077-1150
template <typename T>
SKX_HD void gradient_pion(const T* phi, const T (*F)[4], const T (*S)[4], const PhysParams& P, T g[4]) {
  skyrme_quartic_gradient<T, 4>(F, S, T(P.k2), g);
  g[0] -= T(P.m2);
  project_tangent<T, 4>(g, phi);
}

// ---------------------------------------------------------------- Berger-Skyrme
// omega(a, b) = a0 b3 + a1 b2 - a2 b1 - a3 b0 (_omega_contract,
// theories/berger_skyrme/kernels.py:81-111); Omega v = (v3, v2, -v1, -v0) is
// the vector with a . (Omega v) = omega(a, v) (:659-677).
template <typename T>
SKX_HD T omega4(const T* a, const T* b) {
  return a[0] * b[3] + a[1] * b[2] - a[2] * b[1] - a[3] * b[0];
}

// theories/berger_skyrme/kernels.py:554-737: the nuclear gradient plus, with
// beta = alpha^2 - 1 (P.coup) and u_i = omega(F_i, phi), W_ij = omega(F_i, F_j),
//   - beta [2 sum_i u_i Omega F_i + (sum_i omega(S_ii, phi)) Omega phi]         (:642-682)
//   - kappa^2 beta sum_{i<j} [d_i W_ij Omega F_j - d_j W_ij Omega F_i]           (:705-724)
// with d_k W_ij = omega(S_ik, F_j) + omega(F_i, S_jk), gathered here per Omega F_i.
template <typename T>
SKX_HD void gradient_berger(const T* phi, const T (*F)[4], const T (*S)[4], const PhysParams& P, T g[4]) {
  skyrme_quartic_gradient<T, 4>(F, S, T(P.k2), g);
  const T beta = T(P.coup), kb = T(P.k2) * beta;
  const T du = omega4<T>(S[0], phi) + omega4<T>(S[1], phi) + omega4<T>(S[2], phi);
  const T dWxy_x = omega4<T>(S[0], F[1]) + omega4<T>(F[0], S[3]);
  const T dWxy_y = omega4<T>(S[3], F[1]) + omega4<T>(F[0], S[1]);
  const T dWxz_x = omega4<T>(S[0], F[2]) + omega4<T>(F[0], S[4]);
  const T dWxz_z = omega4<T>(S[4], F[2]) + omega4<T>(F[0], S[2]);
  const T dWyz_y = omega4<T>(S[1], F[2]) + omega4<T>(F[1], S[5]);
  const T dWyz_z = omega4<T>(S[5], F[2]) + omega4<T>(F[1], S[2]);
  const T cx = T(2) * beta * omega4<T>(F[0], phi) - kb * (dWxy_y + dWxz_z);
  const T cy = T(2) * beta * omega4<T>(F[1], phi) + kb * (dWxy_x - dWyz_z);
  const T cz = T(2) * beta * omega4<T>(F[2], phi) + kb * (dWxz_x + dWyz_y);
  const T cp = beta * du;
  // Omega v = (v3, v2, -v1, -v0)
  g[0] -= cx * F[0][3] + cy * F[1][3] + cz * F[2][3] + cp * phi[3];
  g[1] -= cx * F[0][2] + cy * F[1][2] + cz * F[2][2] + cp * phi[2];
  g[2] += cx * F[0][1] + cy * F[1][1] + cz * F[2][1] + cp * phi[1];
  g[3] += cx * F[0][0] + cy * F[1][0] + cz * F[2][0] + cp * phi[0];
  g[0] -= T(P.m2);
  project_tangent<T, 4>(g, phi);
}

// theories/lightly_bound_skyrme/kernels.py:567-683: sigma-model term at weight
// 2 (1 - alpha), quartic term at unit coupling, then
// - 2 (1 - alpha) mpi^2 - 4 alpha (1 - sigma)^3 on sigma; alpha = P.coup.
template <typename T>
SKX_HD void gradient_lightly(const T* phi, const T (*F)[4], const T (*S)[4], const PhysParams& P, T g[4]) {
  const T al = T(P.coup), one_al = T(1) - al;
  skyrme_quartic_gradient_q<T, 4>(F, S, T(2) * one_al, T(1), g);
  const T om = T(1) - phi[0];
  g[0] -= T(2) * one_al * T(P.m2) + T(4) * al * (om * om * om);
  project_tangent<T, 4>(g, phi);
}

// theories/omega_skyrme/kernels.py:626-743
template <typename T>
SKX_HD void gradient_omega(const T* phi, const T (*F)[5], const T (*S)[5], const PhysParams& P, T g[5]) {
  const T c = T(P.coup);
  const T dw[3] = {F[0][4], F[1][4], F[2][4]};
  T cp[4];
  eps_coupling<T, 5>(phi, F, dw, cp);
  const T cf = c * T(1.0 / (4.0 * kPi * kPi));
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) g[mu] = T(-0.25) * (S[0][mu] + S[1][mu] + S[2][mu]) + cf * cp[mu];
  g[0] -= T(0.25) * T(P.m2);
  project_tangent<T, 4>(g, phi);
  g[4] = phi[4] - (S[0][4] + S[1][4] + S[2][4]) + c * chern_simons<T, 5>(phi, F);
}

// theories/coulomb_skyrme/kernels.py:617-767 (the dead store 2g at :684-685 is
// overwritten by the reference and has no effect)
template <typename T>
SKX_HD void gradient_coulomb(const T* phi, const T (*F)[5], const T (*S)[5], const PhysParams& P, T g[5]) {
  const T kap = T(P.coup);
  skyrme_quartic_gradient<T, 5>(F, S, T(1), g);
  g[0] -= T(P.m2);
  const T dV[3] = {F[0][4], F[1][4], F[2][4]};
  T cp[4];
  eps_coupling<T, 5>(phi, F, dV, cp);
  const T cf = kap * T(1.0 / (4.0 * kPi * kPi));
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) g[mu] -= cf * cp[mu];
  project_tangent<T, 4>(g, phi);
  g[4] = -(kap * (S[0][4] + S[1][4] + S[2][4]) + kap * chern_simons<T, 5>(phi, F));
}

// ---------------------------------------------------------------- rho
// rho^a_i lives at component 4 + 3 i + a (theories/rho_skyrme/kernels.py:84-88).
SKX_HD constexpr int rho_c(int i, int a) { return 4 + 3 * i + a; }

// eps_{abc} for a, b, c in 0..2
SKX_HD constexpr double eps3(int a, int b, int c) {
  return (a == b || b == c || a == c) ? 0.0
         : ((a == 0 && b == 1 && c == 2) || (a == 1 && b == 2 && c == 0) || (a == 2 && b == 0 && c == 1)) ? 1.0
                                                                                                             : -1.0;
}

// d_l R^a_{jk} = S_jl,rho(k,a) - S_kl,rho(j,a)   (theories/rho_skyrme/kernels.py:180-184)
template <typename T, int LD>
SKX_HD T rho_dR(const T (*S)[LD], int a, int j, int k, int l) {
  return S[sidx(j, l)][rho_c(k, a)] - S[sidx(k, l)][rho_c(j, a)];
}

// d_l Omega^a_{jk}  (theories/rho_skyrme/kernels.py:200-212)
template <typename T, int LD>
SKX_HD T rho_dOmega(const T (*F)[4], const T (*S)[LD], int a, int j, int k, int l) {
  T t = S[sidx(j, l)][0] * F[k][a + 1] + F[j][0] * S[sidx(k, l)][a + 1] - S[sidx(j, l)][a + 1] * F[k][0] -
        F[j][a + 1] * S[sidx(k, l)][0];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double e = eps3(a, b, c);
      if (e != 0.0) t += T(e) * (S[sidx(j, l)][b + 1] * F[k][c + 1] + F[j][b + 1] * S[sidx(k, l)][c + 1]);
    }
  return t;
}

// theories/rho_skyrme/kernels.py:607-682.  F holds the 4 pion components
// only (the rho gradient reads no rho first derivatives); S holds all 13.
template <typename T>
SKX_HD void gradient_rho(const T* phi, const T (*F)[4], const T (*S)[13], const PhysParams& P, T g[13]) {
  const T mpi2 = T(P.m2), mrho2 = T(P.coup), ca = T(P.calpha);
  // pion part, c2 = 1, c4 = 1/4 written as the reference's loop
  T nn[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) nn[j] = F[j][0] * F[j][0] + F[j][1] * F[j][1] + F[j][2] * F[j][2] + F[j][3] * F[j][3];
  const T ntot = nn[0] + nn[1] + nn[2];
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    T acc = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const T Sii = S[i][mu];
      acc += T(2) * Sii;
      // sum_j sum_nu [S_ii,mu F_jnu^2 + F_imu (S_ij,nu F_jnu - S_jj,nu F_inu) - S_ij,mu F_inu F_jnu]
      T t = Sii * ntot;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T* Sij = S[sidx(i, j)];
        const T* Sjj = S[j];
        const T a1 = Sij[0] * F[j][0] + Sij[1] * F[j][1] + Sij[2] * F[j][2] + Sij[3] * F[j][3];
        const T a2 = Sjj[0] * F[i][0] + Sjj[1] * F[i][1] + Sjj[2] * F[i][2] + Sjj[3] * F[i][3];
        const T dij = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2] + F[i][3] * F[j][3];
        t += F[i][mu] * (a1 - a2) - Sij[mu] * dij;
      }
      acc += T(2) * t;
    }
    g[mu] = -acc;
  }
  g[0] -= T(2) * mpi2;
  // rho back-reaction on the pion components (kernels.py:656-664)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T dR = rho_dR<T, 13>(S, a, i, j, i);
        g[0] += T(16) * ca * F[j][a + 1] * dR;
        g[a + 1] -= T(16) * ca * F[j][0] * dR;
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double e = eps3(b, a, c);
            if (e != 0.0) g[a + 1] += T(16) * ca * T(e) * F[j][c + 1] * rho_dR<T, 13>(S, b, i, j, i);
          }
      }
  // rho components (kernels.py:667-674)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T val = T(8) * mrho2 * phi[rho_c(i, a)];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        val += T(8) * rho_dR<T, 13>(S, a, i, j, j);
        val -= T(16) * ca * rho_dOmega<T, 13>(F, S, a, i, j, j);
      }
      g[rho_c(i, a)] = val;
    }
  project_tangent<T, 4>(g, phi);
}

// The rho gradient of theories/rho_skyrme/kernels.py:607-682 with the rho
// second derivatives entering through
//   D[a][j] = sum_i d_i R^a_ij = sum_{i != j} (S_ii,rho(j,a) - S_ij,rho(i,a)),
// which is how they appear there: the back-reaction sums _d_R_a_ijk(S, a, i, j, i)
// over i (kernels.py:655-664) and the rho equation sums
// _d_R_a_ijk(S, a, i, j, j) = -D[a][i] over j (kernels.py:667-674).  F and S
// hold the 4 pion components only.
template <typename T>
SKX_HD void gradient_rho_split(const T* phi, const T (*F)[4], const T (*S)[4], const T (*D)[3], const PhysParams& P,
                               T g[13]) {
  const T mpi2 = T(P.m2), mrho2 = T(P.coup), ca = T(P.calpha);
  T nn[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) nn[j] = F[j][0] * F[j][0] + F[j][1] * F[j][1] + F[j][2] * F[j][2] + F[j][3] * F[j][3];
  const T ntot = nn[0] + nn[1] + nn[2];
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    T acc = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const T Sii = S[i][mu];
      acc += T(2) * Sii;
      T t = Sii * ntot;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T* Sij = S[sidx(i, j)];
        const T* Sjj = S[j];
        const T a1 = Sij[0] * F[j][0] + Sij[1] * F[j][1] + Sij[2] * F[j][2] + Sij[3] * F[j][3];
        const T a2 = Sjj[0] * F[i][0] + Sjj[1] * F[i][1] + Sjj[2] * F[i][2] + Sjj[3] * F[i][3];
        const T dij = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2] + F[i][3] * F[j][3];
        t += F[i][mu] * (a1 - a2) - Sij[mu] * dij;
      }
      acc += T(2) * t;
    }
    g[mu] = -acc;
  }
  g[0] -= T(2) * mpi2;
  const T c16 = T(16) * ca;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const T d = D[a][j];
      g[0] += c16 * F[j][a + 1] * d;
      g[a + 1] -= c16 * F[j][0] * d;
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double e = eps3(b, a, c);
          if (e != 0.0) g[a + 1] += c16 * T(e) * F[j][c + 1] * D[b][j];
        }
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T om = T(0);
#pragma unroll
      for (int j = 0; j < 3; ++j) om += rho_dOmega<T, 4>(F, S, a, i, j, j);
      g[rho_c(i, a)] = T(8) * mrho2 * phi[rho_c(i, a)] - T(8) * D[a][i] - c16 * om;
    }
  project_tangent<T, 4>(g, phi);
}

// ---------------------------------------------------------------- energy densities
// All return e * vol in double; derivatives are promoted to double first.

// theories/nuclear_skyrme/kernels.py:312-347
template <typename T, int LD>
SKX_HD double energy_pion(const T* phi, const T (*F)[LD], const PhysParams& P) {
  double f[3][4];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a) f[i][a] = double(F[i][a]);
  const double nx = f[0][0] * f[0][0] + f[0][1] * f[0][1] + f[0][2] * f[0][2] + f[0][3] * f[0][3];
  const double ny = f[1][0] * f[1][0] + f[1][1] * f[1][1] + f[1][2] * f[1][2] + f[1][3] * f[1][3];
  const double nz = f[2][0] * f[2][0] + f[2][1] * f[2][1] + f[2][2] * f[2][2] + f[2][3] * f[2][3];
  const double dxy = f[0][0] * f[1][0] + f[0][1] * f[1][1] + f[0][2] * f[1][2] + f[0][3] * f[1][3];
  const double dxz = f[0][0] * f[2][0] + f[0][1] * f[2][1] + f[0][2] * f[2][2] + f[0][3] * f[2][3];
  const double dyz = f[1][0] * f[2][0] + f[1][1] * f[2][1] + f[1][2] * f[2][2] + f[1][3] * f[2][3];
  double e = 0.5 * (nx + ny + nz);
  e += 0.5 * P.k2 * ((nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz));
  e += P.m2 * (1.0 - double(phi[0]));
  return e * P.vol;
}

// theories/berger_skyrme/kernels.py:251-337: the nuclear density plus
// beta/2 sum_i u_i^2 + kappa^2 beta/2 sum_{i<j} W_ij^2.
template <typename T, int LD>
SKX_HD double energy_berger(const T* phi, const T (*F)[LD], const PhysParams& P) {
  double f[3][4], s[4];
  for (int a = 0; a < 4; ++a) {
    s[a] = double(phi[a]);
    for (int i = 0; i < 3; ++i) f[i][a] = double(F[i][a]);
  }
  const double nx = f[0][0] * f[0][0] + f[0][1] * f[0][1] + f[0][2] * f[0][2] + f[0][3] * f[0][3];
  const double ny = f[1][0] * f[1][0] + f[1][1] * f[1][1] + f[1][2] * f[1][2] + f[1][3] * f[1][3];
  const double nz = f[2][0] * f[2][0] + f[2][1] * f[2][1] + f[2][2] * f[2][2] + f[2][3] * f[2][3];
  const double dxy = f[0][0] * f[1][0] + f[0][1] * f[1][1] + f[0][2] * f[1][2] + f[0][3] * f[1][3];
  const double dxz = f[0][0] * f[2][0] + f[0][1] * f[2][1] + f[0][2] * f[2][2] + f[0][3] * f[2][3];
  const double dyz = f[1][0] * f[2][0] + f[1][1] * f[2][1] + f[1][2] * f[2][2] + f[1][3] * f[2][3];
  const double ux = omega4<double>(f[0], s), uy = omega4<double>(f[1], s), uz = omega4<double>(f[2], s);
  const double wxy = omega4<double>(f[0], f[1]), wxz = omega4<double>(f[0], f[2]), wyz = omega4<double>(f[1], f[2]);
  const double beta = P.coup;
  double e = 0.5 * (nx + ny + nz);
  e += 0.5 * beta * (ux * ux + uy * uy + uz * uz);
  e += 0.5 * P.k2 * ((nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz));
  e += 0.5 * P.k2 * beta * (wxy * wxy + wxz * wxz + wyz * wyz);
  e += P.m2 * (1.0 - s[0]);
  return e * P.vol;
}

// theories/lightly_bound_skyrme/kernels.py:275-345:
// (1 - alpha)|d phi|^2 + quartic / 2 + 2 (1 - alpha) mpi^2 (1 - sigma) + alpha (1 - sigma)^4.
template <typename T, int LD>
SKX_HD double energy_lightly(const T* phi, const T (*F)[LD], const PhysParams& P) {
  double f[3][4];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a) f[i][a] = double(F[i][a]);
  const double nx = f[0][0] * f[0][0] + f[0][1] * f[0][1] + f[0][2] * f[0][2] + f[0][3] * f[0][3];
  const double ny = f[1][0] * f[1][0] + f[1][1] * f[1][1] + f[1][2] * f[1][2] + f[1][3] * f[1][3];
  const double nz = f[2][0] * f[2][0] + f[2][1] * f[2][1] + f[2][2] * f[2][2] + f[2][3] * f[2][3];
  const double dxy = f[0][0] * f[1][0] + f[0][1] * f[1][1] + f[0][2] * f[1][2] + f[0][3] * f[1][3];
  const double dxz = f[0][0] * f[2][0] + f[0][1] * f[2][1] + f[0][2] * f[2][2] + f[0][3] * f[2][3];
  const double dyz = f[1][0] * f[2][0] + f[1][1] * f[2][1] + f[1][2] * f[2][2] + f[1][3] * f[2][3];
  const double al = P.coup, om = 1.0 - double(phi[0]);
  double e = (1.0 - al) * (nx + ny + nz);
  e += 0.5 * ((nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz));
  e += (1.0 - al) * 2.0 * P.m2 * om;
  e += al * om * om * om * om;
  return e * P.vol;
}

// theories/omega_skyrme/kernels.py:266-302 (coupling term commented out there)
template <typename T, int LD>
SKX_HD double energy_omega(const T* phi, const T (*F)[LD], const PhysParams& P) {
  double n = 0.0, w2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    for (int a = 0; a < 4; ++a) n += double(F[i][a]) * double(F[i][a]);
    w2 += double(F[i][4]) * double(F[i][4]);
  }
  const double w = double(phi[4]);
  return (0.125 * n + P.m2 * 0.25 * (1.0 - double(phi[0])) + 0.5 * w * w + 0.5 * w2) * P.vol;
}

// theories/coulomb_skyrme/kernels.py:266-306 (coupling term commented out there)
template <typename T, int LD>
SKX_HD double energy_coulomb(const T* phi, const T (*F)[LD], const PhysParams& P) {
  double f[3][4];
  double v2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    for (int a = 0; a < 4; ++a) f[i][a] = double(F[i][a]);
    v2 += double(F[i][4]) * double(F[i][4]);
  }
  const double nx = f[0][0] * f[0][0] + f[0][1] * f[0][1] + f[0][2] * f[0][2] + f[0][3] * f[0][3];
  const double ny = f[1][0] * f[1][0] + f[1][1] * f[1][1] + f[1][2] * f[1][2] + f[1][3] * f[1][3];
  const double nz = f[2][0] * f[2][0] + f[2][1] * f[2][1] + f[2][2] * f[2][2] + f[2][3] * f[2][3];
  const double dxy = f[0][0] * f[1][0] + f[0][1] * f[1][1] + f[0][2] * f[1][2] + f[0][3] * f[1][3];
  const double dxz = f[0][0] * f[2][0] + f[0][1] * f[2][1] + f[0][2] * f[2][2] + f[0][3] * f[2][3];
  const double dyz = f[1][0] * f[2][0] + f[1][1] * f[2][1] + f[1][2] * f[2][2] + f[1][3] * f[2][3];
  double e = nx + ny + nz;
  e += (nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz);
  e += 2.0 * P.m2 * (1.0 - double(phi[0]));
  e += 0.5 * P.coup * v2;
  return e * P.vol;
}

// theories/rho_skyrme/kernels.py:302-347
template <typename T>
SKX_HD double energy_rho(const T* phi, const T (*F)[13], const PhysParams& P) {
  const double mrho2 = P.coup, ca = P.calpha;
  double f[3][4];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a) f[i][a] = double(F[i][a]);
  double n[3], d[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) d[i][j] = f[i][0] * f[j][0] + f[i][1] * f[j][1] + f[i][2] * f[j][2] + f[i][3] * f[j][3];
  for (int i = 0; i < 3; ++i) n[i] = d[i][i];
  double e = n[0] + n[1] + n[2];
  // 1/2 sum_{ij} [ n_i n_j - d_ij^2 ]
  double q = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) q += n[i] * n[j] - d[i][j] * d[i][j];
  e += 0.5 * q;
  e += 2.0 * P.m2 * (1.0 - double(phi[0]));
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i) {
      const double r = double(phi[rho_c(i, a)]);
      e += 4.0 * mrho2 * r * r;
      for (int j = 0; j < 3; ++j) {
        const double R = double(F[i][rho_c(j, a)]) - double(F[j][rho_c(i, a)]);
        double Om = f[i][0] * f[j][a + 1] - f[i][a + 1] * f[j][0];
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) {
            const double ee = eps3(a, b, c);
            if (ee != 0.0) Om += ee * f[i][b + 1] * f[j][c + 1];
          }
        e += 2.0 * R * R;
        e -= 8.0 * ca * R * Om;
      }
    }
  return e * P.vol;
}

// theories/rho_skyrme/kernels.py:302-347 with the rho first derivatives
// entering through R[a][p] = R^a_ij for the pairs p = (x,y), (x,z), (y,z):
// R and Omega are antisymmetric in (i, j), so the ordered-pair sum of the
// reference is twice the sum over these pairs.  F holds the pion components.
template <typename T>
SKX_HD double energy_rho_split(const T* phi, const T (*F)[4], const T (*R)[3], const PhysParams& P) {
  const double mrho2 = P.coup, ca = P.calpha;
  double f[3][4];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a) f[i][a] = double(F[i][a]);
  double n[3], d[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) d[i][j] = f[i][0] * f[j][0] + f[i][1] * f[j][1] + f[i][2] * f[j][2] + f[i][3] * f[j][3];
  for (int i = 0; i < 3; ++i) n[i] = d[i][i];
  double e = n[0] + n[1] + n[2];
  double q = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) q += n[i] * n[j] - d[i][j] * d[i][j];
  e += 0.5 * q;
  e += 2.0 * P.m2 * (1.0 - double(phi[0]));
  constexpr int PI[3] = {0, 0, 1}, PJ[3] = {1, 2, 2};
  for (int a = 0; a < 3; ++a) {
    for (int i = 0; i < 3; ++i) {
      const double r = double(phi[rho_c(i, a)]);
      e += 4.0 * mrho2 * r * r;
    }
    for (int p = 0; p < 3; ++p) {
      const int i = PI[p], j = PJ[p];
      const double Rij = double(R[a][p]);
      double Om = f[i][0] * f[j][a + 1] - f[i][a + 1] * f[j][0];
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          const double ee = eps3(a, b, c);
          if (ee != 0.0) Om += ee * f[i][b + 1] * f[j][c + 1];
        }
      e += 2.0 * (2.0 * Rij * Rij - 8.0 * ca * Rij * Om);
    }
  }
  return e * P.vol;
}

// Baryon density times vol.  nuclear / rho: -det[phi, Fx, Fy, Fz] vol / (2 pi^2)
// (theories/nuclear_skyrme/kernels.py:441-447); omega / Coulomb: vol * CS
// (theories/omega_skyrme/kernels.py:374-375) -- the same number.
template <typename T, int LD>
SKX_HD double baryon_density(const T* phi, const T (*F)[LD], const PhysParams& P) {
  double ph[4], fx[4], fy[4], fz[4], X[4];
  for (int a = 0; a < 4; ++a) {
    ph[a] = double(phi[a]);
    fx[a] = double(F[0][a]);
    fy[a] = double(F[1][a]);
    fz[a] = double(F[2][a]);
  }
  cross4<double>(fx, fy, fz, X);
  const double det = ph[0] * X[0] + ph[1] * X[1] + ph[2] * X[2] + ph[3] * X[3];
  return -det * (P.vol / (2.0 * kPi * kPi));
}

}  // namespace skx
