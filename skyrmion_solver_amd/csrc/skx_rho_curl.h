// Curl pass of the rho theory.
//
// The rho second derivatives enter the gradient of
// theories/rho_skyrme/kernels.py:607-682 only through
//   D^a_j = sum_{i != j} (d_i d_i rho^a_j - d_i d_j rho^a_i)      (-curl curl rho^a)
// and the rho first derivatives enter the energy (kernels.py:302-347) only
// through R^a_ij = d_i rho^a_j - d_j rho^a_i.  For one isospin a the three
// components rho^a_x, rho^a_y, rho^a_z (components 4 + 3 i + a,
// kernels.py:84-88) are marched along x exactly like the stage kernel marches
// its components, and the pass writes per site
//   aux[(3 a + j) N]      D^a_j                      (stages 1-4)
//   aux[(9 + 3 a + p) N]  R^a_ij, p = xy, xz, yz     (stage 0; stage 1 with the energy fused)
// which the stage kernel of the same stage then reads at each site.  Blocks
// cover (isospin, tile): three times the stage kernel's grid.
#pragma once

#include "skx_device.h"

namespace skx {

template <typename T, int STAGE, int TZ, int TY>
__device__ __forceinline__ void rho_curl_body(const StageArgs<T>& A, int blk, int nblk) {
  constexpr int NC = 3;  // rho^a_x, rho^a_y, rho^a_z
  constexpr int NT = TZ * TY;
  constexpr int TW = TZ + 4;
  constexpr int TR = TY + 4;
  constexpr int TE = TW * TR;
  constexpr int LPT = (TE + NT - 1) / NT;
  constexpr int NA = (STAGE <= 1) ? 1 : (STAGE == 2 ? 2 : 3);
  constexpr bool D2 = STAGE >= 1;
  constexpr int FSLOTS = 4;
  constexpr int ZSLOTS = 3;
  static_assert(NT % 64 == 0, "block must be whole wavefronts");
  static_assert(TY >= 4, "halo rows of the z-difference ring are filled by the first four rows of threads");

  __shared__ T fring[FSLOTS][NC][TE];
  // raw z-differences of rho^a_y and rho^a_z for their yz mixed derivatives
  __shared__ T zring[D2 ? ZSLOTS : 1][D2 ? 2 : 1][D2 ? TR * TZ : 1];

  if ((A.flags & CHECK_DONE) && A.st->done) return;
  const bool want_R = (STAGE == 0) || (STAGE == 1 && (A.flags & FUSE_ENERGY));

  const int tid = threadIdx.x;
  const int tz = tid % TZ, ty = tid / TZ;
  int lin = xcd_tile_order(blk, nblk);
  const int per_iso = A.nbz * A.nby * A.nbx;
  const int iso = lin / per_iso;
  lin -= iso * per_iso;
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
  const size_t N = (size_t)A.N;
  const size_t YZ = (size_t)Y * Z;

  const bool gy1 = (y >= halo) && (y < Y - halo), gz1 = (z >= halo) && (z < Z - halo);
  const bool gy2 = (y > halo) && (y < Y - halo), gz2 = (z > halo) && (z < Z - halo);

  const T dt = T(A.P.dt);
  T keep = T(1);
  if constexpr (STAGE >= 2) keep = T(A.st->keep);
  AxisCoef<T> cx, cy, cz;
  cx.set(A.ilsx);
  cy.set(A.ilsy);
  cz.set(A.ilsz);
  const T y2a = gy1 ? cy.d2a : T(0), y2b = gy1 ? cy.d2b : T(0), y2c = gy1 ? cy.d2c : T(0);
  const T z2a = gz1 ? cz.d2a : T(0), z2b = gz1 ? cz.d2b : T(0), z2c = gz1 ? cz.d2c : T(0);
  const T mxy_a = gy2 ? cx.d1a : T(0), mxy_b = gy2 ? cx.d1b : T(0);
  const T mxz_a = gz2 ? cx.d1a : T(0), mxz_b = gz2 ? cx.d1b : T(0);
  const T myz_a = (gy2 && gz2) ? cy.d1a : T(0), myz_b = (gy2 && gz2) ? cy.d1b : T(0);

  uint32_t eoff[LPT];
#pragma unroll
  for (int j = 0; j < LPT; ++j) {
    const int e = tid + j * NT;
    const int r = e / TW, cc = e - r * TW;
    const int gy = y0 - 2 + r, gz = z0 - 2 + cc;
    const bool valid = (e < TE) && gy >= 0 && gy < Y && gz >= 0 && gz < Z;
    eoff[j] = valid ? (uint32_t)(gy * Z + gz) * (uint32_t)sizeof(T) : kOOB;
  }
  const uint32_t own_b = own ? (uint32_t)(y * Z + z) * (uint32_t)sizeof(T) : kOOB;
  const uint32_t plane_bytes = (uint32_t)(YZ * sizeof(T));
  const int hrow = (ty < 2) ? ty : (TY + ty);

  const size_t comp0 = (size_t)(4 + iso) * N;  // rho^a_i = component 4 + 3 i + a
  const T* src[3] = {A.F + comp0, A.V + comp0, A.la ? A.la + comp0 : nullptr};
  T pre[NA][NC][LPT];
  auto issue = [&](int p) {
    if (p < 0 || p >= X) {
#pragma unroll
      for (int k = 0; k < NA; ++k)
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
          for (int j = 0; j < LPT; ++j) pre[k][a][j] = T(0);
      return;
    }
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        const auto r = plane_rsrc(src[k] + 3 * a * N + (size_t)p * YZ, plane_bytes);
#pragma unroll
        for (int j = 0; j < LPT; ++j) pre[k][a][j] = bload<T>(r, eoff[j]);
      }
  };

  T qf[NC][5], qy[NC][5], qz[NC][5];
#pragma unroll
  for (int a = 0; a < NC; ++a)
#pragma unroll
    for (int k = 0; k < 5; ++k) qf[a][k] = qy[a][k] = qz[a][k] = T(0);

  const int C = (ty + 2) * TW + (tz + 2);
  T* const outD = A.aux + (size_t)(3 * iso) * N;
  T* const outR = A.aux + (size_t)(9 + 3 * iso) * N;

  auto body = [&](auto Kc, int p) {
    constexpr int K = decltype(Kc)::value;
    constexpr int sm2 = (K + 1) % 5, sm1 = (K + 2) % 5, s0 = (K + 3) % 5, sp1 = (K + 4) % 5, sp2 = K;
    const int fs = (p + 8) & (FSLOTS - 1);
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
      for (int j = 0; j < LPT; ++j) {
        const int e = tid + j * NT;
        if (j < LPT - 1 || e < TE)
          fring[fs][a][e] = stage_value<STAGE>(pre[0][a][j], NA > 1 ? pre[NA > 1 ? 1 : 0][a][j] : T(0),
                                               NA > 2 ? pre[NA > 2 ? 2 : 0][a][j] : T(0), dt, keep);
      }
    __syncthreads();
    if (p + 1 <= xe + 1) issue(p + 1);

    const int c = p - 2;
    const int zs = (p + 9) % ZSLOTS;
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      const T* t = fring[fs][a];
      qf[a][sp2] = t[C];
      qy[a][sp2] = cd1(t[C - 2 * TW], t[C - TW], t[C + TW], t[C + 2 * TW], cy.d1a, cy.d1b);
      const T dz = cd1(t[C - 2], t[C - 1], t[C + 1], t[C + 2], cz.d1a, cz.d1b);
      qz[a][sp2] = dz;
      if constexpr (D2) {
        if (a >= 1) {
          zring[zs][a - 1][(ty + 2) * TZ + tz] = dz;
          if (ty < 4) {
            const int h = hrow * TW + tz + 2;
            zring[zs][a - 1][hrow * TZ + tz] = cd1(t[h - 2], t[h - 1], t[h + 1], t[h + 2], cz.d1a, cz.d1b);
          }
        }
      }
    }

    if ((c >= xs) && own) {
      const bool gx1 = (c >= halo) && (c < X - halo), gx2 = (c > halo) && (c < X - halo);
      const T x1a = gx1 ? cx.d1a : T(0), x1b = gx1 ? cx.d1b : T(0);
      const T x2a = gx1 ? cx.d2a : T(0), x2b = gx1 ? cx.d2b : T(0), x2c = gx1 ? cx.d2c : T(0);
      const T xya = gx2 ? mxy_a : T(0), xyb = gx2 ? mxy_b : T(0);
      const T xza = gx2 ? mxz_a : T(0), xzb = gx2 ? mxz_b : T(0);
      const int cfs = (c + 8) & (FSLOTS - 1);
      const int czs = (c + 9) % ZSLOTS;
      const uint32_t pb = (uint32_t)c;
      if constexpr (D2) {
        const T* t0 = fring[cfs][0];
        const T* t1 = fring[cfs][1];
        const T* t2 = fring[cfs][2];
        // rho^a_x: d_yy, d_zz, d_xy, d_xz
        const T syy0 = cd2(t0[C - 2 * TW], t0[C - TW], qf[0][s0], t0[C + TW], t0[C + 2 * TW], y2a, y2b, y2c);
        const T szz0 = cd2(t0[C - 2], t0[C - 1], qf[0][s0], t0[C + 1], t0[C + 2], z2a, z2b, z2c);
        const T sxy0 = cd1(qy[0][sm2], qy[0][sm1], qy[0][sp1], qy[0][sp2], xya, xyb);
        const T sxz0 = cd1(qz[0][sm2], qz[0][sm1], qz[0][sp1], qz[0][sp2], xza, xzb);
        // rho^a_y: d_xx, d_zz, d_xy, d_yz
        const T sxx1 = cd2(qf[1][sm2], qf[1][sm1], qf[1][s0], qf[1][sp1], qf[1][sp2], x2a, x2b, x2c);
        const T szz1 = cd2(t1[C - 2], t1[C - 1], qf[1][s0], t1[C + 1], t1[C + 2], z2a, z2b, z2c);
        const T sxy1 = cd1(qy[1][sm2], qy[1][sm1], qy[1][sp1], qy[1][sp2], xya, xyb);
        const T* zr1 = zring[czs][0];
        const T syz1 = cd1(zr1[ty * TZ + tz], zr1[(ty + 1) * TZ + tz], zr1[(ty + 3) * TZ + tz],
                           zr1[(ty + 4) * TZ + tz], myz_a, myz_b);
        // rho^a_z: d_xx, d_yy, d_xz, d_yz
        const T sxx2 = cd2(qf[2][sm2], qf[2][sm1], qf[2][s0], qf[2][sp1], qf[2][sp2], x2a, x2b, x2c);
        const T syy2 = cd2(t2[C - 2 * TW], t2[C - TW], qf[2][s0], t2[C + TW], t2[C + 2 * TW], y2a, y2b, y2c);
        const T sxz2 = cd1(qz[2][sm2], qz[2][sm1], qz[2][sp1], qz[2][sp2], xza, xzb);
        const T* zr2 = zring[czs][1];
        const T syz2 = cd1(zr2[ty * TZ + tz], zr2[(ty + 1) * TZ + tz], zr2[(ty + 3) * TZ + tz],
                           zr2[(ty + 4) * TZ + tz], myz_a, myz_b);
        const T Dx = (syy0 - sxy1) + (szz0 - sxz2);
        const T Dy = (sxx1 - sxy0) + (szz1 - syz2);
        const T Dz = (sxx2 - sxz0) + (syy2 - syz1);
        bstore<T>(Dx, plane_rsrc(outD + pb * YZ, plane_bytes), own_b);
        bstore<T>(Dy, plane_rsrc(outD + N + pb * YZ, plane_bytes), own_b);
        bstore<T>(Dz, plane_rsrc(outD + 2 * N + pb * YZ, plane_bytes), own_b);
      }
      if (want_R) {
        const T dx_y = cd1(qf[1][sm2], qf[1][sm1], qf[1][sp1], qf[1][sp2], x1a, x1b);
        const T dx_z = cd1(qf[2][sm2], qf[2][sm1], qf[2][sp1], qf[2][sp2], x1a, x1b);
        const T dy_x = gy1 ? qy[0][s0] : T(0);
        const T dz_x = gz1 ? qz[0][s0] : T(0);
        const T dy_z = gy1 ? qy[2][s0] : T(0);
        const T dz_y = gz1 ? qz[1][s0] : T(0);
        bstore<T>(dx_y - dy_x, plane_rsrc(outR + pb * YZ, plane_bytes), own_b);
        bstore<T>(dx_z - dz_x, plane_rsrc(outR + N + pb * YZ, plane_bytes), own_b);
        bstore<T>(dy_z - dz_y, plane_rsrc(outR + 2 * N + pb * YZ, plane_bytes), own_b);
      }
    }
  };

  issue(xs - 2);
  const int pend = xe + 1;
  for (int p = xs - 2;;) {
    body(IC<0>{}, p);
    if (++p > pend) break;
    body(IC<1>{}, p);
    if (++p > pend) break;
    body(IC<2>{}, p);
    if (++p > pend) break;
    body(IC<3>{}, p);
    if (++p > pend) break;
    body(IC<4>{}, p);
    if (++p > pend) break;
  }
}

template <typename T, int STAGE, int TZ, int TY>
__global__ void __launch_bounds__(TZ* TY) k_rho_curl(const StageArgs<T> A) {
  rho_curl_body<T, STAGE, TZ, TY>(A, blockIdx.x, gridDim.x);
}

// Batched curl pass: the grid is three copies (one per isospin) of the
// stage's block map `map`; a lattice's blocks of isospin a are numbered
// a * nblocks + (block - first_block) in that map, the order rho_curl_body
// expects.
template <typename T, int STAGE, int TZ, int TY>
__global__ void __launch_bounds__(TZ* TY) k_rho_curl_batch(const BatchSlot<T>* __restrict__ slots,
                                                          const int* __restrict__ blk_inst, int map, int parity,
                                                          int grid) {
  const int iso = (int)blockIdx.x / grid;
  const int tb = (int)blockIdx.x - iso * grid;
  const int inst = blk_inst[tb];
  if (inst < 0) return;
  const BatchSlot<T>& sl = slots[inst];
  if (!sl.active) return;
  if constexpr (STAGE >= 1)
    rho_curl_body<T, STAGE, TZ, TY>(sl.args[STAGE - 1][parity], iso * sl.nblocks[map] + (tb - sl.first_block[map]),
                                    3 * sl.nblocks[map]);
}

}  // namespace skx
