// Two lanes per site: the RK4 stage kernels of the pion theory in FP64 with
// the four sigma-model components of a site split over a lane pair.
//
// Why: in FP64 the one-lane stage kernels (stage_body) need 304-422 registers
// per lane for RK stages 2-4, so they run one wave per SIMD and spill into
// the accumulation registers.  With two components per lane, the x-queues,
// the derivatives and the plane buffers halve, these stages fit in 204-255
// registers, and a CU holds two waves per SIMD.  (Stage 1 keeps one lane per
// site: its fused energy and baryon number would need the full four-vectors
// again, and the pair kernel spilled there.)
//
// Layout: lanes 2k and 2k+1 of a wavefront own site k of the wave; lane h
// holds components h and h + 2 ("slot" q = 0, 1 is component 2q + h).  In the
// cell-major LDS rings a cell stores its components in the order 0, 2, 1, 3,
// so a lane's two components are the 16 contiguous bytes at 16 h.  In memory
// a lane's component 2q + h sits h N elements past component 2q: every load
// and store takes the descriptor of component 2q and adds h N sizeof(T) to
// the lane offset, which the launcher keeps below 2^31 bytes
// (lane_pair_fits).
//
// The contractions the Skyrme gradient needs across components (n_i, d_ij,
// C_k, the tangent projection g.phi; |F|^2 and V.F of the stage-4
// renormalisation) are a partial sum over the lane's
// two components plus the partner lane's partial, read with a DPP lane swap.
// Each lane adds the partner's partial to its own, so both lanes of a pair
// hold the same total (addition commutes).  Everything else -- the rings, the
// guards, the x-march, the exact memory waits, the reductions -- is
// stage_body's, and so is the arithmetic of each component; only the order in
// which the four components enter a contraction differs (two pairs).
#pragma once

#include "skx_device.h"

namespace skx {

// Value of the partner lane (lane ^ 1): DPP quad_perm [1, 0, 3, 2].  Both
// lanes of a pair always take the same branches (they share a site), so the
// source lane is active wherever the reading lane is.
__device__ __forceinline__ double pair_other(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffLL), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0xB1, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double pair_sum(double v) { return v + pair_other(v); }

// The pion gradient on a lane's two components: skyrme_quartic_gradient_q
// (q = 1), the mass term on sigma and the tangent projection
// (theories/nuclear_skyrme/kernels.py:1077-1150).  `sigma_lane`: slot 0 of
// this lane is component 0.
template <typename T>
__device__ __forceinline__ void gradient_pion_pair(const T* phi, const T (*F)[2], const T (*S)[2], const PhysParams& P,
                                                   bool sigma_lane, T g[2]) {
  const T k2 = T(P.k2);
  T n[3], d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) n[i] = pair_sum(F[i][0] * F[i][0] + F[i][1] * F[i][1]);
  d[0] = pair_sum(F[0][0] * F[1][0] + F[0][1] * F[1][1]);
  d[1] = pair_sum(F[0][0] * F[2][0] + F[0][1] * F[2][1]);
  d[2] = pair_sum(F[1][0] * F[2][0] + F[1][1] * F[2][1]);
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
      const T Pp = F[i][0] * Sik[0] + F[i][1] * Sik[1];
      const T Qp = Sii[0] * F[k][0] + Sii[1] * F[k][1];
      c += Pp - Qp;
    }
    C[k] = pair_sum(c);
  }
  const T w0 = T(1) + k2 * (ntot - n[0]);
  const T w1 = T(1) + k2 * (ntot - n[1]);
  const T w2 = T(1) + k2 * (ntot - n[2]);
  const T e01 = T(2) * k2 * d[0], e02 = T(2) * k2 * d[1], e12 = T(2) * k2 * d[2];
  const T c0 = k2 * C[0], c1 = k2 * C[1], c2 = k2 * C[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    g[a] = -(w0 * S[0][a] + w1 * S[1][a] + w2 * S[2][a]) - (c0 * F[0][a] + c1 * F[1][a] + c2 * F[2][a]) +
           (e01 * S[3][a] + e02 * S[4][a] + e12 * S[5][a]);
  }
  if (sigma_lane) g[0] -= T(P.m2);
  const T lm = pair_sum(g[0] * phi[0] + g[1] * phi[1]);
  g[0] -= lm * phi[0];
  g[1] -= lm * phi[1];
}

// RK4 stage 2..4 of the pion theory over block `blk` of `nblk`: a TZ x TY tile
// of sites on 2 TZ TY threads.  Same march, rings and guards as stage_body
// (its comments explain them); each lane handles two components.
template <typename T, int STAGE, int TZ, int TY>
__device__ __forceinline__ void stage_body_pair(const StageArgs<T>& A, int blk, int nblk) {
  static_assert(STAGE >= 2 && STAGE <= 4, "the lane-pair kernel covers RK stages 2-4");
  constexpr int NST = 4;
  constexpr int NS = TZ * TY;  // sites of the tile
  constexpr int NT = 2 * NS;   // threads
  constexpr int TW = TZ + 4;
  constexpr int TR = TY + 4;
  constexpr int TE = TW * TR;
  constexpr int NH = TE - NS;
  constexpr int LPT = 1 + (NH + NS - 1) / NS;
  constexpr int NA = (STAGE == 2) ? 2 : 3;
  constexpr int FSLOTS = 4;
  constexpr int ZSLOTS = 3;
  constexpr int LEAD_LOADS = 2;
  constexpr int ZC = TR * TZ;
  static_assert(NT % 64 == 0, "block must be whole wavefronts");
  static_assert(TY >= 4, "halo rows of the z-difference ring are filled by the first four rows of sites");

  __shared__ __align__(16) T fring[FSLOTS * NST * TE];
  __shared__ __align__(16) T zring[ZSLOTS * NST * ZC];
  // position 2h + q of a cell holds component 2q + h
  auto fi = [&](int slot, int pos, int e) { return (slot * TE + e) * NST + pos; };
  auto zi = [&](int slot, int pos, int e) { return (slot * ZC + e) * NST + pos; };
  __shared__ double red[2 * (NT / 64)];

  if ((A.flags & CHECK_DONE) && A.st->done) return;

  const int tid = threadIdx.x;
  const int hh = tid & 1;   // lane of the pair: components hh and hh + 2
  const int p0 = 2 * hh;    // its first position in a cell
  const int sid = tid >> 1; // site of the tile
  const TileSite ts = tile_site<TZ, TY>(A, blk, nblk, sid);
  const int tz = ts.tz, ty = ts.ty, z0 = ts.z0, y0 = ts.y0, xs = ts.xs, xe = ts.xe, y = ts.y, z = ts.z;
  const bool own = ts.own;
  const int Y = A.Y, Z = A.Z, X = A.X, halo = A.halo;
  const size_t N = (size_t)A.N;
  const size_t YZ = (size_t)Y * Z;

  const T dt = T(A.P.dt);
  const T keep = T(A.st->keep);
  const StageCoef<T, STAGE> sc(dt, keep);
  const SiteCoef<T> gc(A, y, z);
  const AxisCoef<T>& cx = gc.cx;

  // Tile elements of the lane: element 0 is its site, elements 1.. the halo
  // cells sid + (j - 1) NS, numbered as in stage_body.  Offsets carry the
  // lane's component shift hh N sizeof(T).
  int eidx[LPT];
  uint32_t eoff[LPT];
  const uint32_t lane_shift = (uint32_t)((size_t)hh * N * sizeof(T));
#pragma unroll
  for (int j = 0; j < LPT; ++j) {
    int r = ty + 2, cc = tz + 2;
    bool in_tile = true;
    if (j > 0) {
      const int hc = sid + (j - 1) * NS;
      in_tile = hc < NH;
      halo_cell<TW, TR>(hc, r, cc);
    }
    const int gy = y0 - 2 + r, gz = z0 - 2 + cc;
    const bool valid = in_tile && gy >= 0 && gy < Y && gz >= 0 && gz < Z;
    eidx[j] = in_tile ? r * TW + cc : -1;
    eoff[j] = valid ? lane_shift + (uint32_t)(gy * Z + gz) * (uint32_t)sizeof(T) : kOOB;
  }
  const uint32_t own_off = own ? (uint32_t)(y * Z + z) : 0u;
  const uint32_t own_b = own ? lane_shift + own_off * (uint32_t)sizeof(T) : kOOB;
  // a descriptor of component 2q covers component 2q + 1 of the same plane
  const uint32_t range = (uint32_t)((N + YZ) * sizeof(T));
  const int hrow = (ty < 2) ? ty : (TY + ty);

  const T* src[3] = {A.F, A.V, A.la};
  using PlaneBuf = T[NA][2][LPT];
  PlaneBuf pre;
  auto issue = [&](PlaneBuf& buf, int p) {
    const bool pin = (p >= 0) && (p < X);
    const size_t pc = (size_t)min(max(p, 0), X - 1);
    uint32_t off[LPT];
#pragma unroll
    for (int j = 0; j < LPT; ++j) off[j] = pin ? eoff[j] : kOOB;
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const auto r = plane_rsrc(src[k] + 2 * q * N + pc * YZ, range);
#pragma unroll
        for (int j = 0; j < LPT; ++j) buf[k][q][j] = bload<T>(r, off[j]);
      }
  };

  T qf[2][5], qy[2][5], qz[2][5];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int k = 0; k < 5; ++k) qf[q][k] = qy[q][k] = qz[q][k] = T(0);
  // stage 4: the lane's raw F, V, l2, [slot][thread][F, V, l2 x 2], where it
  // fits in LDS beside the rings; otherwise the combine loads them
  constexpr size_t LDS_RINGS =
      sizeof(T) * (size_t(FSLOTS) * NST * TE + size_t(ZSLOTS) * NST * ZC) + 2 * (NT / 64) * sizeof(double);
  constexpr bool RAWQ = (STAGE == 4) && LDS_RINGS + sizeof(T) * 2 * NT * 6 <= 160 * 1024;
  __shared__ __align__(16) T rawq[RAWQ ? 2 * NT * 6 : 1];

  double acc0 = (STAGE == 4) ? -1.0e300 : 0.0;
  double acc1 = 0.0;
  bool bad = false;
  const int C = (ty + 2) * TW + (tz + 2);

  auto body = [&](auto Kc, auto HAS_CENTRE, auto NEXT, int p, PlaneBuf& buf) {
    constexpr int K = decltype(Kc)::value;
    constexpr bool HC = decltype(HAS_CENTRE)::value;
    constexpr int sm2 = (K + 1) % 5, sm1 = (K + 2) % 5, s0 = (K + 3) % 5, sp1 = (K + 4) % 5, sp2 = K;
    const int fs = (p + 8) & (FSLOTS - 1);
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      if (j == 0 || eidx[j] >= 0) {
#pragma unroll
        for (int q = 0; q < 2; ++q)
          fring[fi(fs, p0 + q, eidx[j])] = sc(buf[0][q][j], buf[1][q][j], NA > 2 ? buf[NA > 2 ? 2 : 0][q][j] : T(0));
      }
    }
    T cF[2], cV[2], cl1[2], cl2[2], cl3[2];
    if constexpr (RAWQ) {
      T* rq = rawq + ((p & 1) * NT + tid) * 6;
      if constexpr (HC) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          cF[q] = rq[q];
          cV[q] = rq[2 + q];
          cl2[q] = rq[4 + q];
        }
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        rq[q] = buf[0][q][0];
        rq[2 + q] = buf[1][q][0];
        rq[4 + q] = buf[NA > 2 ? 2 : 0][q][0];
      }
    }

    const int c = p - 2;
    const bool center = HC && (c >= xs) && own;
    const uint32_t coff = (c >= xs) ? own_b : kOOB;
    const size_t pb = (size_t)max(c, 0) * YZ;
    if constexpr (STAGE == 4 && HC) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const size_t o = 2 * q * N + pb;
        if constexpr (!RAWQ) {
          cF[q] = bload<T>(plane_rsrc(A.F + o, range), coff);
          cV[q] = bload<T>(plane_rsrc(A.V + o, range), coff);
          cl2[q] = bload<T>(plane_rsrc(A.la + o, range), coff);
        }
        cl1[q] = bload<T>(plane_rsrc(A.lb + o, range), coff);
        cl3[q] = bload<T>(plane_rsrc(A.lc + o, range), coff);
      }
    }
    __syncthreads();
    if constexpr (decltype(NEXT)::value) issue(pre, p + 1);  // past the end of the chunk: never read

    const int zs = (p + 9) % ZSLOTS;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      auto t = [&](int e) { return fring[fi(fs, p0 + q, e)]; };
      qf[q][sp2] = t(C);
      qy[q][sp2] = cd1(t(C - 2 * TW), t(C - TW), t(C + TW), t(C + 2 * TW), gc.y1a, gc.y1b);
      const T dz = cd1(t(C - 2), t(C - 1), t(C + 1), t(C + 2), gc.z1a, gc.z1b);
      qz[q][sp2] = dz;
      zring[zi(zs, p0 + q, (ty + 2) * TZ + tz)] = dz;
    }
    if (ty < 4) {  // the halo rows of the z-difference plane
      const int hcell = hrow * TW + tz + 2;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        auto t = [&](int e) { return fring[fi(fs, p0 + q, e)]; };
        zring[zi(zs, p0 + q, hrow * TZ + tz)] =
            cd1(t(hcell - 2), t(hcell - 1), t(hcell + 1), t(hcell + 2), gc.z1a, gc.z1b);
      }
    }

    T lv[2] = {T(0), T(0)}, Fn[2] = {T(0), T(0)}, Vn[2] = {T(0), T(0)};
    if (center) {
      const bool gx1 = (c >= halo) && (c < X - halo), gx2 = (c > halo) && (c < X - halo);
      const T x1a = gx1 ? cx.d1a : T(0), x1b = gx1 ? cx.d1b : T(0);
      const T x2a = gx1 ? cx.d2a : T(0), x2b = gx1 ? cx.d2b : T(0), x2c = gx1 ? cx.d2c : T(0);
      const T xya = gx2 ? gc.mxy_a : T(0), xyb = gx2 ? gc.mxy_b : T(0);
      const T xza = gx2 ? gc.mxz_a : T(0), xzb = gx2 ? gc.mxz_b : T(0);
      const int cfs = (c + 8) & (FSLOTS - 1);
      const int czs = (c + 9) % ZSLOTS;
      T phi[2], F[3][2], S[6][2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const T f0 = qf[q][s0];
        phi[q] = f0;
        F[0][q] = cd1(qf[q][sm2], qf[q][sm1], qf[q][sp1], qf[q][sp2], x1a, x1b);
        F[1][q] = qy[q][s0];
        F[2][q] = qz[q][s0];
        auto t = [&](int e) { return fring[fi(cfs, p0 + q, e)]; };
        S[0][q] = cd2(qf[q][sm2], qf[q][sm1], f0, qf[q][sp1], qf[q][sp2], x2a, x2b, x2c);
        S[1][q] = cd2(t(C - 2 * TW), t(C - TW), f0, t(C + TW), t(C + 2 * TW), gc.y2a, gc.y2b, gc.y2c);
        S[2][q] = cd2(t(C - 2), t(C - 1), f0, t(C + 1), t(C + 2), gc.z2a, gc.z2b, gc.z2c);
        S[3][q] = cd1(qy[q][sm2], qy[q][sm1], qy[q][sp1], qy[q][sp2], xya, xyb);
        S[4][q] = cd1(qz[q][sm2], qz[q][sm1], qz[q][sp1], qz[q][sp2], xza, xzb);
        auto zr = [&](int row) { return zring[zi(czs, p0 + q, row * TZ + tz)]; };
        S[5][q] = cd1(zr(ty), zr(ty + 1), zr(ty + 3), zr(ty + 4), gc.myz_a, gc.myz_b);
      }
      const size_t so = (size_t)c * YZ + own_off;
      T g[2];
      gradient_pion_pair<T>(phi, F, S, A.P, hh == 0, g);
      if (A.eg_out) {
#pragma unroll
        for (int q = 0; q < 2; ++q) A.eg_out[(2 * q + hh) * N + so] = double(g[q]);
      }
      if constexpr (STAGE <= 3) {
        lv[0] = -dt * g[0];
        lv[1] = -dt * g[1];
      } else {
        // RK4 combine (core/integrator.py:170-175) in the FP64 form's operand order
        const T rdt = T(A.P.dt);
        T gmax = g[0], gabs = fabs(g[0]);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const T Vk = (keep != T(0)) ? cV[q] : T(0);
          const T l1 = cl1[q], l2 = cl2[q], l3 = cl3[q];
          const T l4 = -dt * g[q];
          Vn[q] = Vk + T(1.0 / 6.0) * (l1 + T(2.0) * l2 + T(2.0) * l3 + l4);
          const T k1 = rdt * Vk;
          const T k2 = rdt * (Vk + T(0.5) * l1);
          const T k3 = rdt * (Vk + T(0.5) * l2);
          const T k4 = rdt * (Vk + l3);
          Fn[q] = cF[q] + T(1.0 / 6.0) * (k1 + T(2.0) * k2 + T(2.0) * k3 + k4);
          gmax = fmax(gmax, g[q]);
          gabs = fmax(gabs, fabs(g[q]));
          bad |= !isfinite(g[q]);
        }
        acc0 = fmax(acc0, double(gmax));
        acc1 = fmax(acc1, double(gabs));
        if (A.unit_length) {
          // S^3 renormalisation and tangent projection (theories/nuclear_skyrme/kernels.py:196-272)
          double f[2] = {double(Fn[0]), double(Fn[1])}, v[2] = {double(Vn[0]), double(Vn[1])};
          double s = pair_sum(f[0] * f[0] + f[1] * f[1]);
          if (s != 0.0) {
            s = sqrt(s);
            f[0] /= s;
            f[1] /= s;
          }
          const double lm = pair_sum(v[0] * f[0] + v[1] * f[1]);
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            v[q] -= lm * f[q];
            Fn[q] = T(f[q]);
            Vn[q] = T(v[q]);
          }
        }
      }
    }
    if constexpr (HC && STAGE <= 3) {
#pragma unroll
      for (int q = 0; q < 2; ++q) bstore<T>(lv[q], plane_rsrc(A.lout + 2 * q * N + pb, range), coff);
    }
    if constexpr (HC && STAGE == 4) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        bstore<T>(Fn[q], plane_rsrc(A.Fout + 2 * q * N + pb, range), coff);
        bstore<T>(Vn[q], plane_rsrc(A.Vout + 2 * q * N + pb, range), coff);
      }
    }
  };

  // lead-in: planes xs-2 .. xs+1, the first two loaded together
  PlaneBuf pre1;
  issue(pre, xs - 2);
  issue(pre1, xs - 1);
  using NoCentre = std::false_type;
  auto lead_in = [&](auto Ic, PlaneBuf& own_buf) {
    constexpr int i = decltype(Ic)::value;
    if constexpr (i < LEAD_LOADS) body(Ic, NoCentre{}, std::bool_constant<i == LEAD_LOADS - 1>{}, xs - 2 + i, own_buf);
    else body(Ic, NoCentre{}, std::true_type{}, xs - 2 + i, pre);
  };
  lead_in(IC<0>{}, pre);
  lead_in(IC<1>{}, pre1);
  lead_in(IC<2>{}, pre);
  lead_in(IC<3>{}, pre);
  const int pend = xe + 1;
  using Centre = std::true_type;
  using Next = std::true_type;
  for (int p = xs + 2;;) {
    body(IC<4>{}, Centre{}, Next{}, p, pre);
    if (++p > pend) break;
    body(IC<0>{}, Centre{}, Next{}, p, pre);
    if (++p > pend) break;
    body(IC<1>{}, Centre{}, Next{}, p, pre);
    if (++p > pend) break;
    body(IC<2>{}, Centre{}, Next{}, p, pre);
    if (++p > pend) break;
    body(IC<3>{}, Centre{}, Next{}, p, pre);
    if (++p > pend) break;
  }

  stage_epilogue<NT, STAGE>(A, acc0, acc1, bad, red, blk, nblk);
}

}  // namespace skx
