// Stage kernel templates of the arrested-Newton-flow step for gfx950
// (pion, omega, Coulomb, rho, Berger and lightly bound theories).
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
// differences; one barrier per plane.  Per-site values along x (the field,
// and the raw y / z first differences of the components with mixed
// derivatives) live in 5-deep register queues, and the raw z-differences of
// every tile row of those components go to a 3-slot LDS ring for the yz
// mixed derivative.  Mixed derivatives are first differences of raw first
// differences, which is the reference's 25-tap tensor stencil
// (core/derivatives.py:121-173) regrouped.
#pragma once

#include "skx_device.h"
#include "skx_rho_curl.h"
#include "skx_stage_pair.h"

namespace skx {

template <int TH>
struct Traits;
// NF components in all; the first NST take the stencil (the others are read
// at the site only); MIXED marks the stencil components needing mixed second
// derivatives; NAUX per-site values come from the rho curl pass.
template <>
struct Traits<PION> {
  static constexpr int NF = 4, NST = 4, NSK = 4, NAUX = 0;
  static constexpr unsigned MIXED = 0xFu;
};
template <>
struct Traits<OMEGA> {
  static constexpr int NF = 5, NST = 5, NSK = 4, NAUX = 0;
  static constexpr unsigned MIXED = 0x0u;
};
template <>
struct Traits<COULOMB> {
  static constexpr int NF = 5, NST = 5, NSK = 4, NAUX = 0;
  static constexpr unsigned MIXED = 0xFu;
};
// rho: the 4 pion components take the stencil; the 9 rho components enter the
// gradient only through D^a_j and the energy only through R^a_ij, both
// formed by the curl pass (k_rho_curl) and read here per site.
template <>
struct Traits<RHO> {
  static constexpr int NF = 13, NST = 4, NSK = 4, NAUX = 9;
  static constexpr unsigned MIXED = 0xFu;
};

// Berger-Skyrme and lightly bound: the pion layout, new point functions only.
template <>
struct Traits<BERGER> {
  static constexpr int NF = 4, NST = 4, NSK = 4, NAUX = 0;
  static constexpr unsigned MIXED = 0xFu;
};
template <>
struct Traits<LIGHTLY> {
  static constexpr int NF = 4, NST = 4, NSK = 4, NAUX = 0;
  static constexpr unsigned MIXED = 0xFu;
};

template <int TH, typename T, int LD>
__device__ __forceinline__ void point_gradient(const T* phi, const T (*F)[LD], const T (*S)[LD], const T (*aux)[3],
                                               const PhysParams& P, T* g) {
  if constexpr (TH == PION) gradient_pion<T>(phi, F, S, P, g);
  if constexpr (TH == OMEGA) gradient_omega<T>(phi, F, S, P, g);
  if constexpr (TH == COULOMB) gradient_coulomb<T>(phi, F, S, P, g);
  if constexpr (TH == RHO) gradient_rho_split<T>(phi, F, S, aux, P, g);
  if constexpr (TH == BERGER) gradient_berger<T>(phi, F, S, P, g);
  if constexpr (TH == LIGHTLY) gradient_lightly<T>(phi, F, S, P, g);
}

template <int TH, typename T, int LD>
__device__ __forceinline__ double point_energy(const T* phi, const T (*F)[LD], const T (*aux)[3],
                                               const PhysParams& P) {
  if constexpr (TH == PION) return energy_pion<T, LD>(phi, F, P);
  if constexpr (TH == OMEGA) return energy_omega<T, LD>(phi, F, P);
  if constexpr (TH == COULOMB) return energy_coulomb<T, LD>(phi, F, P);
  if constexpr (TH == RHO) return energy_rho_split<T>(phi, F, aux, P);
  if constexpr (TH == BERGER) return energy_berger<T, LD>(phi, F, P);
  if constexpr (TH == LIGHTLY) return energy_lightly<T, LD>(phi, F, P);
  return 0.0;
}

// One stage over the block `blk` of `nblk` blocks that cover one lattice.
template <typename T, int TH, int STAGE, int TZ, int TY>
__device__ __forceinline__ void stage_body(const StageArgs<T>& A, int blk, int nblk) {
  using Tr = Traits<TH>;
  constexpr int NF = Tr::NF;
  constexpr int NST = Tr::NST;
  constexpr int NSK = Tr::NSK;
  constexpr int NAUX = Tr::NAUX;
  constexpr int NT = TZ * TY;
  constexpr int TW = TZ + 4;  // tile row length (z with halo)
  constexpr int TR = TY + 4;  // tile rows (y with halo)
  constexpr int TE = TW * TR;
  constexpr int NH = TE - NT;                  // halo cells of the tile
  constexpr int LPT = 1 + (NH + NT - 1) / NT;  // elements staged per thread: its own site + a halo share
  constexpr int NA = (STAGE <= 1) ? 1 : (STAGE == 2 ? 2 : 3);
  constexpr bool D2 = STAGE >= 1;
  constexpr bool ANY_MIXED = D2 && (Tr::MIXED != 0u);
  constexpr int FSLOTS = 4;  // stage-field planes: lead p, p-1, centre p-2, and the slot being refilled
  constexpr int ZSLOTS = 3;  // raw z-differences for the yz mixed derivative: planes p, p-1, p-2
  // lead-in planes whose loads are issued together (registers permitting: the
  // FP64 stage 4 keeps two)
  constexpr int LEAD_LOADS = (sizeof(T) == 8 && NA == 3) ? 2 : 4;
  static_assert(NT % 64 == 0, "block must be whole wavefronts");
  static_assert(TY >= 4, "halo rows of the z-difference ring are filled by the first four rows of threads");

  // LDS layout of the rings.  In the AOS layout the four leading components
  // of one cell are contiguous, so they move in one 16-byte (FP32) access and
  // every neighbour is the thread's own cell address plus a constant; a fifth
  // stencil component is a plane of its own after the cells.  Otherwise each
  // component is a plane of its own.  AOS for every four-component theory;
  // for the five-component ones only in the FP32 stages where it measured
  // faster (wider accesses against more registers held: omega stages 1 and 4
  // 60.4 -> 58.4 and 151.3 -> 148.6 us at 160^3, Coulomb stage 2 43.7 -> 41.5
  // us at 128^3; on Coulomb's stages 3 and 4 it spilled).
  constexpr bool AOS5 = (NST == 5) && sizeof(T) == 4 &&
                        ((TH == COULOMB && STAGE == 2) || (TH == OMEGA && (STAGE == 1 || STAGE == 4)));
  constexpr bool AOS = (NST == 4) || AOS5;
  constexpr int ZC = TR * TZ;  // cells of one z-difference plane (tile rows with halo x TZ)
  __shared__ __align__(16) T fring[FSLOTS * NST * TE];  // (rawq: declared below, stage 4 only)
  // the components with mixed derivatives are the leading NMIX stencil
  // components (Traits::MIXED = 2^NMIX - 1); only they have a z-difference ring
  static_assert((Tr::MIXED & (Tr::MIXED + 1)) == 0, "mixed components must be the leading ones");
  constexpr int NMIX = (Tr::MIXED & 8u) ? 4 : (Tr::MIXED & 4u) ? 3 : (Tr::MIXED & 2u) ? 2 : (Tr::MIXED & 1u) ? 1 : 0;
  static_assert(NMIX <= 4, "the z-difference ring holds at most the four leading components");
  __shared__ __align__(16) T zring[ANY_MIXED ? ZSLOTS * NMIX * ZC : 1];
  auto fi = [&](int slot, int a, int e) {
    if (!AOS) return (slot * NST + a) * TE + e;
    return a < 4 ? (slot * TE + e) * 4 + a : FSLOTS * 4 * TE + (slot * (NST - 4) + (a - 4)) * TE + e;
  };
  auto zi = [&](int slot, int a, int e) {
    return (AOS && NMIX == 4) ? (slot * ZC + e) * 4 + a : (slot * NMIX + a) * ZC + e;
  };
  __shared__ double red[2 * (NT / 64)];

  if ((A.flags & CHECK_DONE) && A.st->done) return;  // uniform per launch

  const int tid = threadIdx.x;
  const TileSite ts = tile_site<TZ, TY>(A, blk, nblk, tid);
  const int tz = ts.tz, ty = ts.ty, z0 = ts.z0, y0 = ts.y0, xs = ts.xs, xe = ts.xe, y = ts.y, z = ts.z;
  const bool own = ts.own;
  const int Y = A.Y, Z = A.Z, X = A.X, halo = A.halo;
  const size_t N = (size_t)A.N;
  const size_t YZ = (size_t)Y * Z;

  const T dt = T(A.P.dt);
  T keep = T(1);
  if constexpr (STAGE >= 2) keep = T(A.st->keep);
  const StageCoef<T, STAGE> sc(dt, keep);
  const SiteCoef<T> gc(A, y, z);
  const AxisCoef<T>& cx = gc.cx;

  // Tile elements this thread stages: element 0 is its own site (tile row
  // ty + 2, column tz + 2), so the raw values of its own column arrive in its
  // own registers; elements 1.. are halo cells h = tid + (j - 1) NT (halo_cell
  // numbering).  eidx = LDS index (-1: no element), eoff = byte offset in the
  // plane (kOOB outside the lattice: loads 0).
  int eidx[LPT];
  uint32_t eoff[LPT];
#pragma unroll
  for (int j = 0; j < LPT; ++j) {
    int r = ty + 2, cc = tz + 2;
    bool in_tile = true;
    if (j > 0) {
      const int h = tid + (j - 1) * NT;
      in_tile = h < NH;
      halo_cell<TW, TR>(h, r, cc);
    }
    const int gy = y0 - 2 + r, gz = z0 - 2 + cc;
    const bool valid = in_tile && gy >= 0 && gy < Y && gz >= 0 && gz < Z;
    eidx[j] = in_tile ? r * TW + cc : -1;
    eoff[j] = valid ? (uint32_t)(gy * Z + gz) * (uint32_t)sizeof(T) : kOOB;
  }
  const uint32_t own_off = own ? (uint32_t)(y * Z + z) : 0u;
  const uint32_t own_b = own ? own_off * (uint32_t)sizeof(T) : kOOB;
  const uint32_t plane_bytes = (uint32_t)(YZ * sizeof(T));
  // halo row of the z-difference ring this thread fills (rows 0,1 and TY+2,TY+3)
  const int hrow = (ty < 2) ? ty : (TY + ty);

  const T* src[3] = {A.F, A.V, A.la};
  using PlaneBuf = T[NA][NST][LPT];
  PlaneBuf pre;
  // Planes outside the lattice load as zeros: the descriptor points at a plane
  // inside it and every offset is out of range.  Every global load and store
  // of the march is issued on every path, so the compiler's memory-counter
  // waits are exact counts: a wait for one plane's loads never also waits for
  // later loads or for stores still in flight.
  auto issue = [&](PlaneBuf& buf, int p) {
    const bool pin = (p >= 0) && (p < X);
    const size_t pc = (size_t)min(max(p, 0), X - 1);
    uint32_t off[LPT];
#pragma unroll
    for (int j = 0; j < LPT; ++j) off[j] = pin ? eoff[j] : kOOB;
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
      for (int a = 0; a < NST; ++a) {
        const auto r = plane_rsrc(src[k] + a * N + pc * YZ, plane_bytes);
#pragma unroll
        for (int j = 0; j < LPT; ++j) buf[k][a][j] = bload<T>(r, off[j]);
      }
  };

  // own-column x-queues, 5 slots rotated by the unrolled plane loop: the stage
  // field of every stencil component, and the raw y / z first differences of
  // the components with mixed derivatives (the others form theirs at the
  // centre from the LDS plane, sharing the taps of their second differences,
  // so that no queue holds them: omega's five components need no mixed
  // derivative, and this keeps its stage kernels at two waves per SIMD)
  T qf[NST][5], qy[NST][5], qz[NST][5];
#pragma unroll
  for (int a = 0; a < NST; ++a)
#pragma unroll
    for (int k = 0; k < 5; ++k) qf[a][k] = qy[a][k] = qz[a][k] = T(0);
  // stage 4: raw F, V, l2 of the own column are parked in a private LDS ring
  // from their plane's arrival until it is the centre, so the RK4 combine
  // does not read them again from memory.  Slot p & 1 holds plane p - 2 when
  // plane p arrives: it is read, then overwritten with plane p.  Each thread
  // touches only its own entries, so the ring needs no barrier.  Used where it
  // fits in the 160 KB of LDS beside the plane rings (all FP32 shapes; FP64
  // pion 32x8 and 16x16, where one block per CU is resident anyway), for
  // every stencil component or, where that does not fit, for the first four
  // (FP64 Coulomb: its fifth component, V, is then read at the centre).
  constexpr size_t LDS_RINGS = sizeof(T) * (size_t(FSLOTS) * NST * TE + (ANY_MIXED ? size_t(ZSLOTS) * NMIX * TR * TZ : 0)) +
                               2 * (NT / 64) * sizeof(double);
  constexpr size_t LDS_MAX = 160 * 1024;
  constexpr bool RAWQ_ALL = sizeof(T) == 4 || LDS_RINGS + sizeof(T) * 2 * 3 * size_t(NST) * NT <= LDS_MAX;
  constexpr bool RAWQ_FOUR = NST > 4 && LDS_RINGS + sizeof(T) * 2 * 3 * size_t(4) * NT <= LDS_MAX;
  constexpr bool RAWQ = (STAGE == 4) && (RAWQ_ALL || RAWQ_FOUR);
  constexpr int NRQ = RAWQ_ALL ? NST : 4;  // components in the ring
  __shared__ __align__(16) T rawq[RAWQ ? 2 * NT * 3 * NRQ : 1];  // [slot][thread][F, V, l2 x NRQ]

  double acc0 = (STAGE == 4) ? -1.0e300 : 0.0;
  double acc1 = 0.0;
  bool bad = false;
  const int C = (ty + 2) * TW + (tz + 2);

  // One plane of the march.  K = phase mod 5: the lead plane p goes to queue
  // slot K; the centre c = p - 2 and its x-neighbours c-2..c+2 sit in slots
  // (K+1)..(K+5) mod 5.  `buf` holds plane p's loads.  HAS_CENTRE = false
  // compiles the centre site out (the four lead-in planes p < xs + 2, whose
  // centre lies before the chunk); NEXT = true issues plane p + 1 into `pre`.
  auto body = [&](auto Kc, auto HAS_CENTRE, auto NEXT, int p, PlaneBuf& buf) {
    constexpr int K = decltype(Kc)::value;
    constexpr bool HC = decltype(HAS_CENTRE)::value;
    constexpr int sm2 = (K + 1) % 5, sm1 = (K + 2) % 5, s0 = (K + 3) % 5, sp1 = (K + 4) % 5, sp2 = K;
    const int fs = (p + 8) & (FSLOTS - 1);  // p >= -2
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      if (j == 0 || eidx[j] >= 0) {
#pragma unroll
        for (int a = 0; a < NST; ++a)
          fring[fi(fs, a, eidx[j])] = sc(buf[0][a][j], NA > 1 ? buf[NA > 1 ? 1 : 0][a][j] : T(0),
                                     NA > 2 ? buf[NA > 2 ? 2 : 0][a][j] : T(0));
      }
    }
    T cF[NF], cV[NF], cl1[NF], cl2[NF], cl3[NF];
    if constexpr (RAWQ) {
      T* rq = rawq + ((p & 1) * NT + tid) * (3 * NRQ);
      if constexpr (HC) {
#pragma unroll
        for (int a = 0; a < NRQ; ++a) {
          cF[a] = rq[a];
          cV[a] = rq[NRQ + a];
          cl2[a] = rq[2 * NRQ + a];
        }
      }
#pragma unroll
      for (int a = 0; a < NRQ; ++a) {
        rq[a] = buf[0][a][0];
        rq[NRQ + a] = buf[1][a][0];
        rq[2 * NRQ + a] = buf[NA > 2 ? 2 : 0][a][0];
      }
    }

    const int c = p - 2;
    const bool center = HC && (c >= xs) && own;
    // centre-site loads, issued before the next plane's loads so that waiting
    // for them does not wait for those (offsets out of range before the chunk)
    const uint32_t coff = (c >= xs) ? own_b : kOOB;
    const size_t pb = (size_t)max(c, 0) * YZ;
    // stage 4: its own site's F, V, l1, l2, l3 for the RK4 combine -- F, V, l2
    // of the stencil components came from the raw ring above, the rest are loaded here
    if constexpr (STAGE == 4 && HC) {
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        const size_t o = a * N + pb;
        if (!RAWQ || a >= NRQ) {
          cF[a] = bload<T>(plane_rsrc(A.F + o, plane_bytes), coff);
          cV[a] = bload<T>(plane_rsrc(A.V + o, plane_bytes), coff);
          cl2[a] = bload<T>(plane_rsrc(A.la + o, plane_bytes), coff);
        }
        cl1[a] = bload<T>(plane_rsrc(A.lb + o, plane_bytes), coff);
        cl3[a] = bload<T>(plane_rsrc(A.lc + o, plane_bytes), coff);
      }
    }
    // site-only components (stages 0-3; stage 4 has them in cF / cV / cl2) and the curl-pass values
    T cP[NA][NF > NST ? NF - NST : 1];
    T cx_aux[NAUX > 0 ? NAUX : 1];
    if constexpr (NF > NST && STAGE <= 3 && HC) {
#pragma unroll
      for (int k = 0; k < NA; ++k)
#pragma unroll
        for (int a = NST; a < NF; ++a) cP[k][a - NST] = bload<T>(plane_rsrc(src[k] + a * N + pb, plane_bytes), coff);
    }
    if constexpr (NAUX > 0 && STAGE >= 1 && HC) {
      // D^a_j at aux[3a + j] (the energy pass reads R^a_ij at aux[9 + 3a + p] below)
#pragma unroll
      for (int q = 0; q < NAUX; ++q) cx_aux[q] = bload<T>(plane_rsrc(A.aux + q * N + pb, plane_bytes), coff);
    }
    __syncthreads();
    if constexpr (decltype(NEXT)::value) issue(pre, p + 1);  // past the end of the chunk: never read

    // lead plane p: own-column value and, for mixed components, raw first differences along y and z
    const int zs = (p + 9) % ZSLOTS;
#pragma unroll
    for (int a = 0; a < NST; ++a) {
      auto t = [&](int e) { return fring[fi(fs, a, e)]; };
      qf[a][sp2] = t(C);
      if ((Tr::MIXED >> a) & 1u) {
        qy[a][sp2] = cd1(t(C - 2 * TW), t(C - TW), t(C + TW), t(C + 2 * TW), gc.y1a, gc.y1b);
        const T dz = cd1(t(C - 2), t(C - 1), t(C + 1), t(C + 2), gc.z1a, gc.z1b);
        qz[a][sp2] = dz;
        if constexpr (ANY_MIXED) zring[zi(zs, a, (ty + 2) * TZ + tz)] = dz;
      }
    }
    if constexpr (ANY_MIXED) {
      if (ty < 4) {  // the halo rows of the z-difference plane
        const int h = hrow * TW + tz + 2;
#pragma unroll
        for (int a = 0; a < NST; ++a) {
          if ((Tr::MIXED >> a) & 1u) {
            auto t = [&](int e) { return fring[fi(fs, a, e)]; };
            zring[zi(zs, a, hrow * TZ + tz)] = cd1(t(h - 2), t(h - 1), t(h + 1), t(h + 2), gc.z1a, gc.z1b);
          }
        }
      }
    }

    // what this site stores (l_s for stages 1-3, F' and V' for stage 4), zero off the centre
    T lv[(D2 && STAGE <= 3) ? NF : 1], Fn[STAGE == 4 ? NF : 1], Vn[STAGE == 4 ? NF : 1];
#pragma unroll
    for (int a = 0; a < ((D2 && STAGE <= 3) ? NF : 1); ++a) lv[a] = T(0);
#pragma unroll
    for (int a = 0; a < (STAGE == 4 ? NF : 1); ++a) Fn[a] = Vn[a] = T(0);
    if (center) {
      // x guards are uniform over the block for this plane
      const bool gx1 = (c >= halo) && (c < X - halo), gx2 = (c > halo) && (c < X - halo);
      const T x1a = gx1 ? cx.d1a : T(0), x1b = gx1 ? cx.d1b : T(0);
      const T x2a = gx1 ? cx.d2a : T(0), x2b = gx1 ? cx.d2b : T(0), x2c = gx1 ? cx.d2c : T(0);
      const T xya = gx2 ? gc.mxy_a : T(0), xyb = gx2 ? gc.mxy_b : T(0);
      const T xza = gx2 ? gc.mxz_a : T(0), xzb = gx2 ? gc.mxz_b : T(0);
      const int cfs = (c + 8) & (FSLOTS - 1);
      const int czs = (c + 9) % ZSLOTS;
      T phi[NF], F[3][NST], S[6][NST];
#pragma unroll
      for (int a = 0; a < NST; ++a) {
        const T f0 = qf[a][s0];
        phi[a] = f0;
        F[0][a] = cd1(qf[a][sm2], qf[a][sm1], qf[a][sp1], qf[a][sp2], x1a, x1b);
        if ((Tr::MIXED >> a) & 1u) {
          F[1][a] = qy[a][s0];
          F[2][a] = qz[a][s0];
        } else {
          auto t = [&](int e) { return fring[fi(cfs, a, e)]; };
          F[1][a] = cd1(t(C - 2 * TW), t(C - TW), t(C + TW), t(C + 2 * TW), gc.y1a, gc.y1b);
          F[2][a] = cd1(t(C - 2), t(C - 1), t(C + 1), t(C + 2), gc.z1a, gc.z1b);
        }
        if constexpr (D2) {
          auto t = [&](int e) { return fring[fi(cfs, a, e)]; };
          S[0][a] = cd2(qf[a][sm2], qf[a][sm1], f0, qf[a][sp1], qf[a][sp2], x2a, x2b, x2c);
          S[1][a] = cd2(t(C - 2 * TW), t(C - TW), f0, t(C + TW), t(C + 2 * TW), gc.y2a, gc.y2b, gc.y2c);
          S[2][a] = cd2(t(C - 2), t(C - 1), f0, t(C + 1), t(C + 2), gc.z2a, gc.z2b, gc.z2c);
          if ((Tr::MIXED >> a) & 1u) {
            S[3][a] = cd1(qy[a][sm2], qy[a][sm1], qy[a][sp1], qy[a][sp2], xya, xyb);
            S[4][a] = cd1(qz[a][sm2], qz[a][sm1], qz[a][sp1], qz[a][sp2], xza, xzb);
            if constexpr (ANY_MIXED) {
              auto zr = [&](int row) { return zring[zi(czs, a, row * TZ + tz)]; };
              S[5][a] = cd1(zr(ty), zr(ty + 1), zr(ty + 3), zr(ty + 4), gc.myz_a, gc.myz_b);
            }
          } else {
            S[3][a] = S[4][a] = S[5][a] = T(0);
          }
        }
      }
      if constexpr (NF > NST) {
#pragma unroll
        for (int a = NST; a < NF; ++a) {
          if constexpr (STAGE == 4) phi[a] = sc(cF[a], cV[a], cl2[a]);
          else
            phi[a] = sc(cP[0][a - NST], NA > 1 ? cP[NA > 1 ? 1 : 0][a - NST] : T(0),
                        NA > 2 ? cP[NA > 2 ? 2 : 0][a - NST] : T(0));
        }
      }
      const size_t so = (size_t)c * YZ + own_off;

      if constexpr (STAGE <= 1) {
        if (STAGE == 0 || (A.flags & FUSE_ENERGY)) {
          T Rv[3][3];
          if constexpr (NAUX > 0) {
#pragma unroll
            for (int q = 0; q < 9; ++q)
              Rv[q / 3][q % 3] = bload<T>(plane_rsrc(A.aux + (9 + q) * N + (size_t)c * YZ, plane_bytes), own_b);
          }
          const double e = point_energy<TH, T, NST>(phi, F, Rv, A.P);
          acc0 += e;
          if (STAGE == 0 && A.dens_e) A.dens_e[so] = e;
        }
        if (A.flags & FUSE_BARYON) {
          const double b = baryon_density<T, NST>(phi, F, A.P);
          acc1 += b;
          if (STAGE == 0 && A.dens_b) A.dens_b[so] = b;
        }
      }
      if constexpr (D2) {
        T g[NF];
        T Dv[3][3];
        if constexpr (NAUX > 0) {
#pragma unroll
          for (int q = 0; q < 9; ++q) Dv[q / 3][q % 3] = cx_aux[q];
        }
        point_gradient<TH, T, NST>(phi, F, S, Dv, A.P, g);
        if (A.eg_out) {
#pragma unroll
          for (int a = 0; a < NF; ++a) A.eg_out[a * N + so] = double(g[a]);
        }
        if constexpr (STAGE <= 3) {
#pragma unroll
          for (int a = 0; a < NF; ++a) lv[a] = -dt * g[a];
        } else {
          // RK4 combine (core/integrator.py:170-175), recomputing k1..k4 from V and l1..l3,
          // in the storage precision: the result is rounded to it when stored
          // (FP64 form: the reference's operand order; FP32 form: the same sum regrouped as
          // F + dt (V + (l1 + l2 + l3) / 6), since k1 + 2 k2 + 2 k3 + k4 = dt (6 V + l1 + l2 + l3))
          using R = T;
          const R rdt = R(A.P.dt);
          // signed max and max |g| of this site in the storage precision: the
          // conversion to FP64 is exact and monotone, so the FP64 reduction
          // sees the same values (NaN is ignored by fmax and flagged by `bad`)
          R gmax = g[0], gabs = fabs(g[0]);
#pragma unroll
          for (int a = 0; a < NF; ++a) {
            const R Fv = cF[a];
            const R Vk = (keep != T(0)) ? cV[a] : R(0);
            const R l1 = cl1[a], l2 = cl2[a], l3 = cl3[a];
            const R l4 = -dt * g[a];
            Vn[a] = Vk + R(1.0 / 6.0) * (l1 + R(2.0) * l2 + R(2.0) * l3 + l4);
            if constexpr (sizeof(T) == 8) {
              const R k1 = rdt * Vk;
              const R k2 = rdt * (Vk + R(0.5) * l1);
              const R k3 = rdt * (Vk + R(0.5) * l2);
              const R k4 = rdt * (Vk + l3);
              Fn[a] = Fv + R(1.0 / 6.0) * (k1 + R(2.0) * k2 + R(2.0) * k3 + k4);
            } else {
              Fn[a] = Fv + rdt * (Vk + R(1.0 / 6.0) * (l1 + l2 + l3));
            }
            gmax = fmax(gmax, g[a]);
            gabs = fmax(gabs, fabs(g[a]));
            bad |= !isfinite(g[a]);
          }
          acc0 = fmax(acc0, double(gmax));
          acc1 = fmax(acc1, double(gabs));
          if (A.unit_length) {
            // S^3 renormalisation of F and tangent projection of V in FP64 in both forms
            // (compute_norm_skyrme_field / project_orthogonal_skyrme_field,
            // theories/nuclear_skyrme/kernels.py:196-272)
            double f[NSK], v[NSK], s = 0.0;
#pragma unroll
            for (int a = 0; a < NSK; ++a) {
              f[a] = double(Fn[a]);
              v[a] = double(Vn[a]);
              s += f[a] * f[a];
            }
            if (s != 0.0) {
              if constexpr (sizeof(T) == 8) {  // FP64 form: sqrt, then one division per component, as the reference
                s = sqrt(s);
#pragma unroll
                for (int a = 0; a < NSK; ++a) f[a] /= s;
              } else {
                // FP32 form (the result is rounded to FP32 on store): 1/sqrt(s) from
                // the hardware FP64 reciprocal square root and one Newton step
                // y <- y (3 - s y^2) / 2, which squares its relative error
                double y = __builtin_amdgcn_rsq(s);
                y = y * (1.5 - (0.5 * s) * (y * y));
#pragma unroll
                for (int a = 0; a < NSK; ++a) f[a] *= y;
              }
            }
            double lm = 0.0;
#pragma unroll
            for (int a = 0; a < NSK; ++a) lm += v[a] * f[a];
#pragma unroll
            for (int a = 0; a < NSK; ++a) {
              v[a] -= lm * f[a];
              Fn[a] = R(f[a]);
              Vn[a] = R(v[a]);
            }
          }
        }
      }
    }
    if constexpr (HC && D2 && STAGE <= 3) {
#pragma unroll
      for (int a = 0; a < NF; ++a) bstore<T>(lv[a], plane_rsrc(A.lout + a * N + pb, plane_bytes), coff);
    }
    if constexpr (HC && STAGE == 4) {
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        bstore<T>(Fn[a], plane_rsrc(A.Fout + a * N + pb, plane_bytes), coff);
        bstore<T>(Vn[a], plane_rsrc(A.Vout + a * N + pb, plane_bytes), coff);
      }
    }
  };

  // Lead-in: the planes xs-2 .. xs+1 fill the rings before the first centre.
  // Their loads (the first LEAD_LOADS of them at once) are in flight
  // together, so the march starts after about one memory latency rather
  // than four; from plane xs+2 on, each plane's loads are issued one plane
  // ahead.  A chunk holds at least one plane, so xs+2 <= xe+1.
  PlaneBuf pre1, pre2, pre3;
  issue(pre, xs - 2);
  if constexpr (LEAD_LOADS > 1) issue(pre1, xs - 1);
  if constexpr (LEAD_LOADS > 2) issue(pre2, xs);
  if constexpr (LEAD_LOADS > 3) issue(pre3, xs + 1);
  using NoCentre = std::false_type;
  auto lead_in = [&](auto Ic, PlaneBuf& own_buf) {
    constexpr int i = decltype(Ic)::value;
    if constexpr (i < LEAD_LOADS) body(Ic, NoCentre{}, std::bool_constant<i == LEAD_LOADS - 1>{}, xs - 2 + i, own_buf);
    else body(Ic, NoCentre{}, std::true_type{}, xs - 2 + i, pre);
  };
  lead_in(IC<0>{}, pre);
  lead_in(IC<1>{}, pre1);
  lead_in(IC<2>{}, pre2);
  lead_in(IC<3>{}, pre3);
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

  // ---- fused reductions
  stage_epilogue<NT, STAGE>(A, acc0, acc1, bad, red, blk, nblk);
}

// Stage kernels with two lanes per site (skx_stage_pair.h): RK stages 2-4 of
// the FP64 pion theory.  LANES = 2 selects them; the launcher falls back to
// one lane per site where the lattice's offsets do not fit (lane_pair_fits) or
// the caller asks for it (SKX_FLAG_ONE_LANE).
template <typename T, int TH, int STAGE>
constexpr bool has_lane_pairs() {
  return std::is_same<T, double>::value && TH == PION && STAGE >= 2;
}

template <typename T, int TH, int STAGE, int TZ, int TY, int LANES>
__global__ void __launch_bounds__((TZ * TY * LANES))
k_stage(const StageArgs<T> A) {
  if constexpr (LANES == 2) stage_body_pair<T, STAGE, TZ, TY>(A, blockIdx.x, gridDim.x);
  else stage_body<T, TH, STAGE, TZ, TY>(A, blockIdx.x, gridDim.x);
}

// Batched launch: many lattices of one theory / precision / tile shape in one
// grid.  blk_inst (block map `map` of the batch) maps every block of the grid
// to a slot (-1 = idle block); the slot carries the lattice's arguments for
// both ping-pong parities and its run of blocks in each map.
template <typename T, int TH, int STAGE, int TZ, int TY, int LANES>
__global__ void __launch_bounds__((TZ * TY * LANES))
k_stage_batch(const BatchSlot<T>* __restrict__ slots, const int* __restrict__ blk_inst, int map, int parity) {
  const int inst = blk_inst[blockIdx.x];
  if (inst < 0) return;
  const BatchSlot<T>& sl = slots[inst];
  if (!sl.active) return;
  if constexpr (STAGE >= 1 && LANES == 2)
    stage_body_pair<T, STAGE, TZ, TY>(sl.args[STAGE - 1][parity], (int)blockIdx.x - sl.first_block[map],
                                      sl.nblocks[map]);
  else if constexpr (STAGE >= 1)
    stage_body<T, TH, STAGE, TZ, TY>(sl.args[STAGE - 1][parity], (int)blockIdx.x - sl.first_block[map],
                                     sl.nblocks[map]);
}

// ------------------------------------------------------------------ per-theory launch
// Compile-time visit of (stage, tile shape) for one precision and theory.
template <typename T, int TH, int TZ, int TY, class Fn>
hipError_t visit_stage(int stage, Fn&& fn) {
  switch (stage) {
    case 0: return fn.template operator()<0, TZ, TY>();
    case 1: return fn.template operator()<1, TZ, TY>();
    case 2: return fn.template operator()<2, TZ, TY>();
    case 3: return fn.template operator()<3, TZ, TY>();
    case 4: return fn.template operator()<4, TZ, TY>();
  }
  return hipErrorInvalidValue;
}
template <typename T, int TH, class Fn>
hipError_t visit_shape(int stage, const LaunchShape& s, Fn&& fn) {
  if (s.tz == 64 && s.ty == 4) return visit_stage<T, TH, 64, 4>(stage, fn);
  if (s.tz == 32 && s.ty == 8) return visit_stage<T, TH, 32, 8>(stage, fn);
  if (s.tz == 16 && s.ty == 16) return visit_stage<T, TH, 16, 16>(stage, fn);
  // 32x16: the FP32 default of some stages (the runtime's stage_tile)
  if constexpr (sizeof(T) == 4 && TH != RHO) {
    if (s.tz == 32 && s.ty == 16) return visit_stage<T, TH, 32, 16>(stage, fn);
  }
  return hipErrorInvalidValue;
}

template <typename T, int TH>
struct StageLaunchFn {
  const StageArgs<T>& a;
  const LaunchShape& s;
  hipStream_t st;
  template <int STAGE, int TZ, int TY>
  hipError_t operator()() {
    StageArgs<T> b = a;
    b.xchunk = s.xchunk;
    b.nbz = (a.Z + TZ - 1) / TZ;
    b.nby = (a.Y + TY - 1) / TY;
    b.nbx = (a.X + s.xchunk - 1) / s.xchunk;
    const int blocks = b.nbz * b.nby * b.nbx;
    if constexpr (TH == RHO) {  // curl pass of the three isospins first (same stream)
      hipLaunchKernelGGL((k_rho_curl<T, STAGE, TZ, TY>), dim3(3 * blocks), dim3(TZ * TY), 0, st, b);
      const hipError_t r = hipGetLastError();
      if (r != hipSuccess) return r;
    }
    if constexpr (has_lane_pairs<T, TH, STAGE>()) {
      if (s.lanes == 2) {
        hipLaunchKernelGGL((k_stage<T, TH, STAGE, TZ, TY, 2>), dim3(blocks), dim3(2 * TZ * TY), 0, st, b);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((k_stage<T, TH, STAGE, TZ, TY, 1>), dim3(blocks), dim3(TZ * TY), 0, st, b);
    return hipGetLastError();
  }
};

template <typename T, int TH>
struct StageBatchLaunchFn {
  const BatchSlot<T>* slots;
  const int* blk_inst;
  int map, parity, grid, lanes;
  hipStream_t st;
  template <int STAGE, int TZ, int TY>
  hipError_t operator()() {
    if constexpr (TH == RHO) {  // curl pass of every lattice's three isospins first
      hipLaunchKernelGGL((k_rho_curl_batch<T, STAGE, TZ, TY>), dim3(3 * grid), dim3(TZ * TY), 0, st, slots,
                         blk_inst, map, parity, grid);
      const hipError_t r = hipGetLastError();
      if (r != hipSuccess) return r;
    }
    if constexpr (has_lane_pairs<T, TH, STAGE>()) {
      if (lanes == 2) {
        hipLaunchKernelGGL((k_stage_batch<T, TH, STAGE, TZ, TY, 2>), dim3(grid), dim3(2 * TZ * TY), 0, st, slots,
                           blk_inst, map, parity);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((k_stage_batch<T, TH, STAGE, TZ, TY, 1>), dim3(grid), dim3(TZ * TY), 0, st, slots, blk_inst, map,
                       parity);
    return hipGetLastError();
  }
};

template <typename T, int TH>
struct StageOccupancyFn {
  int* out;
  int lanes;
  template <int STAGE, int TZ, int TY>
  hipError_t operator()() {
    if constexpr (has_lane_pairs<T, TH, STAGE>()) {
      if (lanes == 2)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            out, reinterpret_cast<const void*>(&k_stage<T, TH, STAGE, TZ, TY, 2>), 2 * TZ * TY, 0);
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        out, reinterpret_cast<const void*>(&k_stage<T, TH, STAGE, TZ, TY, 1>), TZ * TY, 0);
  }
};

template <typename T, int TH>
hipError_t stage_launch(int stage, const StageArgs<T>& a, const LaunchShape& s, hipStream_t st) {
  return visit_shape<T, TH>(stage, s, StageLaunchFn<T, TH>{a, s, st});
}

template <typename T, int TH>
hipError_t stage_launch_batch(int stage, const LaunchShape& s, const BatchSlot<T>* slots, const int* blk_inst,
                              int map, int parity, int grid, hipStream_t st) {
  return visit_shape<T, TH>(stage, s, StageBatchLaunchFn<T, TH>{slots, blk_inst, map, parity, grid, s.lanes, st});
}

template <typename T, int TH>
hipError_t stage_occupancy(int stage, const LaunchShape& s, int* blocks) {
  return visit_shape<T, TH>(stage, s, StageOccupancyFn<T, TH>{blocks, s.lanes});
}

// One translation unit per (precision, theory) instantiates these.
#define SKX_INSTANTIATE_STAGES(T, TH)                                                                          \
  template hipError_t stage_launch<T, TH>(int, const StageArgs<T>&, const LaunchShape&, hipStream_t);          \
  template hipError_t stage_launch_batch<T, TH>(int, const LaunchShape&, const BatchSlot<T>*, const int*, int, \
                                                int, int, hipStream_t);                                        \
  template hipError_t stage_occupancy<T, TH>(int, const LaunchShape&, int*);

}  // namespace skx
