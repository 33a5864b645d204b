// stage.hip -- one fused low-storage RK stage of the SPARC2 right-hand side.
//
// The LDDRK-5 step of step.f90:20-63 is
//   u_1 = a;  K_s = RHS(u_s);  u_{s+1} = a + beta_{s+1} K_s;  a <- a + dt K_5.
// One launch of k_stage computes K_s for every grid point -- the 11-point
// derivatives (derivatives.f90 ddx/ddy/ddz), the CPML auxiliary updates on
// all six faces, the pulse source and the boundary-plane zeroing of
// MP_QUIET_3D (RHS.f90:18-175) or MP_MHD_3D (RHS.f90:363-668) -- and
// immediately writes the next stage value u_{s+1} (or, on the last stage,
// the advanced base a and per-block diagnostic partials).  Stencils therefore
// read a single stage-value array; the base a is read at the point only.
//
// Blocking (2.5-D): a block owns a TX x TY column tile and marches up a chunk
// of z planes.  Fields differentiated in z live in an 11-deep register queue
// per thread; fields differentiated in x or y are staged plane by plane into
// LDS with 5-cell halos in one load phase per plane (each staged cell loads
// every field it feeds once).  The staging addresses, including the periodic
// wrap, are computed once per block.  Boundary rows use the reference's
// one-sided closures.
#include "sparc_dev.h"

namespace sparc {

// Occupancy target (waves per SIMD): the fused MHD kernel's register queue (6
// fields x 11 planes) plus nine LDS-staged fields need about 250 registers,
// so it runs 2 waves per SIMD; the MHD xy kernel (z derivatives from the z
// pass) is held to 3.  The quiet kernel keeps the next plane's values in
// registers, which fits 3 waves only in FP32 without 3-D backgrounds.
// A lane of two FP32 points holds as many register bytes as an FP64 lane, so
// it takes the FP64 occupancy.
template <typename T, bool MHD, bool BG3D, bool ZP, int VW>
constexpr int stage_waves() {
  if (VW == 2) return 2;
  if (MHD) return (ZP && sizeof(T) == 4) ? 3 : 2;
  // quiet with the one-plane-ahead point values: FP64, or FP32 with 3-D
  // backgrounds, need more registers than 3 waves per SIMD allow
  return (sizeof(T) == 8 || BG3D) ? 2 : 3;
}

// One work item: tile `tile` of z chunk `chunk` (logical = chunk*nblk_xy + tile)
// of the run whose arguments Ap points to.  ZP (MHD only): the z derivatives,
// CPML-corrected on the z faces, come from the z pass (Ap->dz) instead of the
// register queue.  VW: x points per lane (2: FP32 on even nx, F2 values); the
// tile is TX lanes = VW*TX points wide.
template <typename T, bool MHD, bool BG3D, int TX, int TY, bool ZP, int VW, bool DMA = false>
__device__ __forceinline__ void stage_item(CArgs<T> Ap, const int logical) {
  static_assert(MHD || !ZP, "the z pass serves the MHD stage only");
  static_assert(!DMA || (MHD && ZP && TX == 16 && TY == 16 && VW * sizeof(T) == 8),
                "direct-to-LDS staging: the MHD xy kernel on 16x16 tiles of 8-byte lanes");
  static_assert(VW == 1 || (sizeof(T) == 4 && (ZP || !MHD)), "two-point lanes: FP32 quiet or MHD xy");
  using V = typename Lane<T, VW>::V;
  constexpr int NT = TX * TY;
  // x halo in lane columns: 5 points, rounded up to whole lanes (DMA: to an
  // even number of lanes, so the 16-byte pairs of lane cells of a row never
  // straddle a periodic wrap)
  constexpr int HX = DMA ? (VW == 1 ? 6 : 4) : (VW == 1 ? 5 : 3);
  constexpr int LX = TX + 2 * HX, LY = TY + 10;
  constexpr int NR = TY * LX + 10 * TX;  // cross-shaped staging region
  constexpr int NS = (NR + NT - 1) / NT;
  constexpr int NQ = ZP ? 1 : (MHD ? 6 : 2);
  constexpr int NV = MHD ? 8 : 5;
  // LDS planes: 0 vx (x), 1 vy (y), 2 p (xy); MHD: 3 F3 (xy), 4 F2 (x),
  // 5 F1 (y), 6 by (x), 7 bz (xy), 8 bx (y).  DMA: planes 3-5 arrive as
  // box, boy, boz and 9 as vz, and the flux pass turns 3-5 into the fluxes.
  constexpr int NL = MHD ? (DMA ? 10 : 9) : 3;
  // SPF (quiet): two LDS plane buffers, the next plane's staging loads held
  // in registers while the current plane computes (one barrier per plane).
  // MHD: one buffer, loads then a barrier per plane -- with nine staged
  // planes a second buffer halves the blocks per CU, which costs more than
  // the prefetch hides (measured in FP32: 1.32 -> 1.37 ms at 32x16,
  // 1.59 -> 1.88 ms at 16x16 on 256^3).
  constexpr bool SPF = !MHD;
  constexpr int NB = SPF ? 2 : 1;
  constexpr int NPRE = MHD ? 10 : 3;  // raw values staged per cell
  // one spare cell per plane takes the writes of slots past the region.
  // DMA: the cross-shaped region packed (band rows of LX cells, then the
  // ten halo rows of TX cells), padded to whole 128-cell wave loads.
  constexpr int LP = DMA ? (NR + 127) / 128 * 128 : LY * LX + 1;
  __shared__ V LB[NB][NL][LP];
  __shared__ double rs[3][NT / 64 + 1];
  __shared__ RowTab<T> rtab;

  const int tid = threadIdx.x;
  rtab.fill(tid, NT);
  __syncthreads();
  const int tx = tid % TX, ty = tid / TX;
  const int nx = Ap->nx, ny = Ap->ny, nz = Ap->nz;
  const uint32_t fs = (uint32_t)Ap->fs;
  const uint32_t nxy = (uint32_t)nx * (uint32_t)ny;
  const bool first = Ap->first, last = Ap->last, periodic = Ap->periodic;

  const int nblk_xy = Ap->nblk_xy, ntx = Ap->ntx, kc = Ap->kc;
  const int tile = logical % nblk_xy, chunk = logical / nblk_xy;
  // i0, i: x point indices (the tile's first point, the lane's first point)
  const int i0 = (tile % ntx) * TX * VW, j0 = (tile / ntx) * TY;
  const int k0 = chunk * kc;
  const int k1 = min(k0 + kc, nz);
  const int i = i0 + VW * tx, j = j0 + ty;
  const bool active = (i < nx) && (j < ny);
  const uint32_t ij = active ? (uint32_t)(i + nx * j) : 0u;

  // ---- staging plan: per slot the LDS cell and the global xy offset (or
  // none: outside the domain of a non-periodic axis, or past the region).
  // Every slot loads and stores unconditionally -- a slot without a source
  // reads 0 and one past the region writes the spare cell -- so the loads of
  // a plane issue back to back behind a single wait.  Values staged into a
  // halo cell that no derivative of that plane reads are harmless. ----
  uint32_t s_off[NS];
  bool s_ok[NS];
  int s_lds[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int r = tid + s * NT;
    int row, col;
    if (r < TY * LX) {
      row = 5 + r / LX;
      col = r % LX;
    } else {
      const int r2 = r - TY * LX;
      const int hr = r2 / TX;
      row = hr < 5 ? hr : hr + TY;
      col = HX + r2 % TX;
    }
    // gi: lane column (VW points from x = VW*gi)
    const int nxl = nx / VW;
    int gi = i0 / VW - HX + col, gj = j0 - 5 + row;
    bool ok = r < NR;
    if (periodic) {
      gi = wrap(gi, nxl);
      gj = wrap(gj, ny);
    } else {
      ok = ok && gi >= 0 && gi < nxl && gj >= 0 && gj < ny;
    }
    s_ok[s] = ok;
    s_off[s] = ok ? (uint32_t)(VW * gi + nx * gj) : 0u;
    s_lds[s] = r < NR ? row * LX + col : LY * LX;
  }

  // DMA plan: one wave load moves 64 x 16 bytes -- two adjacent lane cells
  // (8 bytes each: one FP64 point, or two FP32 points) per lane -- into 128
  // consecutive packed LDS cells; chunk c of a plane is packed cells c*128 ..
  // c*128+127.  Per lane and chunk: the element offset of its cell pair in
  // the plane, or ~0 (the load reads 0: outside the domain of a
  // non-periodic axis, or past the region).  With an even number of lane
  // columns per row (nx even in FP64, a multiple of 4 in FP32), HX even and
  // the tile's first lane column a multiple of 16, a pair starts at an even
  // lane column, so it lies wholly inside or outside the domain and never
  // straddles a periodic wrap.
  constexpr int NCH = DMA ? LP / 128 : 1;
  uint32_t d_off[NCH];
  if constexpr (DMA) {
    const int lane = tid & 63;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int r = c * 128 + 2 * lane;
      int row, col;
      if (r < TY * LX) {
        row = 5 + r / LX;
        col = r % LX;
      } else {
        const int r2 = r - TY * LX;
        const int hr = r2 / TX;
        row = hr < 5 ? hr : hr + TY;
        col = HX + r2 % TX;
      }
      const int nxl = nx / VW;
      int gi = i0 / VW - HX + col, gj = j0 - 5 + row;
      bool ok = r < NR;
      if (periodic) {
        gi = wrap(gi, nxl);
        gj = wrap(gj, ny);
      } else {
        ok = ok && gi >= 0 && gi < nxl && gj >= 0 && gj < ny;
      }
      d_off[c] = ok ? (uint32_t)(VW * gi + nx * gj) : 0xFFFFFFFFu;
    }
  }

  // Array views, rebuilt from the argument pointer where they are used.
  // g: element of the lane's first point; bo: byte offset (Buf::kOut reads 0)
  const T sxs = Ap->sxs, sys = Ap->sys, sxl = Ap->sxs_lo, syl = Ap->sys_lo;
  constexpr uint32_t W = sizeof(T);
  auto UF = [&](int f, uint32_t g) -> V { return ldv<V>(Buf<T>(Ap->u[f], fs), g * W); };
  auto B3 = [&](const T* p, uint32_t g) -> V { return ldv<V>(Buf<T>(p, fs), g * W); };
  auto UFb = [&](int f, uint32_t bo) -> V { return ldv<V>(Buf<T>(Ap->u[f], fs), bo); };
  auto B3b = [&](const T* p, uint32_t bo) -> V { return ldv<V>(Buf<T>(p, fs), bo); };
  // Streams only this block reads or writes (the point values, the RK base,
  // the outputs, the CPML arrays) take the non-temporal policy, so the L2
  // keeps the staged planes for the halo reads of the neighbouring tiles:
  // MHD 256^3 FP64 xy kernel 1.93 -> 1.85 ms, quiet 512x512x300 2.69 ->
  // 2.60 ms (FP64), 1.46 -> 1.41 ms (FP32).  The FP32 MHD xy kernel gains
  // from it on its point loads (1.09 -> 1.06 ms) but loses on its stores
  // (-> 1.19 ms) and on the CPML arrays (-> 1.13 ms), so those keep the
  // default policy there (CPW: stores and CPML).
  constexpr int CPL = kStream, CPW = (MHD && sizeof(T) == 4) ? 0 : kStream;
  auto LDS1 = [&](const T* p, uint32_t bo) -> V { return ldv<V, CPL>(Buf<T>(p, fs), bo); };
  auto slot_bo = [&](int s, uint32_t pk) -> uint32_t {
    return s_ok[s] ? (pk + s_off[s]) * W : Buf<T>::kOut;
  };

  V(*L)[LP] = LB[0];  // the plane buffer in use
  // SPF: staging values of the next plane, held in registers (MHD: v, p,
  // b and B0; the fluxes are formed when the plane is written to LDS)
  V pre[NS][NPRE];
  auto pre_load = [&](int k) {
    const uint32_t pk = (uint32_t)k * nxy;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const uint32_t bo = slot_bo(s, pk);
      pre[s][0] = UFb(VX, bo);
      pre[s][1] = UFb(VY, bo);
      pre[s][2] = UFb(P, bo);
      if constexpr (MHD) {
        pre[s][3] = UFb(VZ, bo);
        pre[s][4] = UFb(BX, bo);
        pre[s][5] = UFb(BY, bo);
        pre[s][6] = UFb(BZ, bo);
        pre[s][7] = B3b(Ap->box, bo);
        pre[s][8] = B3b(Ap->boy, bo);
        pre[s][9] = B3b(Ap->boz, bo);
      }
    }
  };
  auto pre_store = [&]() {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int c = s_lds[s];
      const V vx = pre[s][0], vy = pre[s][1];
      L[0][c] = vx;
      L[1][c] = vy;
      L[2][c] = pre[s][2];
      if constexpr (MHD) {
        const V vz = pre[s][3], b0x = pre[s][7], b0y = pre[s][8], b0z = pre[s][9];
        L[3][c] = -vy * b0x + vx * b0y;  // F3 = -vy box + vx boy
        L[4][c] = -vx * b0z + vz * b0x;  // F2
        L[5][c] = -vz * b0y + vy * b0z;  // F1
        L[6][c] = pre[s][5];
        L[7][c] = pre[s][6];
        L[8][c] = pre[s][4];
      }
    }
  };
  // DMA: the ten raw planes of staging row data, straight to LDS (no
  // registers), the loads spread over the block's waves
  auto dma_plane = [&](int kk) {
    if constexpr (DMA) {
      const uint32_t pk = (uint32_t)kk * nxy;
      uint32_t bo[NCH];
#pragma unroll
      for (int c = 0; c < NCH; ++c) bo[c] = d_off[c] == 0xFFFFFFFFu ? Buf<T>::kOut : (pk + d_off[c]) * W;
      const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
      const T* const src[NL] = {Ap->u[VX], Ap->u[VY], Ap->u[P], Ap->box, Ap->boy,
                                Ap->boz,   Ap->u[BY], Ap->u[BZ], Ap->u[BX], Ap->u[VZ]};
#pragma unroll
      for (int f = 0; f < NL; ++f) {
        const Buf<T> b(src[f], fs);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          if ((f * NCH + c) % (NT / 64) != wave) continue;  // uniform over the wave
          lds_load16(b.r, &LB[0][f][c * 128], bo[c]);
        }
      }
    }
  };
  auto stage_plane = [&](int k) {
    const uint32_t pk = (uint32_t)k * nxy;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const uint32_t bo = slot_bo(s, pk);
      const int c = s_lds[s];
      const V vx = UFb(VX, bo), vy = UFb(VY, bo), p = UFb(P, bo);
      L[0][c] = vx;
      L[1][c] = vy;
      L[2][c] = p;
      if constexpr (MHD) {
        const V vz = UFb(VZ, bo), bx = UFb(BX, bo), by = UFb(BY, bo), bz = UFb(BZ, bo);
        const V b0x = B3b(Ap->box, bo), b0y = B3b(Ap->boy, bo), b0z = B3b(Ap->boz, bo);
        L[3][c] = -vy * b0x + vx * b0y;  // F3 = -vy box + vx boy
        L[4][c] = -vx * b0z + vz * b0x;  // F2
        L[5][c] = -vz * b0y + vy * b0z;  // F1
        L[6][c] = by;
        L[7][c] = bz;
        L[8][c] = bx;
      }
    }
  };
  // LDS index of (staging row, column): the full rectangle, or with DMA the
  // packed cross (rows 5..TY+4 are the band, the rest halo rows of the tile's
  // own TX columns -- the only cells the y derivatives read there)
  auto lidx = [&](int row, int col) -> int {
    if constexpr (!DMA) return row * LX + col;
    else {
      if (row >= 5 && row < 5 + TY) return (row - 5) * LX + col;
      return TY * LX + (row < 5 ? row : row - TY) * TX + (col - HX);
    }
  };
  // 11-point derivative along x / y from LDS (derivatives.f90:38-44,66-75,92-101).
  // x: the centred row for every point of the lane, then the one-sided rows
  // for points within 5 of a non-periodic edge, reading the LDS row as
  // points (point x sits at index x - i0 + VW*HX).
  auto dx_lds = [&](int f) -> V {
    const V* row = &L[f][lidx(ty + 5, 0)];
    V out;
    if constexpr (VW == 1) {
      T acc = 0, acl = 0;
      const int c0 = tx + HX;
#pragma unroll
      for (int m = 1; m <= 5; ++m) {
        const T d = row[c0 + m] - row[c0 - m];
        acc += Coef<T>::c(m - 1) * d;
        if constexpr (sizeof(T) == 4) acl += Coef<T>::cl(m - 1) * d;
      }
      if constexpr (sizeof(T) == 4) acc += acl;
      out = scl(sxs, sxl, acc);
    } else {
      T e[14];  // points i-6 .. i+7
#pragma unroll
      for (int q2 = 0; q2 < 7; ++q2) {
        const V w = row[tx + HX - 3 + q2];
        e[2 * q2] = w.x;
        e[2 * q2 + 1] = w.y;
      }
      T a0 = 0, a1 = 0, l0 = 0, l1 = 0;
#pragma unroll
      for (int m = 1; m <= 5; ++m) {
        const T d0 = e[6 + m] - e[6 - m], d1 = e[7 + m] - e[7 - m];
        a0 += Coef<T>::c(m - 1) * d0;
        a1 += Coef<T>::c(m - 1) * d1;
        if constexpr (sizeof(T) == 4) {
          l0 += Coef<T>::cl(m - 1) * d0;
          l1 += Coef<T>::cl(m - 1) * d1;
        }
      }
      if constexpr (sizeof(T) == 4) {
        a0 += l0;
        a1 += l1;
      }
      out = V(scl(sxs, sxl, a0), scl(sxs, sxl, a1));
    }
    if (!periodic && (i < 5 || i + VW - 1 >= nx - 5)) {
      constexpr int XO = VW * HX, RW = VW * LX;
      const T* rp = (const T*)row;
#pragma unroll
      for (int c = 0; c < VW; ++c) {
        const int p = i + c;
        RowSumT<T> acc(rtab);
        if (p < 5) {
#pragma unroll
          for (int m = 0; m < 11; ++m) acc.add(rp[min(m - i0 + XO, RW - 1)], p, m);
          set_lane_pt(out, c, scl(sxs, sxl, acc.sum()));
        } else if (p >= nx - 5) {
          const int r = nx - 1 - p;
#pragma unroll
          for (int m = 0; m < 11; ++m) acc.add(rp[min(max(nx - 1 - m - i0 + XO, 0), RW - 1)], r, m);
          set_lane_pt(out, c, -scl(sxs, sxl, acc.sum()));
        }
      }
    }
    return out;
  };
  auto dy_lds = [&](int f) -> V {
    const V* pl = L[f];
    const int cx = tx + HX;
    if (!periodic && j < 5) {
      typename RowSumTOf<V>::type e(rtab);
#pragma unroll
      for (int m = 0; m < 11; ++m) e.add(pl[lidx(min(max(m - j0 + 5, 0), LY - 1), cx)], j, m);
      return scl(sys, syl, e.sum());
    }
    if (!periodic && j >= ny - 5) {
      const int r = ny - 1 - j;
      typename RowSumTOf<V>::type e(rtab);
#pragma unroll
      for (int m = 0; m < 11; ++m) e.add(pl[lidx(min(max(ny - 1 - m - j0 + 5, 0), LY - 1), cx)], r, m);
      return -scl(sys, syl, e.sum());
    }
    V acc = 0;
    const int r0 = ty + 5;
    V acl = 0;
#pragma unroll
    for (int m = 1; m <= 5; ++m) {
      const V d = pl[lidx(r0 + m, cx)] - pl[lidx(r0 - m, cx)];
      acc += Coef<T>::c(m - 1) * d;
      if constexpr (sizeof(T) == 4) acl += Coef<T>::cl(m - 1) * d;
    }
    if constexpr (sizeof(T) == 4) acc += acl;
    return scl(sys, syl, acc);
  };

  // z register queue: slot m holds plane k-5+m of each z-differentiated
  // field.  Quiet: vz, p.  MHD: vz, p, bx, by, flux1, flux2.  (Unused with ZP.)
  V q[NQ][11];
  V nq[NQ];  // next plane, loaded one iteration ahead
  // Planes below 0 read as 0: the zero ghosts below the bottom.  A plane
  // kk >= nz, which no centred row reads, is loaded as plane kk-11: from
  // plane nz-5 up the window then holds exactly planes nz-11 .. nz-1, the
  // one-sided rows of the top five planes, so those come from registers
  // too.  (Loading them per top plane costs 11 dependent round trips per
  // plane, which set the length of a launch whose top chunk finishes last.)
  auto fetch = [&](int kk, V* dst) {
    if constexpr (ZP) return;
    kk = kk >= nz ? kk - 11 : kk;
    const uint32_t bo = (active && kk >= 0) ? (ij + (uint32_t)kk * nxy) * W : Buf<T>::kOut;
    dst[0] = UFb(VZ, bo);
    dst[1] = UFb(P, bo);
    if constexpr (MHD) {
      dst[2] = UFb(BX, bo);
      dst[3] = UFb(BY, bo);
      const V vx = UFb(VX, bo), vy = UFb(VY, bo), vz = dst[0];
      dst[4] = -vz * B3b(Ap->boy, bo) + vy * B3b(Ap->boz, bo);
      dst[5] = -vx * B3b(Ap->boz, bo) + vz * B3b(Ap->box, bo);
    }
  };
  if constexpr (!ZP) {
#pragma unroll
    for (int m = 0; m < 10; ++m) {
      V t[NQ];
      fetch(k0 - 5 + m, t);
#pragma unroll
      for (int f = 0; f < NQ; ++f) q[f][m] = t[f];
    }
    fetch(k0 + 5, nq);
  }

  double ke = 0.0, mr = 1.0e300, mp = 1.0e300;
  const T dt = Ap->dt, bnext = Ap->beta_next;
  // background at the point: 3-D array, or the z profile (uniform, scalar load)
  auto BG = [&](const T* p, uint32_t g, int k) -> V {
    if constexpr (BG3D) return B3(p, g);
    else return V(((CPtr<T>)p)[k]);
  };

  // Quiet: the point values of a plane -- rho, the RK base a, the 3-D
  // backgrounds where present, and on the last stage the Kahan compensation
  // (FP32) and p0 -- are loaded one plane ahead too, so every load a plane
  // needs was issued a whole iteration earlier and the iteration waits once,
  // at its top.  (In-order completion counting would otherwise make the point
  // loads wait for the next plane's staging loads as well.)  FP64 runs this
  // at 2 waves per SIMD (its registers would not fit 3).
  constexpr bool PPF = !MHD;
  constexpr int NPB = (PPF && BG3D) ? 6 : 0;  // rho0, rhoinv, sponge, gr0z, c2rho0, gp0z
  constexpr int NPC = (PPF && sizeof(T) == 4) ? NV : 0;  // comp[f], last stage
  constexpr int NPP = (PPF && BG3D) ? 1 : 0;             // p0, last stage
  constexpr int IC = 1 + NV + NPB, IP = IC + NPC;         // slots of comp and p0
  constexpr int NP = PPF ? IP + NPP : 1;
  V pc[NP], pn[NP];
  auto point_load = [&](int kk, V* dst) {
    if constexpr (PPF) {
      const uint32_t bo = active ? (ij + (uint32_t)kk * nxy) * W : Buf<T>::kOut;
      dst[0] = LDS1(Ap->u[RHO], bo);
#pragma unroll
      for (int f = 0; f < NV; ++f) dst[1 + f] = LDS1(Ap->a[f], bo);
      if constexpr (NPB > 0) {
        dst[1 + NV + 0] = LDS1(Ap->rho0, bo);
        dst[1 + NV + 1] = LDS1(Ap->rhoinv, bo);
        dst[1 + NV + 2] = LDS1(Ap->sponge, bo);
        dst[1 + NV + 3] = LDS1(Ap->gr0z, bo);
        dst[1 + NV + 4] = LDS1(Ap->c2rho0, bo);
        dst[1 + NV + 5] = LDS1(Ap->gp0z, bo);
      }
      if (last) {
        if constexpr (NPC > 0) {
#pragma unroll
          for (int f = 0; f < NV; ++f) dst[IC + f] = ldcomp_raw<V, CPL>(Buf<uint16_t>(Ap->comp[f], fs), bo);
        }
        if constexpr (NPP > 0) dst[IP] = LDS1(Ap->p0, bo);
      }
    }
  };
  // MHD xy kernel: a plane's staging loads are issued, then its point values
  // (rho, vz, the six z derivatives, a, the 3-D backgrounds; p0 and the FP32
  // compensation on the last stage), then the wait covers the staging loads
  // only; the LDS write and barrier overlap the point loads' latency.
  constexpr bool MPX = MHD && ZP;
  constexpr int MP_A = 8, MP_BG = 16, MP_P0 = 33, MP_C = 34;
  constexpr int NPM = MPX ? MP_C + (sizeof(T) == 4 ? 8 : 0) : 1;
  V pm[NPM];
  auto mpoint_load = [&](int kk) {
    if constexpr (MPX) {
      const uint32_t bo = active ? (ij + (uint32_t)kk * nxy) * W : Buf<T>::kOut;
      // vz and B0 were just staged for the tile: those hit L2
      pm[0] = LDS1(Ap->u[RHO], bo);
      pm[1] = UFb(VZ, bo);
#pragma unroll
      for (int f = 0; f < 6; ++f) pm[2 + f] = LDS1(Ap->dz[f], bo);
#pragma unroll
      for (int f = 0; f < 8; ++f) pm[MP_A + f] = LDS1(Ap->a[f], bo);
      const T* const bgp[17] = {Ap->rho0, Ap->rhoinv, Ap->sponge, Ap->red, Ap->box, Ap->boy,
                                Ap->boz, Ap->gr0x, Ap->gr0y, Ap->gr0z, Ap->c2rho0, Ap->gp0x,
                                Ap->gp0y, Ap->gp0z, Ap->cbox, Ap->cboy, Ap->cboz};
      // rhoinv (slot 1) reads nothing when the kernel forms it from rho0
      const uint32_t bo_ri = (sizeof(T) == 8 && Ap->rinv_rho0) ? Buf<T>::kOut : bo;
#pragma unroll
      for (int q2 = 0; q2 < 17; ++q2)
        pm[MP_BG + q2] = (q2 >= 4 && q2 <= 6) ? B3b(bgp[q2], bo) : LDS1(bgp[q2], q2 == 1 ? bo_ri : bo);
      if (last) {
        pm[MP_P0] = LDS1(Ap->p0, bo);
        if constexpr (sizeof(T) == 4) {
#pragma unroll
          for (int f = 0; f < 8; ++f) pm[MP_C + f] = ldcomp_raw<V, CPL>(Buf<uint16_t>(Ap->comp[f], fs), bo);
        }
      }
    }
  };
  // background slot at the point in pm (MHD xy kernel)
  auto PMB = [&](int q2) -> V { return pm[MP_BG + q2]; };

  // Source factors constant along z, loaded once per work item: the pulse's
  // amp * sx(i) * sy(j), or forcing(x, y) with PULSE = 0; each plane then
  // multiplies by its z factor, a scalar load.  (Loaded per plane, their
  // wait would also wait for the next plane's prefetched loads, which
  // complete in order before them.)
  const int pdir = Ap->pulse_dir;
  V srcxy = 0;
  if (Ap->fmap) {
    srcxy = ldv<V>(Buf<T>(Ap->fmap, nxy), (uint32_t)(i + nx * j) * W);
  } else if (pdir >= 1 && pdir <= 3) {
    const T amp = Ap->amp_tab ? Ap->amp_tab[*Ap->tstep] : Ap->src_amp;
    srcxy = amp * ldv<V>(Buf<T>(Ap->sx, nx), (uint32_t)i * W) * Buf<T>(Ap->sy, ny)[j];
  }

  if constexpr (SPF) pre_load(k0);
  if constexpr (PPF) point_load(k0, pn);
  // the first plane's loads complete before the loop, so that on entry, as
  // on every later pass, the loop-top wait has only stores left to wait on
  // (vmcnt 0, expcnt and lgkmcnt not waited on)
  if constexpr (SPF) __builtin_amdgcn_s_waitcnt(0x0F70);
  for (int k = k0; k < k1; ++k) {
    SPARC_REFRESH(Ap);
    L = LB[NB == 2 ? (k & 1) : 0];
    if constexpr (SPF) {
      // Every load of the previous iteration (this plane's staging and point
      // values, the window's next plane) completes here, in one wait, before
      // the next plane's loads issue.  Left to itself the compiler may defer
      // part of it to the first use; the in-order load counter then makes
      // that later wait cover the next plane's loads too, and the prefetch
      // stops overlapping the compute (quiet 512x512x300 FP64 stage 2.21 ->
      // 2.63 ms).  The previous plane's NV output stores, the youngest
      // operations, stay in flight.  0x0F70 | NV: vmcnt NV, expcnt and
      // lgkmcnt not waited on.
      static_assert(NV < 16, "vmcnt field");
      __builtin_amdgcn_s_waitcnt(0x0F70 | NV);
      pre_store();
      if constexpr (PPF) {
#pragma unroll
        for (int i = 0; i < NP; ++i) pc[i] = pn[i];
      }
      if (k + 1 < k1) {
        pre_load(k + 1);
        if constexpr (PPF) point_load(k + 1, pn);
      }
    } else if constexpr (MPX && DMA) {
      dma_plane(k);
      __builtin_amdgcn_sched_barrier(0);  // the plane's LDS loads stay older than the point loads
      mpoint_load(k);
      // all but the 33 youngest loads (the point values: 33, 34 on the last
      // stage) done: the plane's LDS loads have landed.  0x8F71: vmcnt 33,
      // expcnt and lgkmcnt not waited on.
      __builtin_amdgcn_s_waitcnt(0x8F71);
      __syncthreads();
      // fluxes from v and B0 at each staged cell, in place of B0 (the same
      // expressions as the register staging)
      for (int r = tid; r < NR; r += NT) {
        const V vx = L[0][r], vy = L[1][r], vz = L[9][r];
        const V b0x = L[3][r], b0y = L[4][r], b0z = L[5][r];
        L[3][r] = -vy * b0x + vx * b0y;  // F3 = -vy box + vx boy
        L[4][r] = -vx * b0z + vz * b0x;  // F2
        L[5][r] = -vz * b0y + vy * b0z;  // F1
      }
    } else if constexpr (MPX) {
      pre_load(k);
      mpoint_load(k);
      pre_store();
    } else {
      stage_plane(k);
    }
    // ---- z derivatives (derivatives.f90:481-487,509-518,536-545) ----
    V dz[NQ];
    if constexpr (!ZP) {
#pragma unroll
      for (int f = 0; f < NQ; ++f) q[f][10] = nq[f];
      fetch(k + 6, nq);
      if (active) {
        const T st = ((CPtr<T>)Ap->stretch)[k];
        T stl = 0;
        if constexpr (sizeof(T) == 4) stl = ((CPtr<T>)Ap->stretch_lo)[k];
        if (k < nz - 5) {
#pragma unroll
          for (int f = 0; f < NQ; ++f) {
            V acc = 0, acl = 0;
#pragma unroll
            for (int m = 1; m <= 5; ++m) {
              const V d = q[f][5 + m] - q[f][5 - m];
              acc += Coef<T>::c(m - 1) * d;
              if constexpr (sizeof(T) == 4) acl += Coef<T>::cl(m - 1) * d;
            }
            if constexpr (sizeof(T) == 4) acc += acl;
            dz[f] = scl(st, stl, acc);
          }
        } else {
          // top block: mirrored one-sided rows over planes nz-1 .. nz-11,
          // all in the window (see fetch): row m of plane nz-1-r sits in
          // slot (r + 16 - m) mod 11, so the sum runs over the slots
          const int r = nz - 1 - k;
#pragma unroll
          for (int f = 0; f < NQ; ++f) {
            typename RowSumTOf<V>::type acc(rtab);
#pragma unroll
            for (int m = 0; m < 11; ++m) acc.add(q[f][m], r, (r + 16 - m) % 11);
            dz[f] = -scl(st, stl, acc.sum());
          }
        }
      }
    }
    __syncthreads();
    SPARC_REFRESH(Ap);

    // The plane's outputs (next stage value, or the new base and its FP32
    // compensation) are stored after the point block by every wave, a lane
    // outside the grid at an offset past the buffer's range (dropped): each
    // iteration then ends with the same stores on every path, and the next
    // plane's wait for its prefetched loads can leave them in flight (loads
    // and stores complete in issue order).
    const uint32_t g = ij + (uint32_t)k * nxy;
    V outv[NV], outc[NV];
#pragma unroll
    for (int f = 0; f < NV; ++f) outv[f] = outc[f] = 0;
    if (active) {
      // xy derivatives and centre values from LDS
      V dvxdx = dx_lds(0), dvydy = dy_lds(1), gpx = dx_lds(2), gpy = dy_lds(2);
      const int cc = lidx(ty + 5, tx + HX);
      const V vx = L[0][cc], vy = L[1][cc];
      const V rho = PPF ? pc[0] : (MPX ? pm[0] : UF(RHO, g));
      const V vz = ZP ? (MPX ? pm[1] : UF(VZ, g)) : q[0][5], p = ZP ? L[2][cc] : q[1][5];
      // z pass output, in CPML family order: vz, p, bx, by, F2, F1
      auto DZ = [&](int f) -> V { return MPX ? pm[2 + f] : ldv<V>(Buf<T>(Ap->dz[f], fs), g * W); };
      V dvzdz = ZP ? DZ(0) : dz[0], gpz = ZP ? DZ(1) : dz[1];
      V dxf3 = 0, dyf3 = 0, dxf2 = 0, dyf1 = 0, dxby = 0, dxbz = 0, dybz = 0, dybx = 0, bz = 0;
      V dzbx = 0, dzby = 0, dzf1 = 0, dzf2 = 0;
      if constexpr (MHD) {
        dxf3 = dx_lds(3);
        dyf3 = dy_lds(3);
        dxf2 = dx_lds(4);
        dyf1 = dy_lds(5);
        dxby = dx_lds(6);
        dxbz = dx_lds(7);
        dybz = dy_lds(7);
        dybx = dy_lds(8);
        bz = L[7][cc];
        if constexpr (ZP) {
          dzbx = DZ(2);
          dzby = DZ(3);
          dzf2 = DZ(4);
          dzf1 = DZ(5);
        } else {
          dzbx = dz[2];
          dzby = dz[3];
          dzf1 = dz[4];
          dzf2 = dz[5];
        }
      }
      // ---- CPML (RHS.f90:50-93, 424-546); with ZP the z faces were done
      // by the z pass ----
#pragma unroll
      for (int fz = ZBOT; fz <= ZTOP; ++fz) {
        if constexpr (ZP) break;
        const auto& F = Ap->face[fz];
        const int kp = k - F.off;
        if (!F.active || kp < 0 || kp >= F.n) continue;
        const uint32_t fp = ij + nxy * (uint32_t)kp;
        if constexpr (MHD) {
          V v[6] = {gpz, dvzdz, dzf2, dzf1, dzbx, dzby};
          constexpr int fam[6] = {1, 0, 4, 5, 2, 3};
          pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
          gpz = v[0], dvzdz = v[1], dzf2 = v[2], dzf1 = v[3], dzbx = v[4], dzby = v[5];
        } else {
          V v[2] = {gpz, dvzdz};
          constexpr int fam[2] = {1, 0};
          pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
          gpz = v[0], dvzdz = v[1];
        }
      }
      // x faces: with two points per lane and an even face width and offset,
      // a lane pair lies wholly inside or outside the face and its face
      // element is even, so the pair goes through at once; otherwise point by
      // point (a pair can straddle the face edge, and an x face row of n
      // points need not start on a lane boundary)
#pragma unroll
      for (int fx = XBOT; fx <= XTOP; ++fx) {
        const auto& F = Ap->face[fx];
        if (!F.active) continue;
        if constexpr (VW == 2) {
          if (((F.n | F.off) & 1) == 0) {
            const int ip = i - F.off;
            if (ip >= 0 && ip < F.n) {
              const uint32_t fp = (uint32_t)(ip + F.n * (j + ny * k));
              if constexpr (MHD) {
                V v[6] = {gpx, dvxdx, dxf3, dxf2, dxby, dxbz};
                constexpr int fam[6] = {1, 0, 4, 5, 2, 3};
                pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
                gpx = v[0], dvxdx = v[1], dxf3 = v[2], dxf2 = v[3], dxby = v[4], dxbz = v[5];
              } else {
                V v[2] = {gpx, dvxdx};
                constexpr int fam[2] = {1, 0};
                pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
                gpx = v[0], dvxdx = v[1];
              }
            }
            continue;
          }
        }
#pragma unroll
        for (int c = 0; c < VW; ++c) {
          const int ip = i + c - F.off;
          if (ip < 0 || ip >= F.n) continue;
          const uint32_t fp = (uint32_t)(ip + F.n * (j + ny * k));
          if constexpr (MHD) {
            T v[6] = {lane_pt(gpx, c),  lane_pt(dvxdx, c), lane_pt(dxf3, c),
                      lane_pt(dxf2, c), lane_pt(dxby, c),  lane_pt(dxbz, c)};
            constexpr int fam[6] = {1, 0, 4, 5, 2, 3};
            pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
            set_lane_pt(gpx, c, v[0]);
            set_lane_pt(dvxdx, c, v[1]);
            set_lane_pt(dxf3, c, v[2]);
            set_lane_pt(dxf2, c, v[3]);
            set_lane_pt(dxby, c, v[4]);
            set_lane_pt(dxbz, c, v[5]);
          } else {
            T v[2] = {lane_pt(gpx, c), lane_pt(dvxdx, c)};
            constexpr int fam[2] = {1, 0};
            pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
            set_lane_pt(gpx, c, v[0]);
            set_lane_pt(dvxdx, c, v[1]);
          }
        }
      }
#pragma unroll
      for (int fy = YBOT; fy <= YTOP; ++fy) {
        const auto& F = Ap->face[fy];
        const int jp = j - F.off;
        if (!F.active || jp < 0 || jp >= F.n) continue;
        const uint32_t fp = (uint32_t)(i + nx * (jp + F.n * k));
        if constexpr (MHD) {
          V v[6] = {gpy, dvydy, dyf3, dyf1, dybx, dybz};
          constexpr int fam[6] = {1, 0, 4, 5, 2, 3};
          pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
          gpy = v[0], dvydy = v[1], dyf3 = v[2], dyf1 = v[3], dybx = v[4], dybz = v[5];
        } else {
          V v[2] = {gpy, dvydy};
          constexpr int fam[2] = {1, 0};
          pml_face<CPW>(F, fp, v, fam, first, last, bnext, dt);
          gpy = v[0], dvydy = v[1];
        }
      }

      // ---- pointwise RHS ----
      // background at the point (quiet 3-D backgrounds come from the prefetch)
      auto BGP = [&](const T* p, int slot) -> V {
        if constexpr (NPB > 0) return pc[1 + NV + slot];
        else if constexpr (MPX) return PMB(slot);  // rho0, rhoinv, sponge
        else return BG(p, g, k);
      };
      const V rho0 = BGP(Ap->rho0, 0), sp = BGP(Ap->sponge, 2);
      V rhoinv = BGP(Ap->rhoinv, 1);
      if constexpr (MPX && sizeof(T) == 8) {
        if (Ap->rinv_rho0) rhoinv = T(1) / rho0;  // IEEE division: the host's bits
      }
      const V div = dvxdx + dvydy + dvzdz;
      V K[8];
      if constexpr (!MHD) {
        const T gk = ((CPtr<T>)Ap->g)[k];
        // RHS.f90:95-101,136,152
        K[RHO] = -BGP(Ap->gr0z, 3) * vz - rho0 * div - sp * rho;
        K[VX] = -rhoinv * gpx - sp * vx;
        K[VY] = -rhoinv * gpy - sp * vy;
        K[VZ] = -rhoinv * (gpz + rho * gk) - vz * sp;
        K[P] = -BGP(Ap->c2rho0, 4) * div - vz * BGP(Ap->gp0z, 5) - sp * p;
      } else {
        const V red = MPX ? PMB(3) : B3(Ap->red, g);
        const V b0x = MPX ? PMB(4) : B3(Ap->box, g), b0y = MPX ? PMB(5) : B3(Ap->boy, g),
                b0z = MPX ? PMB(6) : B3(Ap->boz, g);
        const V bx = ZP ? L[8][cc] : q[2][5], by = ZP ? L[6][cc] : q[3][5];
        // RHS.f90:548-580
        K[BX] = red * (dyf3 - dzf2);
        K[BY] = red * (dzf1 - dxf3);
        K[BZ] = red * (dxf2 - dyf1);
        const V cbx = dybz - dzby, cby = -dxbz + dzbx, cbz = dxby - dybx;
        const V gr0x = MPX ? PMB(7) : B3(Ap->gr0x, g), gr0y = MPX ? PMB(8) : B3(Ap->gr0y, g),
                gr0z = MPX ? PMB(9) : B3(Ap->gr0z, g), c2r = MPX ? PMB(10) : B3(Ap->c2rho0, g),
                gp0x = MPX ? PMB(11) : B3(Ap->gp0x, g), gp0y = MPX ? PMB(12) : B3(Ap->gp0y, g),
                gp0z = MPX ? PMB(13) : B3(Ap->gp0z, g);
        K[RHO] = -gr0x * vx - gr0y * vy - gr0z * vz - rho0 * div - sp * rho;
        K[P] = -c2r * div - vx * gp0x - vy * gp0y - vz * gp0z - sp * p;
        // CROSS(curl b, B0) and CROSS(curl B0, b), derivatives.f90:646-655
        const V c1 = cby * b0z - cbz * b0y, c2 = b0x * cbz - cbx * b0z, c3 = cbx * b0y - cby * b0x;
        const V cox = MPX ? PMB(14) : B3(Ap->cbox, g), coy = MPX ? PMB(15) : B3(Ap->cboy, g),
                coz = MPX ? PMB(16) : B3(Ap->cboz, g);
        const V e1 = coy * bz - coz * by;
        const V e2 = bx * coz - cox * bz;
        const V e3 = cox * by - coy * bx;
        K[VX] = -rhoinv * (gpx - c1 * red) - sp * vx + rhoinv * e1 * red;
        K[VY] = -rhoinv * (gpy - c2 * red) - sp * vy + rhoinv * e2 * red;
        K[VZ] = -rhoinv * (gpz - c3 * red) - sp * vz + rhoinv * e3 * red - rhoinv * rho * ((CPtr<T>)Ap->g)[k];
      }
      // pulse source (driver.f90:137-148 separable form; RHS.f90:138-145,632-639),
      // or with PULSE = 0 forcing(x, y) * source_dep(z) on vz (RHS.f90:146-147, 640-641)
      if (Ap->fmap) {
        K[VZ] += srcxy * ((CPtr<T>)Ap->sdep)[k];
      } else if (pdir >= 1 && pdir <= 3) {
        const V s = srcxy * ((CPtr<T>)Ap->sz)[k];
        if (pdir == 1) K[VX] += s;
        else if (pdir == 2) K[VY] += s;
        else K[VZ] += s;
      }
      // boundary-plane zeroing (RHS.f90:154-164,647-657)
      const bool zero = (k == 0) || (k == nz - 1) || (!periodic && (j == 0 || j == ny - 1));
#pragma unroll
      for (int c = 0; c < VW; ++c) {
        if (zero || (!periodic && (i + c == 0 || i + c == nx - 1))) {
#pragma unroll
          for (int f = 0; f < NV; ++f) set_lane_pt(K[f], c, T(0));
        }
      }
      // FFT damping (RHS.f90:166-172, 659-665): after the zeroing, RHSv -=
      // damp_velocity(v), the terms computed from this stage's u beforehand
      if (Ap->damp[0]) {
#pragma unroll
        for (int f = 0; f < 3; ++f) K[VX + f] = K[VX + f] - ldv<V>(Buf<T>(Ap->damp[f], fs), g * W);
      }
      if (!last) {
        // next stage value, step.f90:33
#pragma unroll
        for (int f = 0; f < NV; ++f) {
          const V a0 = PPF ? pc[1 + f] : (MPX ? pm[MP_A + f] : ldv<V>(Buf<T>(Ap->a[f], fs), g * W));
          outv[f] = a0 + scl(bnext, Ap->beta_lo, K[f]);
        }
      } else {
        V (&an)[NV] = outv;
#pragma unroll
        for (int f = 0; f < NV; ++f) {
          const V a0 = PPF ? pc[1 + f] : (MPX ? pm[MP_A + f] : ldv<V>(Buf<T>(Ap->a[f], fs), g * W));
          if constexpr (sizeof(T) == 4) {
            // compensated update: a - c carries the bits a + dt K loses
            V c0;
            if constexpr (NPC > 0) c0 = comp_of_raw(pc[IC + f]);
            else if constexpr (MPX) c0 = comp_of_raw(pm[MP_C + f]);
            else c0 = ldcomp<V>(Buf<uint16_t>(Ap->comp[f], fs), g * W);
            const V y = scl(dt, Ap->dt_lo, K[f]) - c0;
            const V t = a0 + y;
            outc[f] = (t - a0) - y;
            an[f] = t;
          } else {
            an[f] = a0 + dt * K[f];
          }
        }
        // diagnostics: KE norm on the new state, minima on the stage-5 state
        V p0v;
        if constexpr (NPP > 0) p0v = pc[IP];
        else if constexpr (MPX) p0v = pm[MP_P0];
        else p0v = BG(Ap->p0, g, k);
#pragma unroll
        for (int c = 0; c < VW; ++c) {
          const double r0 = lane_pt(rho0, c), ux = lane_pt(an[VX], c), uy = lane_pt(an[VY], c),
                       uz = lane_pt(an[VZ], c);
          const double m = (r0 + (double)lane_pt(an[RHO], c)) * (ux * ux + uy * uy + uz * uz) / 2.0;
          ke += m * m;
          mr = fmin(mr, (double)lane_pt(rho, c) + r0);
          mp = fmin(mp, (double)lane_pt(p, c) + (double)lane_pt(p0v, c));
        }
      }
    }
    {
      // element offsets whose byte offsets are Buf::kOut (past every range)
      const uint32_t ge = active ? g : Buf<T>::kOut / (uint32_t)sizeof(T);
      const uint32_t gc = active ? g : Buf<T>::kOut / 2u;
      // un is a itself on the last stage
      if constexpr (sizeof(T) == 4) {
        if (last) {
#pragma unroll
          for (int f = 0; f < NV; ++f) stcomp<CPW>(Buf<uint16_t>(Ap->comp[f], fs), gc, outc[f]);
        }
      }
#pragma unroll
      for (int f = 0; f < NV; ++f) stv<CPW>(Buf<T>(Ap->un[f], fs), ge, outv[f]);
    }
    // single buffer: the next plane's staging must wait for these reads
    if constexpr (NB == 1) __syncthreads();

    if constexpr (!ZP) {
#pragma unroll
      for (int f = 0; f < NQ; ++f)
#pragma unroll
        for (int m = 0; m < 10; ++m) q[f][m] = q[f][m + 1];
    }
  }

  if (last) {
    // block reduction: wave shuffles, then one LDS slot per wave
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      ke += __shfl_xor(ke, o, 64);
      mr = fmin(mr, __shfl_xor(mr, o, 64));
      mp = fmin(mp, __shfl_xor(mp, o, 64));
    }
    const int w = tid / 64, lane = tid % 64;
    if (lane == 0) {
      rs[0][w] = ke;
      rs[1][w] = mr;
      rs[2][w] = mp;
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0, r = 1.0e300, pp = 1.0e300;
      for (int q2 = 0; q2 < NT / 64; ++q2) {
        s += rs[0][q2];
        r = fmin(r, rs[1][q2]);
        pp = fmin(pp, rs[2][q2]);
      }
      double* const part = Ap->partial;
      part[3 * (long)logical + 0] = s;
      part[3 * (long)logical + 1] = r;
      part[3 * (long)logical + 2] = pp;
    }
  }
}


// ---------------------------------------------------------------------------
// z pass of the MHD stage: the six z derivatives of ddz (derivatives.f90:
// 451-564) -- vz, p, bx, by and the fluxes F2 = -vx box... (in CPML family
// order: vz, p, bx, by, F2, F1) -- with the z-face CPML applied
// (RHS.f90:424-546), written to Ap->dz for the xy kernel.  One column per
// thread, NT consecutive columns per block, marching a z chunk with an
// 11-deep register queue; no LDS and no barriers, so the waves of a block
// stream independently.  Work item: column block cb of z chunk `chunk`
// (logical = chunk*ncb + cb).
// ---------------------------------------------------------------------------
constexpr int ZP_NT = 256;

template <typename T>
__device__ __forceinline__ void zpass_item(CArgs<T> Ap, const int logical) {
  constexpr int NT = ZP_NT;
  const int nz = Ap->nz;
  const uint32_t fs = (uint32_t)Ap->fs;
  const uint32_t nxy = (uint32_t)Ap->nx * (uint32_t)Ap->ny;
  const int ncb = (int)((nxy + NT - 1) / NT), kc = Ap->zkc;
  const int cb = logical % ncb, chunk = logical / ncb;
  const uint32_t col = (uint32_t)cb * NT + threadIdx.x;
  const bool active = col < nxy;
  const uint32_t ij = active ? col : 0u;
  const int k0 = chunk * kc, k1 = min(k0 + kc, nz);
  const bool first = Ap->first, last = Ap->last;
  const T dt = Ap->dt, bnext = Ap->beta_next;
  constexpr uint32_t W = sizeof(T);

  // a column's values are read once (the queue), and the results written
  // once: all non-temporal (256^3 MHD z pass 0.481 -> 0.477 ms FP64)
  auto UFs = [&](int f, uint32_t bo) -> T { return ldv<T, kStream>(Buf<T>(Ap->u[f], fs), bo); };
  auto B3s = [&](const T* p, uint32_t bo) -> T { return ldv<T, kStream>(Buf<T>(p, fs), bo); };
  // the six fields at plane kk in family order; below plane 0 they read 0
  // (the zero ghosts below the bottom); a plane kk >= nz, which no centred
  // row reads, is loaded as plane kk-11, so from plane nz-5 up the window
  // holds planes nz-11 .. nz-1, the top five planes' one-sided rows
  auto fetch = [&](int kk, T* dst) {
    kk = kk >= nz ? kk - 11 : kk;
    const uint32_t bo = (active && kk >= 0) ? (ij + (uint32_t)kk * nxy) * W : Buf<T>::kOut;
    const T vx = UFs(VX, bo), vy = UFs(VY, bo), vz = UFs(VZ, bo);
    const T b0x = B3s(Ap->box, bo), b0y = B3s(Ap->boy, bo), b0z = B3s(Ap->boz, bo);
    dst[0] = vz;
    dst[1] = UFs(P, bo);
    dst[2] = UFs(BX, bo);
    dst[3] = UFs(BY, bo);
    dst[4] = -vx * b0z + vz * b0x;  // F2
    dst[5] = -vz * b0y + vy * b0z;  // F1
  };
  T q[6][11], nq[6];
#pragma unroll
  for (int m = 0; m < 10; ++m) {
    T t[6];
    fetch(k0 - 5 + m, t);
#pragma unroll
    for (int f = 0; f < 6; ++f) q[f][m] = t[f];
  }
  fetch(k0 + 5, nq);

  for (int k = k0; k < k1; ++k) {
    SPARC_REFRESH(Ap);
#pragma unroll
    for (int f = 0; f < 6; ++f) q[f][10] = nq[f];
    fetch(k + 6, nq);
    const T st = ((CPtr<T>)Ap->stretch)[k];
    T stl = 0;
    if constexpr (sizeof(T) == 4) stl = ((CPtr<T>)Ap->stretch_lo)[k];
    T d[6];
    if (k < nz - 5) {
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        T acc = 0, acl = 0;
#pragma unroll
        for (int m = 1; m <= 5; ++m) {
          const T dd = q[f][5 + m] - q[f][5 - m];
          acc += Coef<T>::c(m - 1) * dd;
          if constexpr (sizeof(T) == 4) acl += Coef<T>::cl(m - 1) * dd;
        }
        if constexpr (sizeof(T) == 4) acc += acl;
        d[f] = scl(st, stl, acc);
      }
    } else {
      // top block: mirrored one-sided rows over planes nz-1 .. nz-11, all
      // in the window (see fetch): row m of plane nz-1-r sits in slot
      // (r + 16 - m) mod 11
      const int r = nz - 1 - k;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        RowSum<T> acc;
#pragma unroll
        for (int m = 0; m < 11; ++m) acc.add(q[f][m], r, (r + 16 - m) % 11);
        d[f] = -scl(st, stl, acc.sum());
      }
    }
    if (active) {
      const uint32_t g = ij + (uint32_t)k * nxy;
#pragma unroll
      for (int fz = ZBOT; fz <= ZTOP; ++fz) {
        const auto& F = Ap->face[fz];
        const int kp = k - F.off;
        if (!F.active || kp < 0 || kp >= F.n) continue;
        const uint32_t fp = ij + nxy * (uint32_t)kp;
        constexpr int fam[6] = {0, 1, 2, 3, 4, 5};
        pml_face<kStream>(F, fp, d, fam, first, last, bnext, dt);  // z-face CPML: used once
      }
#pragma unroll
      for (int f = 0; f < 6; ++f) stv<kStream>(Buf<T>(Ap->dz[f], fs), g, d[f]);
    }
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
      for (int m = 0; m < 10; ++m) q[f][m] = q[f][m + 1];
  }
}

template <typename T>
__global__ __launch_bounds__(ZP_NT, 2) void k_zpass(const StageArgs<T> A) {
  const int nblk = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3, base = nblk >> 3, rem = nblk & 7;
  const int logical = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const CArgs<T> Ap = (CArgs<T>)__builtin_amdgcn_kernarg_segment_ptr();
  if (halted(Ap->halt)) return;  // the run's failure guard stopped it
  zpass_item<T>(Ap, logical);
}

// Work distribution of the persistent batch kernels: the item list is cut
// into eight contiguous ranges, one per XCD (blocks with blockIdx % 8 == x,
// which the hardware places on XCD x), and the blocks of a range take its
// items in order from an atomic counter -- work[x] -- so a block that drew
// cheap items (no CPML face, no edge closure) takes more of them.  The last
// block to leave zeroes the counters for the next launch (work[8] counts the
// blocks that left), so no memset is needed between launches or graph
// replays.  Returns the next item of this block's range, or -1.
__device__ __forceinline__ int next_item(int* work, int lo, int hi, int* slot) {
  __syncthreads();  // every thread has read the previous item
  if (threadIdx.x == 0) *slot = lo + atomicAdd(&work[blockIdx.x & 7], 1);
  __syncthreads();
  const int item = *slot;
  return item < hi ? item : -1;
}

__device__ __forceinline__ void leave(int* work) {
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&work[8], 1) == (int)gridDim.x - 1) {
      for (int x = 0; x < 9; ++x) atomicExch(&work[x], 0);
      __threadfence();
    }
  }
}

template <typename T>
__global__ __launch_bounds__(ZP_NT, 2) void k_zpass_batch(const StageArgs<T>* __restrict__ runs,
                                                           const int* __restrict__ item_start, int nrun,
                                                           int* __restrict__ work) {
  __shared__ int slot;
  const int total = item_start[nrun];
  const int xcd = blockIdx.x & 7;
  const int lo = (int)((long)total * xcd / 8), hi = (int)((long)total * (xcd + 1) / 8);
  // a block's items increase, so its run index only moves forward
  int r = 0;
  for (int item = next_item(work, lo, hi, &slot); item >= 0; item = next_item(work, lo, hi, &slot)) {
    while (r + 1 < nrun && item_start[r + 1] <= item) ++r;
    const CArgs<T> Ap = (CArgs<T>)(runs + r);
    if (!halted(Ap->halt)) zpass_item<T>(Ap, item - item_start[r]);
  }
  leave(work);
}

// Single run: one block per work item.  The hardware hands consecutive block
// ids to the 8 XCDs in turn; remap so each XCD gets a contiguous run of tiles
// and the halos the tiles share stay in that XCD's L2.
template <typename T, bool MHD, bool BG3D, int TX, int TY, bool ZP, int VW>
__global__ __launch_bounds__(TX* TY, (stage_waves<T, MHD, BG3D, ZP, VW>())) void k_stage(const StageArgs<T> A) {
  const int nblk = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3, base = nblk >> 3, rem = nblk & 7;
  const int logical = xcd * base + (xcd < rem ? xcd : rem) + loc;
  // A sits at offset 0 of the kernel-argument segment
  const CArgs<T> Ap = (CArgs<T>)__builtin_amdgcn_kernarg_segment_ptr();
  if (halted(Ap->halt)) return;  // the run's failure guard stopped it
  stage_item<T, MHD, BG3D, TX, TY, ZP, VW>(Ap, logical);
}

// The MHD xy kernel with direct-to-LDS staging (16x16 tiles of 8-byte
// lanes: FP64, or FP32 with two points per lane; nx even, in FP32 a
// multiple of 4): the staged planes take no registers and a tenth of the
// load instructions (one 16-byte load per lane moves two cells); 2 waves per
// SIMD (the point values and derivatives need ~230 registers; held to 3
// waves the FP64 kernel spilled and ran 16% slower).  256^3 FP64 MHD xy
// kernel 1.787 -> 1.702 ms.
template <typename T>
constexpr int dma_vw() {
  return sizeof(T) == 4 ? 2 : 1;
}
template <typename T>
__global__ __launch_bounds__(256, 2) void k_stage_dma(const StageArgs<T> A) {
  const int nblk = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3, base = nblk >> 3, rem = nblk & 7;
  const int logical = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const CArgs<T> Ap = (CArgs<T>)__builtin_amdgcn_kernarg_segment_ptr();
  if (halted(Ap->halt)) return;
  stage_item<T, true, true, 16, 16, true, dma_vw<T>(), true>(Ap, logical);
}

// The same, persistent over the work items of a batch group.
template <typename T>
__global__ __launch_bounds__(256, 2) void k_stage_batch_dma(const StageArgs<T>* __restrict__ runs,
                                                             const int* __restrict__ item_start,
                                                             int nrun, int* __restrict__ work) {
  __shared__ int slot;
  const int total = item_start[nrun];
  const int xcd = blockIdx.x & 7;
  const int lo = (int)((long)total * xcd / 8), hi = (int)((long)total * (xcd + 1) / 8);
  int r = 0;
  for (int item = next_item(work, lo, hi, &slot); item >= 0; item = next_item(work, lo, hi, &slot)) {
    while (r + 1 < nrun && item_start[r + 1] <= item) ++r;
    const CArgs<T> Ap = (CArgs<T>)(runs + r);
    if (!halted(Ap->halt)) stage_item<T, true, true, 16, 16, true, dma_vw<T>(), true>(Ap, item - item_start[r]);
  }
  leave(work);
}

// Many runs in one persistent launch: the work items of all runs form one
// list (item_start[r] = first item of run r, item_start[nrun] = total),
// handed out per XCD range by next_item.  The grid size must be a multiple
// of 8.
template <typename T, bool MHD, bool BG3D, int TX, int TY, bool ZP, int VW>
__global__ __launch_bounds__(TX* TY, (stage_waves<T, MHD, BG3D, ZP, VW>())) void k_stage_batch(
    const StageArgs<T>* __restrict__ runs, const int* __restrict__ item_start, int nrun,
    int* __restrict__ work) {
  __shared__ int slot;
  const int total = item_start[nrun];
  const int xcd = blockIdx.x & 7;
  const int lo = (int)((long)total * xcd / 8), hi = (int)((long)total * (xcd + 1) / 8);
  // a block's items increase, so its run index only moves forward
  int r = 0;
  for (int item = next_item(work, lo, hi, &slot); item >= 0; item = next_item(work, lo, hi, &slot)) {
    while (r + 1 < nrun && item_start[r + 1] <= item) ++r;
    // no kernel writes the argument array while the batch runs; an item of
    // a run its failure guard stopped is skipped (uniform over the block)
    const CArgs<T> Ap = (CArgs<T>)(runs + r);
    if (!halted(Ap->halt)) stage_item<T, MHD, BG3D, TX, TY, ZP, VW>(Ap, item - item_start[r]);
  }
  leave(work);
}

// ---------------------------------------------------------------------------
// Host launchers (explicit instantiations over the supported tile shapes).
// Variants: MHD xy kernel (z pass), fused MHD, quiet with 3-D backgrounds,
// quiet with z profiles; two-point lanes for FP32 quiet and MHD xy only.
// ---------------------------------------------------------------------------
template <typename T, bool MHD, bool BG3D, int TX, int TY, bool ZP, int VW>
static hipError_t launch_one(const StageArgs<T>& A, hipStream_t s) {
  const int nblk = A.nblk_xy * A.nchunk;
  hipLaunchKernelGGL((k_stage<T, MHD, BG3D, TX, TY, ZP, VW>), dim3(nblk), dim3(TX * TY), 0, s, A);
  return hipGetLastError();
}

template <typename T, bool MHD, bool BG3D, bool ZP, int VW>
static hipError_t launch_tile(int tx, int ty, const StageArgs<T>& A, hipStream_t s) {
  if (tx == 64 && ty == 4) return launch_one<T, MHD, BG3D, 64, 4, ZP, VW>(A, s);
  if (tx == 64 && ty == 8) return launch_one<T, MHD, BG3D, 64, 8, ZP, VW>(A, s);
  if (tx == 32 && ty == 8) return launch_one<T, MHD, BG3D, 32, 8, ZP, VW>(A, s);
  if (tx == 32 && ty == 16) return launch_one<T, MHD, BG3D, 32, 16, ZP, VW>(A, s);
  if (tx == 16 && ty == 16) return launch_one<T, MHD, BG3D, 16, 16, ZP, VW>(A, s);
  return hipErrorInvalidValue;
}

template <typename T, bool MHD, bool BG3D, int TX, int TY, bool ZP, int VW>
static hipError_t launch_batch_one(const StageArgs<T>* runs, const int* item_start, int nrun,
                                   int nblk, int* work, hipStream_t s) {
  hipLaunchKernelGGL((k_stage_batch<T, MHD, BG3D, TX, TY, ZP, VW>), dim3(nblk), dim3(TX * TY), 0, s,
                     runs, item_start, nrun, work);
  return hipGetLastError();
}

template <typename T, bool MHD, bool BG3D, bool ZP, int VW>
static hipError_t launch_batch_tile(int tx, int ty, const StageArgs<T>* runs, const int* item_start,
                                    int nrun, int nblk, int* work, hipStream_t s) {
#define SPARC_BL(X, Y)                                                                               \
  if (tx == X && ty == Y)                                                                            \
    return launch_batch_one<T, MHD, BG3D, X, Y, ZP, VW>(runs, item_start, nrun, nblk, work, s);
  SPARC_BL(64, 4)
  SPARC_BL(64, 8)
  SPARC_BL(32, 8)
  SPARC_BL(32, 16)
  SPARC_BL(16, 16)
#undef SPARC_BL
  return hipErrorInvalidValue;
}

// Two-point lanes exist for FP32 only; asked of FP64 (or of the fused MHD
// kernel) the launch fails.
template <typename T>
static bool lanes_ok(bool mhd, bool zp, int vw) {
  return vw == 1 || (vw == 2 && sizeof(T) == 4 && (zp || !mhd));
}

template <typename T>
hipError_t launch_stage_batch(bool mhd, bool bg3d, bool zp, int vw, int tx, int ty,
                              const StageArgs<T>* runs, const int* item_start, int nrun, int nblk,
                              int* work, hipStream_t s, bool dma) {
  if (!lanes_ok<T>(mhd, zp, vw) || nblk % 8 != 0) return hipErrorInvalidValue;
  if (dma) {
    if (!(mhd && zp && vw == dma_vw<T>() && tx == 16 && ty == 16)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_stage_batch_dma<T>), dim3(nblk), dim3(256), 0, s, runs, item_start, nrun, work);
    return hipGetLastError();
  }
#define SPARC_BT(M, B, Z, V) launch_batch_tile<T, M, B, Z, V>(tx, ty, runs, item_start, nrun, nblk, work, s)
  if constexpr (sizeof(T) == 4) {
    if (vw == 2) {
      if (mhd) return SPARC_BT(true, true, true, 2);
      if (bg3d) return SPARC_BT(false, true, false, 2);
      return SPARC_BT(false, false, false, 2);
    }
  }
  if (mhd && zp) return SPARC_BT(true, true, true, 1);
  if (mhd) return SPARC_BT(true, true, false, 1);
  if (bg3d) return SPARC_BT(false, true, false, 1);
  return SPARC_BT(false, false, false, 1);
#undef SPARC_BT
}

template hipError_t launch_stage_batch<double>(bool, bool, bool, int, int, int, const StageArgs<double>*,
                                               const int*, int, int, int*, hipStream_t, bool);
template hipError_t launch_stage_batch<float>(bool, bool, bool, int, int, int, const StageArgs<float>*,
                                              const int*, int, int, int*, hipStream_t, bool);

// Persistent grid for a batch launch: the blocks of this kernel instance that
// are resident at once (the runtime's occupancy for its registers and LDS),
// rounded up to a multiple of 8 (one share per XCD).
template <typename K>
static int resident_grid(K kernel, int threads, int ncu) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int n = ncu * per_cu;
  return (n + 7) / 8 * 8;
}

template <typename T, bool MHD, bool BG3D, bool ZP, int VW>
static int batch_blocks_tile(int tx, int ty, int ncu) {
#define SPARC_OCC(X, Y) \
  if (tx == X && ty == Y) return resident_grid(k_stage_batch<T, MHD, BG3D, X, Y, ZP, VW>, X * Y, ncu);
  SPARC_OCC(64, 4)
  SPARC_OCC(64, 8)
  SPARC_OCC(32, 8)
  SPARC_OCC(32, 16)
  SPARC_OCC(16, 16)
#undef SPARC_OCC
  return ncu;
}

template <typename T>
static int batch_blocks_t(bool mhd, bool bg3d, bool zp, int vw, int tx, int ty, int ncu) {
  if constexpr (sizeof(T) == 4) {
    if (vw == 2) {
      if (mhd) return batch_blocks_tile<T, true, true, true, 2>(tx, ty, ncu);
      if (bg3d) return batch_blocks_tile<T, false, true, false, 2>(tx, ty, ncu);
      return batch_blocks_tile<T, false, false, false, 2>(tx, ty, ncu);
    }
  }
  if (mhd && zp) return batch_blocks_tile<T, true, true, true, 1>(tx, ty, ncu);
  if (mhd) return batch_blocks_tile<T, true, true, false, 1>(tx, ty, ncu);
  if (bg3d) return batch_blocks_tile<T, false, true, false, 1>(tx, ty, ncu);
  return batch_blocks_tile<T, false, false, false, 1>(tx, ty, ncu);
}

int stage_batch_blocks(bool f32, bool mhd, bool bg3d, bool zp, int vw, int tx, int ty, int ncu, bool dma) {
  if (dma)
    return f32 ? resident_grid(k_stage_batch_dma<float>, 256, ncu) : resident_grid(k_stage_batch_dma<double>, 256, ncu);
  return f32 ? batch_blocks_t<float>(mhd, bg3d, zp, vw, tx, ty, ncu)
             : batch_blocks_t<double>(mhd, bg3d, zp, vw, tx, ty, ncu);
}

template <typename T>
hipError_t launch_stage(bool mhd, bool bg3d, bool zp, int vw, int tx, int ty, const StageArgs<T>& A,
                        hipStream_t s) {
  if (!lanes_ok<T>(mhd, zp, vw)) return hipErrorInvalidValue;
  if (A.dma) {
    if (!(mhd && zp && vw == dma_vw<T>() && tx == 16 && ty == 16 && A.nx % (2 * vw) == 0))
      return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_stage_dma<T>), dim3(A.nblk_xy * A.nchunk), dim3(256), 0, s, A);
    return hipGetLastError();
  }
  if constexpr (sizeof(T) == 4) {
    if (vw == 2) {
      if (mhd) return launch_tile<T, true, true, true, 2>(tx, ty, A, s);
      if (bg3d) return launch_tile<T, false, true, false, 2>(tx, ty, A, s);
      return launch_tile<T, false, false, false, 2>(tx, ty, A, s);
    }
  }
  if (mhd && zp) return launch_tile<T, true, true, true, 1>(tx, ty, A, s);
  if (mhd) return launch_tile<T, true, true, false, 1>(tx, ty, A, s);
  if (bg3d) return launch_tile<T, false, true, false, 1>(tx, ty, A, s);
  return launch_tile<T, false, false, false, 1>(tx, ty, A, s);
}

template hipError_t launch_stage<double>(bool, bool, bool, int, int, int, const StageArgs<double>&,
                                         hipStream_t);
template hipError_t launch_stage<float>(bool, bool, bool, int, int, int, const StageArgs<float>&,
                                        hipStream_t);

bool stage_tile_supported(int tx, int ty) {
  return (tx == 64 && (ty == 4 || ty == 8)) || (tx == 32 && (ty == 8 || ty == 16)) ||
         (tx == 16 && ty == 16);
}

// z pass: work items per run (column blocks x z chunks), single-run and
// batched launches, and the batch grid (two 256-thread blocks per CU).
long zpass_items(int nx, int ny, int nchunk) {
  const long nxy = (long)nx * ny;
  return (nxy + ZP_NT - 1) / ZP_NT * nchunk;
}

template <typename T>
hipError_t launch_zpass(const StageArgs<T>& A, hipStream_t s) {
  const long nblk = zpass_items(A.nx, A.ny, A.znchunk);
  hipLaunchKernelGGL((k_zpass<T>), dim3((unsigned)nblk), dim3(ZP_NT), 0, s, A);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_zpass_batch(const StageArgs<T>* runs, const int* item_start, int nrun, int nblk,
                              int* work, hipStream_t s) {
  if (nblk % 8 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_zpass_batch<T>), dim3(nblk), dim3(ZP_NT), 0, s, runs, item_start, nrun, work);
  return hipGetLastError();
}

int zpass_batch_blocks(bool f32, int ncu) {
  return f32 ? resident_grid(k_zpass_batch<float>, ZP_NT, ncu)
             : resident_grid(k_zpass_batch<double>, ZP_NT, ncu);
}

template hipError_t launch_zpass<double>(const StageArgs<double>&, hipStream_t);
template hipError_t launch_zpass<float>(const StageArgs<float>&, hipStream_t);
template hipError_t launch_zpass_batch<double>(const StageArgs<double>*, const int*, int, int, int*,
                                               hipStream_t);
template hipError_t launch_zpass_batch<float>(const StageArgs<float>*, const int*, int, int, int*,
                                              hipStream_t);

}  // namespace sparc
