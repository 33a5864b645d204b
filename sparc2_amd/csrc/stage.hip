// stage.hip -- one fused low-storage RK stage of the SPARC2 right-hand side.
//
// One launch computes, for every grid point,
//   u   = a + beta_s K_{s-1}            (stage value, step.f90:27-37)
//   K_s = RHS(u)                        (MP_QUIET_3D RHS.f90:18-175 or
//                                        MP_MHD_3D RHS.f90:363-668)
// including the 11-point derivatives (derivatives.f90 ddx/ddy/ddz), the CPML
// auxiliary updates on all six faces, the pulse source and the zeroing of the
// boundary planes.  The last stage writes a + dt K_5 (step.f90:58-61) instead
// of K_5, advances the CPML bases in place and leaves per-block diagnostic
// partials (driver.f90:161,180).
//
// Blocking (2.5-D): a block owns a TX x TY column tile and marches up a chunk
// of z planes.  Fields differentiated in z live in an 11-deep register queue
// per thread (each plane is loaded once); fields differentiated in x or y are
// staged plane by plane through LDS with 5-cell halos (periodic wrap at the
// load).  Boundary rows use the reference's one-sided closures.
#include "sparc_dev.h"

namespace sparc {

enum Kind { K_RHO = 0, K_VX, K_VY, K_VZ, K_P, K_BX, K_BY, K_BZ, K_F1, K_F2, K_F3 };

template <typename T, bool MHD, bool BG3D, int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_stage(const StageArgs<T> A) {
  constexpr int NT = TX * TY;
  constexpr int LX = TX + 10, LY = TY + 10;
  constexpr int NQ = MHD ? 6 : 2;
  __shared__ T L[3][LY][LX];
  __shared__ double rs[3][NT / 64 + 1];

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const long fs = A.fs;
  const bool first = A.first, last = A.last, periodic = A.periodic;
  const T beta = A.beta;

  // XCD-aware block order: hardware hands consecutive block ids to the 8
  // XCDs in turn; remap so each XCD gets a contiguous run of tiles (shared
  // halos stay in that XCD's L2).
  const int nblk = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3, base = nblk >> 3, rem = nblk & 7;
  const int logical = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const int tile = logical % A.nblk_xy, chunk = logical / A.nblk_xy;
  const int i0 = (tile % A.ntx) * TX, j0 = (tile / A.ntx) * TY;
  const int k0 = chunk * A.kc;
  const int k1 = min(k0 + A.kc, nz);
  const int i = i0 + tx, j = j0 + ty;
  const bool active = (i < nx) && (j < ny);

  const T* __restrict__ a = A.a;
  const T* __restrict__ kin = A.kin;

  auto U = [&](int f, long g) -> T {
    T v = a[f * fs + g];
    if (!first) v = v + beta * kin[f * fs + g];
    return v;
  };
  auto val = [&](int kind, long g) -> T {
    if constexpr (MHD) {
      if (kind == K_F1) return -U(VZ, g) * A.boy[g] + U(VY, g) * A.boz[g];
      if (kind == K_F2) return -U(VX, g) * A.boz[g] + U(VZ, g) * A.box[g];
      if (kind == K_F3) return -U(VY, g) * A.box[g] + U(VX, g) * A.boy[g];
    }
    return U(kind, g);
  };
  // Global index of (gi, gj, k) with periodic wrap; -1 if outside a
  // non-periodic domain (such cells are never read by a derivative).
  auto gidx = [&](int gi, int gj, int k) -> long {
    if (periodic) {
      gi = wrap(gi, nx);
      gj = wrap(gj, ny);
    } else if (gi < 0 || gi >= nx || gj < 0 || gj >= ny) {
      return -1;
    }
    return (long)gi + (long)nx * ((long)gj + (long)ny * k);
  };
  auto ld = [&](int kind, int gi, int gj, int k) -> T {
    long g = gidx(gi, gj, k);
    return g < 0 ? T(0) : val(kind, g);
  };
  // LDS staging of one plane of one field: X = x halo, Y = y halo, XY = both.
  auto stage_x = [&](int bf, int kind, int k) {
    for (int l = tid; l < TY * LX; l += NT) {
      int r = l / LX, c = l % LX;
      L[bf][r + 5][c] = ld(kind, i0 - 5 + c, j0 + r, k);
    }
  };
  auto stage_y = [&](int bf, int kind, int k) {
    for (int l = tid; l < LY * TX; l += NT) {
      int r = l / TX, c = l % TX;
      L[bf][r][c + 5] = ld(kind, i0 + c, j0 - 5 + r, k);
    }
  };
  auto stage_xy = [&](int bf, int kind, int k) {
    stage_x(bf, kind, k);
    for (int l = tid; l < 10 * TX; l += NT) {
      int r = l / TX, c = l % TX;
      int rr = r < 5 ? r : r + TY;
      L[bf][rr][c + 5] = ld(kind, i0 + c, j0 - 5 + rr, k);
    }
  };
  // 11-point derivative along x / y from LDS (derivatives.f90:38-44,66-75,92-101)
  auto dx_lds = [&](int bf) -> T {
    const T* row = &L[bf][ty + 5][0];
    T acc = 0;
    if (!periodic && i < 5) {
#pragma unroll
      for (int m = 0; m < 11; ++m) acc += row[min(m - i0 + 5, LX - 1)] * Coef<T>::b(i, m);
      return A.sxs * acc;
    }
    if (!periodic && i >= nx - 5) {
      const int r = nx - 1 - i;
#pragma unroll
      for (int m = 0; m < 11; ++m) {
        int c = min(max(nx - 1 - m - i0 + 5, 0), LX - 1);
        acc += row[c] * Coef<T>::b(r, m);
      }
      return -A.sxs * acc;
    }
    const int c0 = tx + 5;
#pragma unroll
    for (int m = 1; m <= 5; ++m) acc += Coef<T>::c(m - 1) * (row[c0 + m] - row[c0 - m]);
    return A.sxs * acc;
  };
  auto dy_lds = [&](int bf) -> T {
    const int c0 = tx + 5;
    T acc = 0;
    if (!periodic && j < 5) {
#pragma unroll
      for (int m = 0; m < 11; ++m) acc += L[bf][min(max(m - j0 + 5, 0), LY - 1)][c0] * Coef<T>::b(j, m);
      return A.sys * acc;
    }
    if (!periodic && j >= ny - 5) {
      const int r = ny - 1 - j;
#pragma unroll
      for (int m = 0; m < 11; ++m) {
        int rr = min(max(ny - 1 - m - j0 + 5, 0), LY - 1);
        acc += L[bf][rr][c0] * Coef<T>::b(r, m);
      }
      return -A.sys * acc;
    }
    const int r0 = ty + 5;
#pragma unroll
    for (int m = 1; m <= 5; ++m) acc += Coef<T>::c(m - 1) * (L[bf][r0 + m][c0] - L[bf][r0 - m][c0]);
    return A.sys * acc;
  };

  // z register queue: slot m holds plane k-5+m of each z-differentiated field.
  // Quiet: vz, p.  MHD: vz, p, bx, by, flux1, flux2.
  T q[NQ][11];
  auto load_plane = [&](int kk, int slot) {
    if (active && kk >= 0 && kk < nz) {
      const long g = (long)i + (long)nx * ((long)j + (long)ny * kk);
      const T vz = U(VZ, g), p = U(P, g);
      q[0][slot] = vz;
      q[1][slot] = p;
      if constexpr (MHD) {
        const T vx = U(VX, g), vy = U(VY, g);
        q[2][slot] = U(BX, g);
        q[3][slot] = U(BY, g);
        q[4][slot] = -vz * A.boy[g] + vy * A.boz[g];
        q[5][slot] = -vx * A.boz[g] + vz * A.box[g];
      }
    } else {
#pragma unroll
      for (int f = 0; f < NQ; ++f) q[f][slot] = 0;  // zero ghosts below the bottom
    }
  };
#pragma unroll
  for (int m = 0; m < 10; ++m) load_plane(k0 - 5 + m, m);

  constexpr int qkind[6] = {K_VZ, K_P, K_BX, K_BY, K_F1, K_F2};
  double ke = 0.0, mr = 1.0e300, mp = 1.0e300;

  for (int k = k0; k < k1; ++k) {
    load_plane(k + 5, 10);

    // ---- z derivatives (derivatives.f90:481-487,509-518,536-545) ----
    T dz[NQ];
    if (active) {
      const T st = A.stretch[k];
      if (k < nz - 5) {
#pragma unroll
        for (int f = 0; f < NQ; ++f) {
          T acc = 0;
#pragma unroll
          for (int m = 1; m <= 5; ++m) acc += Coef<T>::c(m - 1) * (q[f][5 + m] - q[f][5 - m]);
          dz[f] = st * acc;
        }
      } else {
        const int r = nz - 1 - k;
#pragma unroll
        for (int f = 0; f < NQ; ++f) {
          T acc = 0;
          for (int m = 0; m < 11; ++m) {
            const long g = (long)i + (long)nx * ((long)j + (long)ny * (nz - 1 - m));
            acc += val(qkind[f], g) * Coef<T>::b(r, m);
          }
          dz[f] = -st * acc;
        }
      }
    }

    // ---- x/y derivatives through LDS ----
    T dvxdx, dvydy, gpx, gpy, vx, vy;
    T dxf3, dyf3, dxf2, dyf1, dxby, dxbz, dybz, dybx, bz;
    stage_x(0, K_VX, k);
    stage_y(1, K_VY, k);
    stage_xy(2, K_P, k);
    __syncthreads();
    if (active) {
      dvxdx = dx_lds(0);
      dvydy = dy_lds(1);
      gpx = dx_lds(2);
      gpy = dy_lds(2);
      vx = L[0][ty + 5][tx + 5];
      vy = L[1][ty + 5][tx + 5];
    }
    __syncthreads();
    if constexpr (MHD) {
      stage_xy(0, K_F3, k);
      stage_x(1, K_F2, k);
      stage_y(2, K_F1, k);
      __syncthreads();
      if (active) {
        dxf3 = dx_lds(0);
        dyf3 = dy_lds(0);
        dxf2 = dx_lds(1);
        dyf1 = dy_lds(2);
      }
      __syncthreads();
      stage_x(0, K_BY, k);
      stage_xy(1, K_BZ, k);
      stage_y(2, K_BX, k);
      __syncthreads();
      if (active) {
        dxby = dx_lds(0);
        dxbz = dx_lds(1);
        dybz = dy_lds(1);
        dybx = dy_lds(2);
        bz = L[1][ty + 5][tx + 5];
      }
      __syncthreads();
    }

    if (active) {
      const long g = (long)i + (long)nx * ((long)j + (long)ny * k);
      const long gb = BG3D ? g : (long)k;
      const T rho = U(RHO, g);
      const T vz = q[0][5], p = q[1][5];
      T dvzdz = dz[0], gpz = dz[1];
      T dzbx = 0, dzby = 0, dzf1 = 0, dzf2 = 0;
      if constexpr (MHD) {
        dzbx = dz[2];
        dzby = dz[3];
        dzf1 = dz[4];
        dzf2 = dz[5];
      }
      // ---- CPML (RHS.f90:50-93, 424-546) ----
      const T dt = A.dt;
      for (int fz = ZBOT; fz <= ZTOP; ++fz) {
        const Face<T>& F = A.face[fz];
        if (!F.active) continue;
        const int kp = k - F.off;
        if (kp < 0 || kp >= F.n) continue;
        const long fp = (long)i + (long)nx * ((long)j + (long)ny * kp);
        pml_family(F, fp, 1, gpz, first, last, beta, dt);
        pml_family(F, fp, 0, dvzdz, first, last, beta, dt);
        if constexpr (MHD) {
          pml_family(F, fp, 4, dzf2, first, last, beta, dt);
          pml_family(F, fp, 5, dzf1, first, last, beta, dt);
          pml_family(F, fp, 2, dzbx, first, last, beta, dt);
          pml_family(F, fp, 3, dzby, first, last, beta, dt);
        }
      }
      for (int fx = XBOT; fx <= XTOP; ++fx) {
        const Face<T>& F = A.face[fx];
        if (!F.active) continue;
        const int ip = i - F.off;
        if (ip < 0 || ip >= F.n) continue;
        const long fp = (long)ip + (long)F.n * ((long)j + (long)ny * k);
        pml_family(F, fp, 1, gpx, first, last, beta, dt);
        pml_family(F, fp, 0, dvxdx, first, last, beta, dt);
        if constexpr (MHD) {
          pml_family(F, fp, 4, dxf3, first, last, beta, dt);
          pml_family(F, fp, 5, dxf2, first, last, beta, dt);
          pml_family(F, fp, 2, dxby, first, last, beta, dt);
          pml_family(F, fp, 3, dxbz, first, last, beta, dt);
        }
      }
      for (int fy = YBOT; fy <= YTOP; ++fy) {
        const Face<T>& F = A.face[fy];
        if (!F.active) continue;
        const int jp = j - F.off;
        if (jp < 0 || jp >= F.n) continue;
        const long fp = (long)i + (long)nx * ((long)jp + (long)F.n * k);
        pml_family(F, fp, 1, gpy, first, last, beta, dt);
        pml_family(F, fp, 0, dvydy, first, last, beta, dt);
        if constexpr (MHD) {
          pml_family(F, fp, 4, dyf3, first, last, beta, dt);
          pml_family(F, fp, 5, dyf1, first, last, beta, dt);
          pml_family(F, fp, 2, dybx, first, last, beta, dt);
          pml_family(F, fp, 3, dybz, first, last, beta, dt);
        }
      }

      // ---- pointwise RHS ----
      const T rho0 = A.rho0[gb], rhoinv = A.rhoinv[gb], sp = A.sponge[gb];
      const T div = dvxdx + dvydy + dvzdz;
      T K[8];
      if constexpr (!MHD) {
        // RHS.f90:95-101,136,152
        K[RHO] = -A.gr0z[gb] * vz - rho0 * div - sp * rho;
        K[VX] = -rhoinv * gpx - sp * vx;
        K[VY] = -rhoinv * gpy - sp * vy;
        K[VZ] = -rhoinv * (gpz + rho * A.g[k]) - vz * sp;
        K[P] = -A.c2rho0[gb] * div - vz * A.gp0z[gb] - sp * p;
      } else {
        const T red = A.red[g];
        const T b0x = A.box[g], b0y = A.boy[g], b0z = A.boz[g];
        const T bx = q[2][5], by = q[3][5];
        // RHS.f90:548-580
        K[BX] = red * (dyf3 - dzf2);
        K[BY] = red * (dzf1 - dxf3);
        K[BZ] = red * (dxf2 - dyf1);
        const T cbx = dybz - dzby, cby = -dxbz + dzbx, cbz = dxby - dybx;
        K[RHO] = -A.gr0x[g] * vx - A.gr0y[g] * vy - A.gr0z[g] * vz - rho0 * div - sp * rho;
        K[P] = -A.c2rho0[g] * div - vx * A.gp0x[g] - vy * A.gp0y[g] - vz * A.gp0z[g] - sp * p;
        // CROSS(curl b, B0) and CROSS(curl B0, b), derivatives.f90:646-655
        const T c1 = cby * b0z - cbz * b0y, c2 = b0x * cbz - cbx * b0z, c3 = cbx * b0y - cby * b0x;
        const T e1 = A.cboy[g] * bz - A.cboz[g] * by;
        const T e2 = bx * A.cboz[g] - A.cbox[g] * bz;
        const T e3 = A.cbox[g] * by - A.cboy[g] * bx;
        K[VX] = -rhoinv * (gpx - c1 * red) - sp * vx + rhoinv * e1 * red;
        K[VY] = -rhoinv * (gpy - c2 * red) - sp * vy + rhoinv * e2 * red;
        K[VZ] = -rhoinv * (gpz - c3 * red) - sp * vz + rhoinv * e3 * red - rhoinv * rho * A.g[k];
      }
      // pulse source (driver.f90:137-148 separable form; RHS.f90:138-145,632-639)
      if (A.pulse_dir >= 1 && A.pulse_dir <= 3) {
        const T s = A.src_amp * A.sx[i] * A.sy[j] * A.sz[k];
        if (A.pulse_dir == 1) K[VX] += s;
        else if (A.pulse_dir == 2) K[VY] += s;
        else K[VZ] += s;
      }
      // boundary-plane zeroing (RHS.f90:154-164,647-657)
      constexpr int NV = MHD ? 8 : 5;
      const bool zero = (k == 0) || (k == nz - 1) ||
                        (!periodic && (i == 0 || i == nx - 1 || j == 0 || j == ny - 1));
      if (zero) {
#pragma unroll
        for (int f = 0; f < NV; ++f) K[f] = 0;
      }
      if (!last) {
#pragma unroll
        for (int f = 0; f < NV; ++f) A.kout[f * fs + g] = K[f];
      } else {
        T an[NV];
#pragma unroll
        for (int f = 0; f < NV; ++f) {
          const T a0 = a[f * fs + g];
          if constexpr (sizeof(T) == 4) {
            // compensated update: the running sum a - c carries the lost bits
            const T c0 = A.comp[f * fs + g];
            const T y = A.dt * K[f] - c0;
            const T t = a0 + y;
            A.comp[f * fs + g] = (t - a0) - y;
            an[f] = t;
          } else {
            an[f] = a0 + A.dt * K[f];
          }
          A.aout[f * fs + g] = an[f];
        }
        // diagnostics: KE norm on the new state, minima on the stage-5 state
        const double m = ((double)rho0 + (double)an[RHO]) *
                         ((double)an[VX] * an[VX] + (double)an[VY] * an[VY] + (double)an[VZ] * an[VZ]) /
                         2.0;
        ke += m * m;
        mr = fmin(mr, (double)rho + (double)rho0);
        mp = fmin(mp, (double)p + (double)A.p0[gb]);
      }
    }

#pragma unroll
    for (int f = 0; f < NQ; ++f)
#pragma unroll
      for (int m = 0; m < 10; ++m) q[f][m] = q[f][m + 1];
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
      A.partial[3 * (long)logical + 0] = s;
      A.partial[3 * (long)logical + 1] = r;
      A.partial[3 * (long)logical + 2] = pp;
    }
  }
}

// ---------------------------------------------------------------------------
// Host launchers (explicit instantiations over the supported tile shapes).
// ---------------------------------------------------------------------------
template <typename T, bool MHD, bool BG3D, int TX, int TY>
static hipError_t launch_one(const StageArgs<T>& A, hipStream_t s) {
  const int nblk = A.nblk_xy * A.nchunk;
  hipLaunchKernelGGL((k_stage<T, MHD, BG3D, TX, TY>), dim3(nblk), dim3(TX * TY), 0, s, A);
  return hipGetLastError();
}

template <typename T, bool MHD, bool BG3D>
static hipError_t launch_tile(int tx, int ty, const StageArgs<T>& A, hipStream_t s) {
  if (tx == 64 && ty == 4) return launch_one<T, MHD, BG3D, 64, 4>(A, s);
  if (tx == 64 && ty == 8) return launch_one<T, MHD, BG3D, 64, 8>(A, s);
  if (tx == 32 && ty == 8) return launch_one<T, MHD, BG3D, 32, 8>(A, s);
  if (tx == 32 && ty == 16) return launch_one<T, MHD, BG3D, 32, 16>(A, s);
  if (tx == 16 && ty == 16) return launch_one<T, MHD, BG3D, 16, 16>(A, s);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_stage(bool mhd, bool bg3d, int tx, int ty, const StageArgs<T>& A, hipStream_t s) {
  if (mhd) return launch_tile<T, true, true>(tx, ty, A, s);
  if (bg3d) return launch_tile<T, false, true>(tx, ty, A, s);
  return launch_tile<T, false, false>(tx, ty, A, s);
}

template hipError_t launch_stage<double>(bool, bool, int, int, const StageArgs<double>&, hipStream_t);
template hipError_t launch_stage<float>(bool, bool, int, int, const StageArgs<float>&, hipStream_t);

bool stage_tile_supported(int tx, int ty) {
  return (tx == 64 && (ty == 4 || ty == 8)) || (tx == 32 && (ty == 8 || ty == 16)) ||
         (tx == 16 && ty == 16);
}

}  // namespace sparc
