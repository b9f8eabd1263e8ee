// Time-domain Yee updates (FDTD) on real fields.
//
// maxwell_h: h -= dt curl_forward(e) / mu   (meanas/fdtd/base.py:67-125)
// maxwell_e: e += dt curl_back(h) / epsilon (meanas/fdtd/base.py:13-64)
// with curl_forward / curl_back as in meanas/fdmath/functional.py: periodic forward
// differences (f[n+1] - f[n]) / dx_e[n] and backward differences (f[n] - f[n-1]) / dx_h[n],
// unit widths when the grid gives none.  Materials divide, as the reference does.
//
// One thread per cell updates its three components; the neighbours one cell away are
// re-read from the caches (k+-1 is the next lane; j+-1 and i+-1 rows and planes are in
// L2 from the neighbouring threads).  HBM-bound: per step the H update reads e and h and
// writes h, the E update reads h and e (and eps) and writes e.
#include <cstdlib>
#include <string>

#include "common.hpp"
#include "../../include/meanas_fdfd.h"

namespace mx {

int set_error(const std::string& msg);

namespace {

constexpr int kFB = 256;

struct RGrid {
  int nx, ny, nz;
  const double* dxe[3];
  const double* dxh[3];
};

__device__ __forceinline__ double width(const double* w, int n) { return w ? w[n] : 1.0; }
// width n of a table with n uniform over the block: a scalar load (constant address
// space) into SGPRs
__device__ __forceinline__ double width_u(const double* w, int n) {
  return w ? ((const MX_CONST double*)w)[n] : 1.0;
}

template <typename T>
__global__ __launch_bounds__(kFB) void k_fdtd_h(RGrid g, T dt, const T* __restrict__ e, T* __restrict__ h,
                                                const T* __restrict__ mu, T mu_s) {
  const int nc = g.nx * g.ny * g.nz;
  for (int cell = blockIdx.x * kFB + threadIdx.x; cell < nc; cell += gridDim.x * kFB) {
    const int k = cell % g.nz, t = cell / g.nz, j = t % g.ny, i = t / g.ny;
    const int sx = (i + 1 == g.nx ? -i : 1) * g.ny * g.nz;   // i+1 (periodic)
    const int sy = (j + 1 == g.ny ? -j : 1) * g.nz;
    const int sz = (k + 1 == g.nz ? -k : 1);
    const T dx = (T)width(g.dxe[0], i), dy = (T)width(g.dxe[1], j), dz = (T)width(g.dxe[2], k);
    const T ex = e[cell], ey = e[nc + cell], ez = e[2 * nc + cell];
    const T cx = (e[2 * nc + cell + sy] - ez) / dy - (e[nc + cell + sz] - ey) / dz;
    const T cy = (e[cell + sz] - ex) / dz - (e[2 * nc + cell + sx] - ez) / dx;
    const T cz = (e[nc + cell + sx] - ey) / dx - (e[cell + sy] - ex) / dy;
    const T c[3] = {cx, cy, cz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T m = mu ? mu[q * nc + cell] : mu_s;
      h[q * nc + cell] -= dt * c[q] / m;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kFB) void k_fdtd_e(RGrid g, T dt, T* __restrict__ e, const T* __restrict__ h,
                                                const T* __restrict__ eps, T eps_s) {
  const int nc = g.nx * g.ny * g.nz;
  for (int cell = blockIdx.x * kFB + threadIdx.x; cell < nc; cell += gridDim.x * kFB) {
    const int k = cell % g.nz, t = cell / g.nz, j = t % g.ny, i = t / g.ny;
    const int sx = (i == 0 ? g.nx - 1 : -1) * g.ny * g.nz;   // i-1 (periodic)
    const int sy = (j == 0 ? g.ny - 1 : -1) * g.nz;
    const int sz = (k == 0 ? g.nz - 1 : -1);
    const T dx = (T)width(g.dxh[0], i), dy = (T)width(g.dxh[1], j), dz = (T)width(g.dxh[2], k);
    const T hx = h[cell], hy = h[nc + cell], hz = h[2 * nc + cell];
    const T cx = (hz - h[2 * nc + cell + sy]) / dy - (hy - h[nc + cell + sz]) / dz;
    const T cy = (hx - h[cell + sz]) / dz - (hz - h[2 * nc + cell + sx]) / dx;
    const T cz = (hy - h[nc + cell + sx]) / dx - (hx - h[cell + sy]) / dy;
    const T c[3] = {cx, cy, cz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T m = eps ? eps[q * nc + cell] : eps_s;
      e[q * nc + cell] += dt * c[q] / m;
    }
  }
}

// ---- fused leapfrog step (mx_fdtd_run) ----------------------------------------------
// One pass per step: H' = H - dt curl_f(E) / mu, then E' = E + dt curl_b(H') / eps, read
// from (e0, h0) and written to (e1, h1).  The two buffer pairs alternate between steps,
// so no block ever reads a half-updated field: 5 field streams per step instead of the
// 7 of two separate updates.  A block owns a TY x 64 (y, z) tile and marches along x
// over `slab` planes.  Per plane i:
//   - E(i) of the tile plus a one-cell halo sits in LDS; plane i+1 comes from the
//     registers of the threads that loaded it: each thread loads exactly the plane-i+1
//     values its own H computations need (own cell; wave 0 the row j0-1 whose H it forms;
//     lanes 0..TY-1 of wave TY-1 the column k0-1 whose H they form);
//   - H'(i) is formed on the tile, the -y row and the -z column (the halo cells
//     redundantly with the neighbouring blocks, by the same expressions) into LDS;
//   - E'(i) takes H'(i) from LDS and H'(i-1) of its own cell from registers.
// The expressions are those of k_fdtd_h and k_fdtd_e, so a run equals the step-by-step
// updates bit for bit.  Partial tiles wrap their row and column indices (periodic grid),
// so every load is in bounds and every neighbour read is the periodic one.
// Tile rows (waves per block) of the fused step.  At 256^3: 8 rows 429 us (float64) /
// 263 us (float32) per step, 4 rows 433 / 282 us, 2 rows 452 / 291 us (fewer re-read
// halo rows per computed row).
#ifndef MX_FDTD_TY
#define MX_FDTD_TY 8
#endif
constexpr int kStepTY = MX_FDTD_TY;
constexpr int kStepSlab = 16;

__device__ __forceinline__ int wrap(int v, int n) {
  const int r = v % n;
  return r < 0 ? r + n : r;
}

template <typename T>
__global__ __launch_bounds__(64 * kStepTY) void k_fdtd_step(RGrid g, T dt, const T* __restrict__ e0,
                                                            const T* __restrict__ h0, T* __restrict__ e1,
                                                            T* __restrict__ h1, const T* __restrict__ eps, T eps_s,
                                                            const T* __restrict__ mu, T mu_s, int slab) {
  constexpr int TY = kStepTY, TZ = 64, LY = TY + 2, LZ = TZ + 2;
  static_assert(2 * TY + 2 <= 64, "the z-halo loads take 2 TY + 2 lanes of one wave");
  __shared__ T sE[3][LY][LZ];
  __shared__ T sH[3][TY + 1][TZ + 1];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int nc = nx * ny * nz, plane = ny * nz;
  const int ntj = (ny + TY - 1) / TY, ntk = (nz + TZ - 1) / TZ, tiles = ntj * ntk;
  const int nwork = ((nx + slab - 1) / slab) * tiles;
  // workgroups are dealt round-robin to the 8 XCDs: XCD x takes work items
  // [x * per, (x + 1) * per), neighbouring tiles of one slab, so halos hit its own L2
  const int per = (nwork + 7) / 8;
  const int w = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (w >= nwork) return;
  const int slab_id = w / tiles, tile = w - slab_id * tiles;
  const int j0 = (tile / ntk) * TY, k0 = (tile % ntk) * TZ;
  const int ib = slab_id * slab, ie = min(ib + slab, nx);
  const int tz = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int j = j0 + ty, k = k0 + tz;
  const bool valid = j < ny && k < nz;
  const int jw = wrap(j, ny), kw = wrap(k, nz);

  // load sites: own cell; y-halo row (wave 0: row j0-1, wave TY-1: row j0+TY); z-halo
  // cells (wave TY-1: lanes 0..TY column k0-1 at LDS rows 1..TY+1, lanes TY+1..2TY+1
  // column k0+TZ at LDS rows 0..TY)
  const int own_off = jw * nz + kw;
  const bool has_row = ty == 0 || ty == TY - 1;
  const int row_j = wrap(ty == 0 ? j0 - 1 : j0 + TY, ny);
  const int row_off = row_j * nz + kw;
  const int row_r = ty == 0 ? 0 : TY + 1;
  const bool has_col = ty == TY - 1 && tz < 2 * TY + 2;
  const int col_r = tz < TY + 1 ? tz + 1 : tz - (TY + 1);
  const int col_q = tz < TY + 1 ? 0 : TZ + 1;
  const int col_j = wrap(j0 - 1 + col_r, ny), col_k = wrap(col_q ? k0 + TZ : k0 - 1, nz);
  const int col_off = col_j * nz + col_k;
  const bool row_h = ty == 0;                    // forms H' of row j0-1 (x, z components)
  const bool col_h = ty == TY - 1 && tz < TY;    // forms H' of column k0-1, row j0+tz (x, y)

  // widths of the cells this thread forms H' / E' at
  const T dye = (T)width(g.dxe[1], jw), dze = (T)width(g.dxe[2], kw);
  const T dyh = (T)width(g.dxh[1], jw), dzh = (T)width(g.dxh[2], kw);
  const T dye_r = (T)width(g.dxe[1], row_j);
  const T dye_c = (T)width(g.dxe[1], col_j), dze_c = (T)width(g.dxe[2], col_k);

  // What step i needs besides LDS: E(i+1) at the load sites, H(i) and mu(i) where this
  // thread forms H', eps(i) at its own cell.
  struct Bundle {
    T eo[3], er[3], ec[3];   // E(i+1): own, row, column site
    T ho[3], hr[2], hc[2];   // H(i): own; row site (x, z); column site (x, y)
    T mo[3], mr[2], mc[2];   // mu(i) at the same cells
    T ep[3];                 // eps(i), own cell
  };
  auto mat = [&](const T* m, T ms, int idx) -> T { return m ? m[idx] : ms; };
  auto fetch = [&](Bundle& B, int i) {   // plane i+1 of E, plane i of H, mu and eps
    const int pn = wrap(i + 1, nx) * plane, pi = i * plane;
#pragma unroll
    for (int c = 0; c < 3; ++c) B.eo[c] = e0[c * nc + pn + own_off];
    if (has_row) {
#pragma unroll
      for (int c = 0; c < 3; ++c) B.er[c] = e0[c * nc + pn + row_off];
    }
    if (has_col) {
#pragma unroll
      for (int c = 0; c < 3; ++c) B.ec[c] = e0[c * nc + pn + col_off];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      B.ho[c] = h0[c * nc + pi + own_off];
      B.mo[c] = mat(mu, mu_s, c * nc + pi + own_off);
      B.ep[c] = mat(eps, eps_s, c * nc + pi + own_off);
    }
    if (row_h) {
      B.hr[0] = h0[pi + row_off];
      B.hr[1] = h0[2 * nc + pi + row_off];
      B.mr[0] = mat(mu, mu_s, pi + row_off);
      B.mr[1] = mat(mu, mu_s, 2 * nc + pi + row_off);
    }
    if (col_h) {
      B.hc[0] = h0[pi + col_off];
      B.hc[1] = h0[nc + pi + col_off];
      B.mc[0] = mat(mu, mu_s, pi + col_off);
      B.mc[1] = mat(mu, mu_s, nc + pi + col_off);
    }
  };
  auto put_e = [&](int i) {   // E(i) at the load sites into LDS (plane-i values of the bundle)
    const int p = i * plane;
#pragma unroll
    for (int c = 0; c < 3; ++c) sE[c][ty + 1][tz + 1] = e0[c * nc + p + own_off];
    if (has_row) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sE[c][row_r][tz + 1] = e0[c * nc + p + row_off];
    }
    if (has_col) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sE[c][col_r][col_q] = e0[c * nc + p + col_off];
    }
  };
  auto commit = [&](const Bundle& B) {   // E(i+1) from the bundle into LDS
#pragma unroll
    for (int c = 0; c < 3; ++c) sE[c][ty + 1][tz + 1] = B.eo[c];
    if (has_row) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sE[c][row_r][tz + 1] = B.er[c];
    }
    if (has_col) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sE[c][col_r][col_q] = B.ec[c];
    }
  };

  Bundle B;
  T hpy, hpz;   // H'(i-1) y, z at the own cell
  const int r = ty + 1, q = tz + 1;
  // ---- prologue: H'(ib-1) at the own cell (y, z), plane ib into LDS
  {
    const int im1 = wrap(ib - 1, nx);
    put_e(im1);
    fetch(B, im1);   // E(ib), H(ib-1), mu(ib-1)
    __syncthreads();
    const T dx = (T)width(g.dxe[0], im1);
    const T cy = (sE[0][r][q + 1] - sE[0][r][q]) / dze - (B.eo[2] - sE[2][r][q]) / dx;
    const T cz = (B.eo[1] - sE[1][r][q]) / dx - (sE[0][r + 1][q] - sE[0][r][q]) / dye;
    hpy = B.ho[1] - dt * cy / B.mo[1];
    hpz = B.ho[2] - dt * cz / B.mo[2];
    __syncthreads();
    commit(B);       // plane ib
    fetch(B, ib);    // E(ib+1), H(ib), mu(ib), eps(ib)
  }

  // float32: the x widths of the next plane are read one step ahead with scalar loads
  // (as vector loads at the top of the step they are waited for right there, a memory
  // round trip per plane): 246 against 255-257 us per 256^3 step.  float64 keeps the
  // vector loads: ahead of time its step measured 432 against 427 us, and carried in the
  // prefetch bundle the widths took it over 128 VGPRs (a wave per SIMD less, 504 us).
  constexpr bool kAheadWidths = sizeof(T) == 4;
  double wxe_n = kAheadWidths ? width_u(g.dxe[0], ib) : 0.0, wxh_n = kAheadWidths ? width_u(g.dxh[0], ib) : 0.0;
#pragma unroll 1
  for (int i = ib; i < ie; ++i) {
    __syncthreads();   // E(i) is in LDS; the previous step's reads of sH are done
    const T dxe_i = (T)(kAheadWidths ? wxe_n : width(g.dxe[0], i));
    const T dxh_i = (T)(kAheadWidths ? wxh_n : width(g.dxh[0], i));
    const T e_own[3] = {sE[0][r][q], sE[1][r][q], sE[2][r][q]};
    const T epsv[3] = {B.ep[0], B.ep[1], B.ep[2]};
    // H'(i): own cell
    {
      const T cx = (sE[2][r + 1][q] - e_own[2]) / dye - (sE[1][r][q + 1] - e_own[1]) / dze;
      const T cy = (sE[0][r][q + 1] - e_own[0]) / dze - (B.eo[2] - e_own[2]) / dxe_i;
      const T cz = (B.eo[1] - e_own[1]) / dxe_i - (sE[0][r + 1][q] - e_own[0]) / dye;
      const T hn[3] = {B.ho[0] - dt * cx / B.mo[0], B.ho[1] - dt * cy / B.mo[1], B.ho[2] - dt * cz / B.mo[2]};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sH[c][r][q] = hn[c];
        if (valid) h1[c * nc + i * plane + own_off] = hn[c];
      }
    }
    if (row_h) {   // row j0-1 (LDS row 0): x and z
      const T cx = (sE[2][1][q] - sE[2][0][q]) / dye_r - (sE[1][0][q + 1] - sE[1][0][q]) / dze;
      const T cz = (B.er[1] - sE[1][0][q]) / dxe_i - (sE[0][1][q] - sE[0][0][q]) / dye_r;
      sH[0][0][q] = B.hr[0] - dt * cx / B.mr[0];
      sH[2][0][q] = B.hr[1] - dt * cz / B.mr[1];
    }
    if (col_h) {   // column k0-1 (LDS column 0), row j0+tz: x and y
      const int rr = tz + 1;
      const T cx = (sE[2][rr + 1][0] - sE[2][rr][0]) / dye_c - (sE[1][rr][1] - sE[1][rr][0]) / dze_c;
      const T cy = (sE[0][rr][1] - sE[0][rr][0]) / dze_c - (B.ec[2] - sE[2][rr][0]) / dxe_i;
      sH[0][rr][0] = B.hc[0] - dt * cx / B.mc[0];
      sH[1][rr][0] = B.hc[1] - dt * cy / B.mc[1];
    }
    __syncthreads();   // H'(i) is in LDS; every read of E(i) is done
    if (i + 1 < ie) {
      commit(B);          // E(i+1) replaces E(i)
      fetch(B, i + 1);    // E(i+2), H(i+1), mu(i+1), eps(i+1)
    }
    // E'(i): own cell
    {
      const T hx = sH[0][r][q], hy = sH[1][r][q], hz = sH[2][r][q];
      const T cx = (hz - sH[2][r - 1][q]) / dyh - (hy - sH[1][r][q - 1]) / dzh;
      const T cy = (hx - sH[0][r][q - 1]) / dzh - (hz - hpz) / dxh_i;
      const T cz = (hy - hpy) / dxh_i - (hx - sH[0][r - 1][q]) / dyh;
      const T c3[3] = {cx, cy, cz};
      if (valid) {
#pragma unroll
        for (int c = 0; c < 3; ++c) e1[c * nc + i * plane + own_off] = e_own[c] + dt * c3[c] / epsv[c];
      }
      hpy = hy;
      hpz = hz;
    }
    // the next plane's widths, at the end of the step: scalar loads share the wait
    // counter of LDS reads, so issued at the top they made its first LDS read wait
    if (kAheadWidths && i + 1 < ie) {
      wxe_n = width_u(g.dxe[0], i + 1);
      wxh_n = width_u(g.dxh[0], i + 1);
    }
  }
}

// ---- convolutional PML (fdtd/pml.py:96-230) ---------------------------------------
// One thread per cell.  The six derivatives d_a f_c (a != c) use dx_h along every axis.
// Each layer the cell lies in (in the table's order: axis 0..2, polarity -1 then +1)
// scales the two derivatives along its axis by p2, advances the cell's psi_u, psi_v
// (u < v the components other than the axis) with p0 and p1, stores them in the psi
// precision, and adds s psi_u to the auxiliary term of u and -s psi_v to v's, in the psi
// precision, s = (-1, 1, -1)[axis].  Precisions follow the reference's array code with
// float64 widths: differences of field values in the field precision, divided by the
// float64 width, so derivatives, psi updates and the final sum are double; dt / material
// is taken in the material's (field) precision, dt / scalar in double.
struct CpmlLayer {
  int axis, start, thickness;
  const double* pe;
  const double* ph;
  void* psie;
  void* psih;
};
struct CpmlTable {
  int n;
  CpmlLayer l[MX_CPML_MAX_LAYERS];
};

template <typename T, typename TP, bool EUPD>
__global__ __launch_bounds__(kFB) void k_cpml(RGrid g, T dt, T* __restrict__ e, T* __restrict__ h,
                                              const T* __restrict__ mat, double mat_s, double dt64, CpmlTable tab) {
  const int nc = g.nx * g.ny * g.nz;
  const int dims[3] = {g.nx, g.ny, g.nz};
  const int strides[3] = {g.ny * g.nz, g.nz, 1};
  const T* f = EUPD ? h : e;
  T* out = EUPD ? e : h;
  for (int cell = blockIdx.x * kFB + threadIdx.x; cell < nc; cell += gridDim.x * kFB) {
    const int k = cell % g.nz, t = cell / g.nz, j = t % g.ny, i = t / g.ny;
    const int idx[3] = {i, j, k};
    // d[a][c]: derivative along a of component c (E update: backward of h; H: forward of e)
    double d[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double w = width(g.dxh[a], idx[a]);
      const int s = EUPD ? (idx[a] == 0 ? (dims[a] - 1) * strides[a] : -strides[a])
                         : (idx[a] + 1 == dims[a] ? -idx[a] * strides[a] : strides[a]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (c == a) continue;
        const T here = f[c * nc + cell], there = f[c * nc + cell + s];
        d[a][c] = (double)(EUPD ? here - there : there - here) / w;
      }
    }
    TP aux[3] = {TP(0), TP(0), TP(0)};
    for (int q = 0; q < tab.n; ++q) {
      const CpmlLayer& L = tab.l[q];
      const int a = L.axis;
      const int pos = idx[a] - L.start;
      if (pos < 0 || pos >= L.thickness) continue;
      const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
      const double* p = EUPD ? L.pe : L.ph;
      const double p0 = p[pos], p1 = p[L.thickness + pos], p2 = p[2 * L.thickness + pos];
      d[a][u] *= p2;
      d[a][v] *= p2;
      // index in the layer's C-order sub-grid (its axis `thickness` cells long)
      int sub[3] = {i, j, k}, sd[3] = {g.nx, g.ny, g.nz};
      sub[a] = pos;
      sd[a] = L.thickness;
      const int r = (sub[0] * sd[1] + sub[1]) * sd[2] + sub[2];
      const int rn = sd[0] * sd[1] * sd[2];
      TP* psi = reinterpret_cast<TP*>(EUPD ? L.psie : L.psih);
      TP pu = (TP)((double)psi[r] * p0), pv = (TP)((double)psi[rn + r] * p0);
      pu = (TP)((double)pu + p1 * d[a][v]);
      pv = (TP)((double)pv + p1 * d[a][u]);
      psi[r] = pu;
      psi[rn + r] = pv;
      const TP sg = a == 1 ? TP(1) : TP(-1);
      aux[u] = aux[u] + sg * pu;
      aux[v] = aux[v] + (-sg) * pv;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      const double curl = (d[c1][c2] - d[c2][c1]) + (double)aux[c];
      const double scale = mat ? (double)(dt / mat[c * nc + cell]) : dt64 / mat_s;
      const double o = (double)out[c * nc + cell];
      out[c * nc + cell] = (T)(EUPD ? o + scale * curl : o - scale * curl);
    }
  }
}

// ---- conducting boundary (fdtd/boundaries.py:12-78) ---------------------------------
// One face's planes along `dir` are rewritten in place, one thread per (u, v) position
// of the face (u < v the other two axes).  A thread reads its column before it writes
// it and touches no other column, so threads need no ordering.
//   polarity -1, planes 0 and 1:
//     E: f_d[0] = 0, f_u[0] = f_u[1], f_v[0] = f_v[1]
//     H: f_d[0] = f_d[1], f_u[0] = f_v[0] = 0
//   polarity +1, planes L = n - 1, L - 1, L - 2:
//     E: f_d[L] = -f_d[L-2], f_d[L-1] = 0, f_u[L] = f_u[L-1], f_v[L] = f_v[L-1]
//     H: f_d[L] = f_d[L-1], f_u[L] = -f_u[L-2], f_u[L-1] = 0, and the same for v
// These are the image values of a mirror plane at 1/2 (polarity -1) or n - 3/2 (+1)
// across which tangential E and normal H are even, normal E and tangential H odd.
template <typename T>
__global__ __launch_bounds__(kFB) void k_conducting(int nx, int ny, int nz, int dir, int pol, int is_h,
                                                    T* __restrict__ f) {
  const int dims[3] = {nx, ny, nz};
  const long long strides[3] = {(long long)ny * nz, nz, 1};
  const int u = dir == 0 ? 1 : 0, v = dir == 2 ? 1 : 2;
  const long long nc = (long long)nx * ny * nz;
  const int nplane = dims[u] * dims[v];
  const int last = dims[dir] - 1;
  for (int t = blockIdx.x * kFB + threadIdx.x; t < nplane; t += gridDim.x * kFB) {
    const long long base = (t / dims[v]) * strides[u] + (t % dims[v]) * strides[v];
    auto at = [&](int comp, int plane) -> T& { return f[comp * nc + base + plane * strides[dir]]; };
    if (pol < 0) {
      if (!is_h) {
        const T fu = at(u, 1), fv = at(v, 1);
        at(dir, 0) = T(0);
        at(u, 0) = fu;
        at(v, 0) = fv;
      } else {
        const T fd = at(dir, 1);
        at(dir, 0) = fd;
        at(u, 0) = T(0);
        at(v, 0) = T(0);
      }
    } else {
      if (!is_h) {
        const T fd = at(dir, last - 2), fu = at(u, last - 1), fv = at(v, last - 1);
        at(dir, last) = -fd;
        at(dir, last - 1) = T(0);
        at(u, last) = fu;
        at(v, last) = fv;
      } else {
        const T fd = at(dir, last - 1), fu = at(u, last - 2), fv = at(v, last - 2);
        at(dir, last) = fd;
        at(u, last) = -fu;
        at(u, last - 1) = T(0);
        at(v, last) = -fv;
        at(v, last - 1) = T(0);
      }
    }
  }
}

// ---- energy and Poynting flux (fdtd/energy.py) ------------------------------------
// Rounding follows the reference's array expressions: the field products and the
// component sum in the field precision (in component order, starting from the x term),
// the widths in double (the float64 width arrays promote) multiplied one axis at a time,
// and nothing contracted into fused multiply-adds.  `form` is one of MX_ENERGY_*.

template <typename T>
__global__ __launch_bounds__(kFB) void k_energy(RGrid g, int form, T dt, const T* __restrict__ f0,
                                                const T* __restrict__ f1, const T* __restrict__ f2,
                                                const T* __restrict__ f3, const T* __restrict__ eps, T eps_s,
                                                const T* __restrict__ mu, T mu_s, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int nc = g.nx * g.ny * g.nz;
  for (int cell = blockIdx.x * kFB + threadIdx.x; cell < nc; cell += gridDim.x * kFB) {
    const int k = cell % g.nz, t = cell / g.nz, j = t % g.ny, i = t / g.ny;
    T se = T(0), sh = T(0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int q = c * nc + cell;
      T ee, hh = T(0);
      switch (form) {
        case MX_ENERGY_DXMUL: ee = f0[q]; hh = f1[q]; break;
        case MX_ENERGY_HSTEP: ee = f0[q] * f2[q]; hh = f1[q] * f1[q]; break;
        case MX_ENERGY_ESTEP: ee = f1[q] * f1[q]; hh = f0[q] * f2[q]; break;
        case MX_ENERGY_H2E: ee = f2[q] * (f2[q] - f0[q]) / dt; hh = f1[q] * (f3[q] - f1[q]) / dt; break;
        case MX_ENERGY_E2H: ee = f1[q] * (f3[q] - f1[q]) / dt; hh = f2[q] * (f2[q] - f0[q]) / dt; break;
        default: ee = f0[q] * f1[q]; break;   // j0 . e1, no material
      }
      const T we = form == MX_ENERGY_J ? ee : ee * (eps ? eps[q] : eps_s);
      const T wh = hh * (mu ? mu[q] : mu_s);
      se = c == 0 ? we : se + we;
      sh = c == 0 ? wh : sh + wh;
    }
    const double ue = (double)se * width(g.dxe[0], i) * width(g.dxe[1], j) * width(g.dxe[2], k);
    out[cell] = form == MX_ENERGY_J ? ue : ue + (double)sh * width(g.dxh[0], i) * width(g.dxh[1], j) * width(g.dxh[2], k);
  }
}

// s = E(+1) x H with both width-weighted (fdtd/energy.py:10-87): s_x = ey(i+1) hz - ez(i+1) hy
// and cyclically, E taken at the next cell along the component's axis (periodic), in
// double, stored in the field precision.
template <typename T>
__global__ __launch_bounds__(kFB) void k_poynting(RGrid g, const T* __restrict__ e, const T* __restrict__ h,
                                                  T* __restrict__ s) {
#pragma clang fp contract(off)
  const int nc = g.nx * g.ny * g.nz;
  for (int cell = blockIdx.x * kFB + threadIdx.x; cell < nc; cell += gridDim.x * kFB) {
    const int k = cell % g.nz, t = cell / g.nz, j = t % g.ny, i = t / g.ny;
    const int sx = (i + 1 == g.nx ? -i : 1) * g.ny * g.nz;
    const int sy = (j + 1 == g.ny ? -j : 1) * g.nz;
    const int sz = (k + 1 == g.nz ? -k : 1);
    const double wex = width(g.dxe[0], i), wey = width(g.dxe[1], j), wez = width(g.dxe[2], k);
    const double hx = (double)h[cell] * width(g.dxh[0], i);
    const double hy = (double)h[nc + cell] * width(g.dxh[1], j);
    const double hz = (double)h[2 * nc + cell] * width(g.dxh[2], k);
    const double s0 = ((double)e[nc + cell + sx] * wey) * hz - ((double)e[2 * nc + cell + sx] * wez) * hy;
    const double s1 = ((double)e[2 * nc + cell + sy] * wez) * hx - ((double)e[cell + sy] * wex) * hz;
    const double s2 = ((double)e[cell + sz] * wex) * hy - ((double)e[nc + cell + sz] * wey) * hx;
    s[cell] = (T)s0;
    s[nc + cell] = (T)s1;
    s[2 * nc + cell] = (T)s2;
  }
}

// ds = sum_a (s_a - s_a(previous cell along a)), unit spacing (deriv_back() without
// widths, fdtd/energy.py:90-125), in the field precision.
template <typename T>
__global__ __launch_bounds__(kFB) void k_poynting_div(RGrid g, const T* __restrict__ s, T* __restrict__ ds) {
#pragma clang fp contract(off)
  const int nc = g.nx * g.ny * g.nz;
  for (int cell = blockIdx.x * kFB + threadIdx.x; cell < nc; cell += gridDim.x * kFB) {
    const int k = cell % g.nz, t = cell / g.nz, j = t % g.ny, i = t / g.ny;
    const int sx = (i == 0 ? g.nx - 1 : -1) * g.ny * g.nz;   // i-1 (periodic)
    const int sy = (j == 0 ? g.ny - 1 : -1) * g.nz;
    const int sz = (k == 0 ? g.nz - 1 : -1);
    const T dx = s[cell] - s[cell + sx];
    const T dy = s[nc + cell] - s[nc + cell + sy];
    const T dz = s[2 * nc + cell] - s[2 * nc + cell + sz];
    ds[cell] = dx + dy + dz;
  }
}

bool rgrid(const mx_fdtd_grid* g, RGrid* v, std::string* why) {
  if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1) { *why = "bad grid"; return false; }
  if (3LL * g->nx * g->ny * g->nz >= (1LL << 31)) { *why = "grid too large (3N must be < 2^31)"; return false; }
  v->nx = g->nx; v->ny = g->ny; v->nz = g->nz;
  for (int a = 0; a < 3; ++a) {
    v->dxe[a] = g->dxe[a];
    v->dxh[a] = g->dxh[a];
  }
  return true;
}

unsigned blocks(const RGrid& v) {
  const long long nc = (long long)v.nx * v.ny * v.nz;
  const long long nb = (nc + kFB - 1) / kFB;
  return (unsigned)(nb > 16384 ? 16384 : nb);
}

hipError_t launch_h(const RGrid& v, int prec, double dt, const void* e, void* h, const void* mu, double mu_s,
                    hipStream_t s) {
  if (prec == MX_REAL64)
    hipLaunchKernelGGL(k_fdtd_h<double>, dim3(blocks(v)), dim3(kFB), 0, s, v, dt, (const double*)e, (double*)h,
                       (const double*)mu, mu_s);
  else
    hipLaunchKernelGGL(k_fdtd_h<float>, dim3(blocks(v)), dim3(kFB), 0, s, v, (float)dt, (const float*)e, (float*)h,
                       (const float*)mu, (float)mu_s);
  return hipGetLastError();
}

hipError_t launch_e(const RGrid& v, int prec, double dt, void* e, const void* h, const void* eps, double eps_s,
                    hipStream_t s) {
  if (prec == MX_REAL64)
    hipLaunchKernelGGL(k_fdtd_e<double>, dim3(blocks(v)), dim3(kFB), 0, s, v, dt, (double*)e, (const double*)h,
                       (const double*)eps, eps_s);
  else
    hipLaunchKernelGGL(k_fdtd_e<float>, dim3(blocks(v)), dim3(kFB), 0, s, v, (float)dt, (float*)e, (const float*)h,
                       (const float*)eps, (float)eps_s);
  return hipGetLastError();
}

hipError_t launch_conducting(const RGrid& v, int prec, int dir, int pol, int is_h, void* f, hipStream_t s) {
  const int dims[3] = {v.nx, v.ny, v.nz};
  const long long nplane = (long long)dims[dir == 0 ? 1 : 0] * dims[dir == 2 ? 1 : 2];
  const long long nb = (nplane + kFB - 1) / kFB;
  const unsigned grid = (unsigned)(nb > 16384 ? 16384 : nb);
  if (prec == MX_REAL64)
    hipLaunchKernelGGL(k_conducting<double>, dim3(grid), dim3(kFB), 0, s, v.nx, v.ny, v.nz, dir, pol, is_h,
                       (double*)f);
  else
    hipLaunchKernelGGL(k_conducting<float>, dim3(grid), dim3(kFB), 0, s, v.nx, v.ny, v.nz, dir, pol, is_h,
                       (float*)f);
  return hipGetLastError();
}

hipError_t launch_energy(const RGrid& v, int prec, int form, double dt, const void* const* f, const void* eps,
                         double eps_s, const void* mu, double mu_s, double* out, hipStream_t s) {
  if (prec == MX_REAL64)
    hipLaunchKernelGGL(k_energy<double>, dim3(blocks(v)), dim3(kFB), 0, s, v, form, dt, (const double*)f[0],
                       (const double*)f[1], (const double*)f[2], (const double*)f[3], (const double*)eps, eps_s,
                       (const double*)mu, mu_s, out);
  else
    hipLaunchKernelGGL(k_energy<float>, dim3(blocks(v)), dim3(kFB), 0, s, v, form, (float)dt, (const float*)f[0],
                       (const float*)f[1], (const float*)f[2], (const float*)f[3], (const float*)eps, (float)eps_s,
                       (const float*)mu, (float)mu_s, out);
  return hipGetLastError();
}

hipError_t launch_poynting(const RGrid& v, int prec, const void* e, const void* h, void* out, hipStream_t s) {
  if (prec == MX_REAL64)
    hipLaunchKernelGGL(k_poynting<double>, dim3(blocks(v)), dim3(kFB), 0, s, v, (const double*)e, (const double*)h,
                       (double*)out);
  else
    hipLaunchKernelGGL(k_poynting<float>, dim3(blocks(v)), dim3(kFB), 0, s, v, (const float*)e, (const float*)h,
                       (float*)out);
  return hipGetLastError();
}

hipError_t launch_poynting_div(const RGrid& v, int prec, const void* sv, void* ds, hipStream_t s) {
  if (prec == MX_REAL64)
    hipLaunchKernelGGL(k_poynting_div<double>, dim3(blocks(v)), dim3(kFB), 0, s, v, (const double*)sv, (double*)ds);
  else
    hipLaunchKernelGGL(k_poynting_div<float>, dim3(blocks(v)), dim3(kFB), 0, s, v, (const float*)sv, (float*)ds);
  return hipGetLastError();
}

// Layer table of a CPML update: faces in increasing (axis, polarity) order (the order
// the reference accumulates in), each inside the grid, with its arrays.
bool cpml_table(const RGrid& v, const mx_cpml_layer* layers, int n, CpmlTable* tab, std::string* why) {
  if (n < 0 || n > MX_CPML_MAX_LAYERS) { *why = "nlayers must be 0..6"; return false; }
  if (n > 0 && !layers) { *why = "null layer table"; return false; }
  const int dims[3] = {v.nx, v.ny, v.nz};
  tab->n = n;
  int prev = -1;
  for (int q = 0; q < n; ++q) {
    const mx_cpml_layer& L = layers[q];
    if (L.axis < 0 || L.axis > 2) { *why = "layer axis must be 0, 1 or 2"; return false; }
    if (L.polarity != -1 && L.polarity != 1) { *why = "layer polarity must be -1 or +1"; return false; }
    if (L.thickness < 1 || L.thickness > dims[L.axis]) { *why = "layer thickness must be 1..n along its axis"; return false; }
    const int key = 2 * L.axis + (L.polarity > 0);
    if (key <= prev) { *why = "layers must be in increasing (axis, polarity) order, one per face"; return false; }
    prev = key;
    if (!L.p_e || !L.p_h || !L.psi_e || !L.psi_h) { *why = "null layer array"; return false; }
    tab->l[q] = CpmlLayer{L.axis, L.polarity < 0 ? 0 : dims[L.axis] - L.thickness, L.thickness, L.p_e, L.p_h,
                          L.psi_e, L.psi_h};
  }
  return true;
}

template <bool EUPD>
hipError_t launch_cpml(const RGrid& v, int prec, int psi_prec, double dt, const void* e, const void* h,
                       const void* mat, double mat_s, const CpmlTable& tab, hipStream_t s) {
#define MX_CPML_CASE(T, TP)                                                                                  \
  hipLaunchKernelGGL((k_cpml<T, TP, EUPD>), dim3(blocks(v)), dim3(kFB), 0, s, v, (T)dt, (T*)e, (T*)h,       \
                     (const T*)mat, mat_s, dt, tab)
  if (prec == MX_REAL64) {
    if (psi_prec == MX_REAL64) MX_CPML_CASE(double, double);
    else MX_CPML_CASE(double, float);
  } else {
    if (psi_prec == MX_REAL64) MX_CPML_CASE(float, double);
    else MX_CPML_CASE(float, float);
  }
#undef MX_CPML_CASE
  return hipGetLastError();
}

// `steps` fused steps, (e, h) -> (e2, h2) -> (e, h) -> ...
hipError_t launch_steps(const RGrid& v, int prec, double dt, void* e, void* h, void* e2, void* h2, const void* eps,
                        double eps_s, const void* mu, double mu_s, int steps, hipStream_t s) {
  static const int forced = [] {   // MEANAS_AMD_FDTD_SLAB overrides the slab depth (tuning runs)
    const char* env = getenv("MEANAS_AMD_FDTD_SLAB");
    return env ? atoi(env) : 0;
  }();
  const int want = forced > 0 ? forced : kStepSlab;
  const int slab = v.nx < want ? v.nx : want;
  const long long tiles = (long long)((v.ny + kStepTY - 1) / kStepTY) * ((v.nz + 63) / 64);
  const long long nwork = (long long)((v.nx + slab - 1) / slab) * tiles;
  const unsigned nb = (unsigned)(((nwork + 7) / 8) * 8);
  for (int t = 0; t < steps; ++t) {
    void* ei = (t & 1) ? e2 : e;
    void* hi = (t & 1) ? h2 : h;
    void* eo = (t & 1) ? e : e2;
    void* ho = (t & 1) ? h : h2;
    if (prec == MX_REAL64)
      hipLaunchKernelGGL(k_fdtd_step<double>, dim3(nb), dim3(64, kStepTY), 0, s, v, dt, (const double*)ei,
                         (const double*)hi, (double*)eo, (double*)ho, (const double*)eps, eps_s, (const double*)mu,
                         mu_s, slab);
    else
      hipLaunchKernelGGL(k_fdtd_step<float>, dim3(nb), dim3(64, kStepTY), 0, s, v, (float)dt, (const float*)ei,
                         (const float*)hi, (float*)eo, (float*)ho, (const float*)eps, (float)eps_s, (const float*)mu,
                         (float)mu_s, slab);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

int check_args(const mx_fdtd_grid* g, int prec, const void* a, const void* b, RGrid* v, const char* who) {
  std::string why;
  if (!rgrid(g, v, &why)) return set_error(std::string(who) + ": " + why);
  if (prec != MX_REAL32 && prec != MX_REAL64) return set_error(std::string(who) + ": prec must be MX_REAL32 or MX_REAL64");
  if (!a || !b) return set_error(std::string(who) + ": null field");
  return 0;
}

}  // namespace
}  // namespace mx

extern "C" {

int mx_fdtd_h(const mx_fdtd_grid* g, int32_t prec, double dt, const void* e, void* h, const void* mu,
              double mu_scalar, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, e, h, &v, "mx_fdtd_h")) return -1;
  const hipError_t err = mx::launch_h(v, prec, dt, e, h, mu, mu_scalar, reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_h: ") + hipGetErrorString(err));
}

int mx_fdtd_e(const mx_fdtd_grid* g, int32_t prec, double dt, void* e, const void* h, const void* eps,
              double eps_scalar, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, e, h, &v, "mx_fdtd_e")) return -1;
  const hipError_t err = mx::launch_e(v, prec, dt, e, h, eps, eps_scalar, reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_e: ") + hipGetErrorString(err));
}

int mx_fdtd_run(const mx_fdtd_grid* g, int32_t prec, double dt, void* e, void* h, const void* eps,
                double eps_scalar, const void* mu, double mu_scalar, int32_t steps, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, e, h, &v, "mx_fdtd_run")) return -1;
  if (steps < 0) return mx::set_error("mx_fdtd_run: steps < 0");
  if (steps == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // fused steps alternate between (e, h) and a second pair, allocated in stream order
  const size_t bytes = (size_t)3 * v.nx * v.ny * v.nz * (prec == MX_REAL64 ? 8 : 4);
  void* e2 = nullptr;
  void* h2 = nullptr;
  hipError_t err = hipMallocAsync(&e2, bytes, s);
  if (err == hipSuccess) err = hipMallocAsync(&h2, bytes, s);
  if (err == hipSuccess) err = mx::launch_steps(v, prec, dt, e, h, e2, h2, eps, eps_scalar, mu, mu_scalar, steps, s);
  if (err == hipSuccess && (steps & 1)) {   // the last step wrote the second pair
    err = hipMemcpyAsync(e, e2, bytes, hipMemcpyDeviceToDevice, s);
    if (err == hipSuccess) err = hipMemcpyAsync(h, h2, bytes, hipMemcpyDeviceToDevice, s);
  }
  if (e2) (void)hipFreeAsync(e2, s);
  if (h2) (void)hipFreeAsync(h2, s);
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_run: ") + hipGetErrorString(err));
}

int mx_fdtd_cpml_e(const mx_fdtd_grid* g, int32_t prec, int32_t psi_prec, double dt, void* e, const void* h,
                   const void* eps, double eps_scalar, const mx_cpml_layer* layers, int32_t nlayers, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, e, h, &v, "mx_fdtd_cpml_e")) return -1;
  if (psi_prec != MX_REAL32 && psi_prec != MX_REAL64) return mx::set_error("mx_fdtd_cpml_e: psi_prec must be MX_REAL32 or MX_REAL64");
  mx::CpmlTable tab;
  std::string why;
  if (!mx::cpml_table(v, layers, nlayers, &tab, &why)) return mx::set_error("mx_fdtd_cpml_e: " + why);
  const hipError_t err = mx::launch_cpml<true>(v, prec, psi_prec, dt, e, h, eps, eps_scalar, tab,
                                               reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_cpml_e: ") + hipGetErrorString(err));
}

int mx_fdtd_cpml_h(const mx_fdtd_grid* g, int32_t prec, int32_t psi_prec, double dt, const void* e, void* h,
                   const void* mu, double mu_scalar, const mx_cpml_layer* layers, int32_t nlayers, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, e, h, &v, "mx_fdtd_cpml_h")) return -1;
  if (psi_prec != MX_REAL32 && psi_prec != MX_REAL64) return mx::set_error("mx_fdtd_cpml_h: psi_prec must be MX_REAL32 or MX_REAL64");
  mx::CpmlTable tab;
  std::string why;
  if (!mx::cpml_table(v, layers, nlayers, &tab, &why)) return mx::set_error("mx_fdtd_cpml_h: " + why);
  const hipError_t err = mx::launch_cpml<false>(v, prec, psi_prec, dt, e, h, mu, mu_scalar, tab,
                                                reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_cpml_h: ") + hipGetErrorString(err));
}

int mx_fdtd_conducting_boundary(const mx_fdtd_grid* g, int32_t prec, int32_t direction, int32_t polarity,
                                int32_t field, void* f, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, f, f, &v, "mx_fdtd_conducting_boundary")) return -1;
  if (direction < 0 || direction > 2) return mx::set_error("mx_fdtd_conducting_boundary: direction must be 0, 1 or 2");
  if (polarity != -1 && polarity != 1) return mx::set_error("mx_fdtd_conducting_boundary: polarity must be -1 or +1");
  if (field != MX_FIELD_E && field != MX_FIELD_H)
    return mx::set_error("mx_fdtd_conducting_boundary: field must be MX_FIELD_E or MX_FIELD_H");
  const int dims[3] = {v.nx, v.ny, v.nz};
  if (dims[direction] < (polarity < 0 ? 2 : 3))
    return mx::set_error("mx_fdtd_conducting_boundary: the axis needs at least 2 cells (polarity -1) or 3 (+1)");
  const hipError_t err = mx::launch_conducting(v, prec, direction, polarity, field == MX_FIELD_H, f,
                                               reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_conducting_boundary: ") + hipGetErrorString(err));
}

int mx_fdtd_energy(const mx_fdtd_grid* g, int32_t prec, int32_t form, double dt, const void* const* f,
                   const void* eps, double eps_scalar, const void* mu, double mu_scalar, double* out, void* stream) {
  static const int kNeeds[6] = {2, 3, 3, 4, 4, 2};   // fields each form reads
  mx::RGrid v;
  if (!f) return mx::set_error("mx_fdtd_energy: null field table");
  if (mx::check_args(g, prec, f[0], out, &v, "mx_fdtd_energy")) return -1;
  if (form < MX_ENERGY_DXMUL || form > MX_ENERGY_J) return mx::set_error("mx_fdtd_energy: unknown form");
  for (int q = 0; q < kNeeds[form]; ++q)
    if (!f[q]) return mx::set_error("mx_fdtd_energy: null field for this form");
  if ((form == MX_ENERGY_H2E || form == MX_ENERGY_E2H) && !(dt != 0.0)) return mx::set_error("mx_fdtd_energy: dt must be nonzero");
  const void* fs[4] = {f[0], f[1], kNeeds[form] > 2 ? f[2] : nullptr, kNeeds[form] > 3 ? f[3] : nullptr};
  const hipError_t err = mx::launch_energy(v, prec, form, dt, fs, eps, eps_scalar, mu, mu_scalar, out,
                                           reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_energy: ") + hipGetErrorString(err));
}

int mx_fdtd_poynting(const mx_fdtd_grid* g, int32_t prec, const void* e, const void* h, void* s, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, e, h, &v, "mx_fdtd_poynting")) return -1;
  if (!s) return mx::set_error("mx_fdtd_poynting: null output");
  const hipError_t err = mx::launch_poynting(v, prec, e, h, s, reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_poynting: ") + hipGetErrorString(err));
}

int mx_fdtd_poynting_divergence(const mx_fdtd_grid* g, int32_t prec, const void* s, void* ds, void* stream) {
  mx::RGrid v;
  if (mx::check_args(g, prec, s, ds, &v, "mx_fdtd_poynting_divergence")) return -1;
  const hipError_t err = mx::launch_poynting_div(v, prec, s, ds, reinterpret_cast<hipStream_t>(stream));
  return err == hipSuccess ? 0 : mx::set_error(std::string("mx_fdtd_poynting_divergence: ") + hipGetErrorString(err));
}

}  // extern "C"
