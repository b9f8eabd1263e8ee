// Fields after the solve: H from E, and the Poynting cross product.
//
// e2h: H = curl_forward(E) / (-i omega) / mu, zero on PMC components
//   (meanas/fdfd/operators.py:253-282 sparse; meanas/fdfd/functional.py:89-117
//   matrix-free).  curl_forward = [[0, -Dz, Dy], [Dz, 0, -Dx], [-Dy, Dx, 0]] with the
//   periodic forward difference (D f)[n] = (f[n+1] - f[n]) / dx_e[n]
//   (meanas/fdmath/operators.py:99-171, 222-235).
// poynting_e_cross_h: s = E x H with E shifted one cell forward along the axis of the
//   cross-product row and both fields weighted by their cell widths along their own
//   component axis (meanas/fdfd/functional.py:187-221; the same product as
//   poynting_e_cross(e, dxes) @ h, meanas/fdfd/operators.py:310-336):
//     s_x = ey(i+1) hz - ez(i+1) hy,  s_y = ez(j+1) hx - ex(j+1) hz,  s_z = ex(k+1) hy - ey(k+1) hx
//   with e_c = E_c dx_e,c and h_c = H_c dx_h,c.
//
// Both are one pass over the grid, one thread per cell: HBM-bound.  The forward
// neighbours are re-read from the caches (k+1 is the next lane; j+1 and i+1 are nz and
// ny*nz cells away and are still in L2 from the neighbouring row / plane of threads).
#include <string>

#include "common.hpp"
#include "../../include/meanas_fdfd.h"

namespace mx {

int set_error(const std::string& msg);

namespace {

constexpr int kPB = 256;

struct GridView {
  int nx, ny, nz;
  const cplx<double>* dxe[3];
  const cplx<double>* dxh[3];
};

__device__ __forceinline__ void cell_coords(int cell, int ny, int nz, int* i, int* j, int* k) {
  *k = cell % nz;
  const int t = cell / nz;
  *j = t % ny;
  *i = t / ny;
}

__global__ __launch_bounds__(kPB) void k_e2h(GridView g, cplx<double> c, const cplx<double>* __restrict__ inv_mu,
                                             const uint8_t* __restrict__ pmc, const cplx<double>* __restrict__ e,
                                             cplx<double>* __restrict__ h) {
  const int nc = g.nx * g.ny * g.nz;
  for (int cell = blockIdx.x * kPB + threadIdx.x; cell < nc; cell += gridDim.x * kPB) {
    int i, j, k;
    cell_coords(cell, g.ny, g.nz, &i, &j, &k);
    const int sx = (i + 1 == g.nx ? -i : 1) * g.ny * g.nz;   // offset of the i+1 neighbour (periodic)
    const int sy = (j + 1 == g.ny ? -j : 1) * g.nz;
    const int sz = (k + 1 == g.nz ? -k : 1);
    const cplx<double> ex = e[cell], ey = e[nc + cell], ez = e[2 * nc + cell];
    const cplx<double> rx = cdiv(cplx<double>{1.0, 0.0}, g.dxe[0][i]);
    const cplx<double> ry = cdiv(cplx<double>{1.0, 0.0}, g.dxe[1][j]);
    const cplx<double> rz = cdiv(cplx<double>{1.0, 0.0}, g.dxe[2][k]);
    // forward differences of the components the curl needs
    const cplx<double> dy_ez = ry * (e[2 * nc + cell + sy] - ez), dz_ey = rz * (e[nc + cell + sz] - ey);
    const cplx<double> dz_ex = rz * (e[cell + sz] - ex), dx_ez = rx * (e[2 * nc + cell + sx] - ez);
    const cplx<double> dx_ey = rx * (e[nc + cell + sx] - ey), dy_ex = ry * (e[cell + sy] - ex);
    cplx<double> out[3] = {dy_ez - dz_ey, dz_ex - dx_ez, dx_ey - dy_ex};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cplx<double> v = c * out[q];
      if (inv_mu) v = inv_mu[q * nc + cell] * v;
      if (pmc && pmc[q * nc + cell]) v = cplx<double>{0.0, 0.0};
      h[q * nc + cell] = v;
    }
  }
}

// J = curl_back(M / mu) / (-i omega) (fdfd/functional.py:120-151, m2j): backward
// differences on the H-grid widths, periodic; M / mu is formed at the cell and at the
// three backward neighbours the curl reads.
__global__ __launch_bounds__(kPB) void k_m2j(GridView g, cplx<double> c, const cplx<double>* __restrict__ mu,
                                             const cplx<double>* __restrict__ m, cplx<double>* __restrict__ jout) {
  const int nc = g.nx * g.ny * g.nz;
  auto v = [&](int q) { return mu ? cdiv(m[q], mu[q]) : m[q]; };
  for (int cell = blockIdx.x * kPB + threadIdx.x; cell < nc; cell += gridDim.x * kPB) {
    int i, j, k;
    cell_coords(cell, g.ny, g.nz, &i, &j, &k);
    const int sx = (i == 0 ? g.nx - 1 : -1) * g.ny * g.nz;   // offset of the i-1 neighbour (periodic)
    const int sy = (j == 0 ? g.ny - 1 : -1) * g.nz;
    const int sz = (k == 0 ? g.nz - 1 : -1);
    const cplx<double> vx = v(cell), vy = v(nc + cell), vz = v(2 * nc + cell);
    const cplx<double> rx = cdiv(cplx<double>{1.0, 0.0}, g.dxh[0][i]);
    const cplx<double> ry = cdiv(cplx<double>{1.0, 0.0}, g.dxh[1][j]);
    const cplx<double> rz = cdiv(cplx<double>{1.0, 0.0}, g.dxh[2][k]);
    const cplx<double> dy_vz = ry * (vz - v(2 * nc + cell + sy)), dz_vy = rz * (vy - v(nc + cell + sz));
    const cplx<double> dz_vx = rz * (vx - v(cell + sz)), dx_vz = rx * (vz - v(2 * nc + cell + sx));
    const cplx<double> dx_vy = rx * (vy - v(nc + cell + sx)), dy_vx = ry * (vx - v(cell + sy));
    jout[cell] = c * (dy_vz - dz_vy);
    jout[nc + cell] = c * (dz_vx - dx_vz);
    jout[2 * nc + cell] = c * (dx_vy - dy_vx);
  }
}

__global__ __launch_bounds__(kPB) void k_exh(GridView g, int conj_h, const cplx<double>* __restrict__ e,
                                             const cplx<double>* __restrict__ hin, cplx<double>* __restrict__ s) {
  const int nc = g.nx * g.ny * g.nz;
  for (int cell = blockIdx.x * kPB + threadIdx.x; cell < nc; cell += gridDim.x * kPB) {
    int i, j, k;
    cell_coords(cell, g.ny, g.nz, &i, &j, &k);
    const int sx = (i + 1 == g.nx ? -i : 1) * g.ny * g.nz;
    const int sy = (j + 1 == g.ny ? -j : 1) * g.nz;
    const int sz = (k + 1 == g.nz ? -k : 1);
    cplx<double> hv[3] = {hin[cell], hin[nc + cell], hin[2 * nc + cell]};
    if (conj_h)
      for (int q = 0; q < 3; ++q) hv[q] = conj(hv[q]);
    const cplx<double> hx = hv[0] * g.dxh[0][i], hy = hv[1] * g.dxh[1][j], hz = hv[2] * g.dxh[2][k];
    // E components at the forward neighbour, weighted by the width along their own axis
    const cplx<double> ey_x = e[nc + cell + sx] * g.dxe[1][j], ez_x = e[2 * nc + cell + sx] * g.dxe[2][k];
    const cplx<double> ez_y = e[2 * nc + cell + sy] * g.dxe[2][k], ex_y = e[cell + sy] * g.dxe[0][i];
    const cplx<double> ex_z = e[cell + sz] * g.dxe[0][i], ey_z = e[nc + cell + sz] * g.dxe[1][j];
    s[cell] = ey_x * hz - ez_x * hy;
    s[nc + cell] = ez_y * hx - ex_y * hz;
    s[2 * nc + cell] = ex_z * hy - ey_z * hx;
  }
}

// s = poynting_h_cross(h, dxes) @ e (fdfd/operators.py:339-363): the width-weighted H at
// the cell against the width-weighted E at the forward neighbour, s = -(E(+1) x H); e is
// conjugated first when conj_e is set.
__global__ __launch_bounds__(kPB) void k_hxe(GridView g, int conj_e, const cplx<double>* __restrict__ hin,
                                             const cplx<double>* __restrict__ e, cplx<double>* __restrict__ s) {
  const int nc = g.nx * g.ny * g.nz;
  auto ev = [&](int q) { return conj_e ? conj(e[q]) : e[q]; };
  for (int cell = blockIdx.x * kPB + threadIdx.x; cell < nc; cell += gridDim.x * kPB) {
    int i, j, k;
    cell_coords(cell, g.ny, g.nz, &i, &j, &k);
    const int sx = (i + 1 == g.nx ? -i : 1) * g.ny * g.nz;
    const int sy = (j + 1 == g.ny ? -j : 1) * g.nz;
    const int sz = (k + 1 == g.nz ? -k : 1);
    const cplx<double> hx = hin[cell] * g.dxh[0][i], hy = hin[nc + cell] * g.dxh[1][j];
    const cplx<double> hz = hin[2 * nc + cell] * g.dxh[2][k];
    const cplx<double> ey_x = ev(nc + cell + sx) * g.dxe[1][j], ez_x = ev(2 * nc + cell + sx) * g.dxe[2][k];
    const cplx<double> ez_y = ev(2 * nc + cell + sy) * g.dxe[2][k], ex_y = ev(cell + sy) * g.dxe[0][i];
    const cplx<double> ex_z = ev(cell + sz) * g.dxe[0][i], ey_z = ev(nc + cell + sz) * g.dxe[1][j];
    s[cell] = hy * ez_x - hz * ey_x;
    s[nc + cell] = hz * ex_y - hx * ez_y;
    s[2 * nc + cell] = hx * ey_z - hy * ex_z;
  }
}

bool grid_view(const mx_grid* g, GridView* v, std::string* why) {
  if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1) { *why = "bad grid"; return false; }
  if (3LL * g->nx * g->ny * g->nz * 16 >= (1LL << 31)) { *why = "grid too large (3N complex128 must be < 2 GiB)"; return false; }
  v->nx = g->nx; v->ny = g->ny; v->nz = g->nz;
  for (int a = 0; a < 3; ++a) {
    if (!g->dxe[a] || !g->dxh[a]) { *why = "missing cell widths"; return false; }
    v->dxe[a] = reinterpret_cast<const cplx<double>*>(g->dxe[a]);
    v->dxh[a] = reinterpret_cast<const cplx<double>*>(g->dxh[a]);
  }
  return true;
}

unsigned grid_blocks(const GridView& v) {
  const long long nc = (long long)v.nx * v.ny * v.nz;
  long long nb = (nc + kPB - 1) / kPB;
  return (unsigned)(nb > 8192 ? 8192 : nb);
}

}  // namespace
}  // namespace mx

extern "C" {

int mx_e2h(const mx_grid* g, double omega_re, double omega_im, const void* inv_mu, const uint8_t* pmc,
           const void* e, void* h, void* stream) {
  using namespace mx;
  GridView v;
  std::string why;
  if (!grid_view(g, &v, &why)) return set_error("mx_e2h: " + why);
  if (!e || !h) return set_error("mx_e2h: null field");
  const cplx<double> w{omega_re, omega_im};
  if (w.re == 0.0 && w.im == 0.0) return set_error("mx_e2h: omega = 0");
  const cplx<double> c = cdiv(cplx<double>{1.0, 0.0}, cplx<double>{w.im, -w.re});   // 1 / (-i omega)
  hipLaunchKernelGGL(k_e2h, dim3(grid_blocks(v)), dim3(kPB), 0, reinterpret_cast<hipStream_t>(stream), v, c,
                     reinterpret_cast<const cplx<double>*>(inv_mu), pmc, reinterpret_cast<const cplx<double>*>(e),
                     reinterpret_cast<cplx<double>*>(h));
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? 0 : set_error(std::string("mx_e2h: ") + hipGetErrorString(err));
}

int mx_m2j(const mx_grid* g, double omega_re, double omega_im, const void* mu, const void* m, void* j,
           void* stream) {
  using namespace mx;
  GridView v;
  std::string why;
  if (!grid_view(g, &v, &why)) return set_error("mx_m2j: " + why);
  if (!m || !j) return set_error("mx_m2j: null field");
  const cplx<double> w{omega_re, omega_im};
  if (w.re == 0.0 && w.im == 0.0) return set_error("mx_m2j: omega = 0");
  const cplx<double> c = cdiv(cplx<double>{1.0, 0.0}, cplx<double>{w.im, -w.re});   // 1 / (-i omega)
  hipLaunchKernelGGL(k_m2j, dim3(grid_blocks(v)), dim3(kPB), 0, reinterpret_cast<hipStream_t>(stream), v, c,
                     reinterpret_cast<const cplx<double>*>(mu), reinterpret_cast<const cplx<double>*>(m),
                     reinterpret_cast<cplx<double>*>(j));
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? 0 : set_error(std::string("mx_m2j: ") + hipGetErrorString(err));
}

int mx_poynting_e_cross_h(const mx_grid* g, const void* e, const void* h, int32_t conj_h, void* s, void* stream) {
  using namespace mx;
  GridView v;
  std::string why;
  if (!grid_view(g, &v, &why)) return set_error("mx_poynting_e_cross_h: " + why);
  if (!e || !h || !s) return set_error("mx_poynting_e_cross_h: null field");
  hipLaunchKernelGGL(k_exh, dim3(grid_blocks(v)), dim3(kPB), 0, reinterpret_cast<hipStream_t>(stream), v, (int)conj_h,
                     reinterpret_cast<const cplx<double>*>(e), reinterpret_cast<const cplx<double>*>(h),
                     reinterpret_cast<cplx<double>*>(s));
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? 0 : set_error(std::string("mx_poynting_e_cross_h: ") + hipGetErrorString(err));
}

int mx_poynting_h_cross_e(const mx_grid* g, const void* h, const void* e, int32_t conj_e, void* s, void* stream) {
  using namespace mx;
  GridView v;
  std::string why;
  if (!grid_view(g, &v, &why)) return set_error("mx_poynting_h_cross_e: " + why);
  if (!e || !h || !s) return set_error("mx_poynting_h_cross_e: null field");
  hipLaunchKernelGGL(k_hxe, dim3(grid_blocks(v)), dim3(kPB), 0, reinterpret_cast<hipStream_t>(stream), v, (int)conj_e,
                     reinterpret_cast<const cplx<double>*>(h), reinterpret_cast<const cplx<double>*>(e),
                     reinterpret_cast<cplx<double>*>(s));
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? 0 : set_error(std::string("mx_poynting_h_cross_e: ") + hipGetErrorString(err));
}

}  // extern "C"
