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
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (int t = 0; t < steps; ++t) {
    hipError_t err = mx::launch_h(v, prec, dt, e, h, mu, mu_scalar, s);
    if (err == hipSuccess) err = mx::launch_e(v, prec, dt, e, h, eps, eps_scalar, s);
    if (err != hipSuccess) return mx::set_error(std::string("mx_fdtd_run: ") + hipGetErrorString(err));
  }
  return 0;
}

}  // extern "C"
