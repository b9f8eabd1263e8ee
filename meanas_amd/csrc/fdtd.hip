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

}  // extern "C"
