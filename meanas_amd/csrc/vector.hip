// Fused Krylov vector kernels for complex-symmetric COCG with residual replacement,
// plus the diagonal maps J -> b and x -> E of meanas.fdfd.solvers.generic
// (fdfd/solvers.py:115,124-125,136-139).
//
// COCG (van der Vorst & Melissen 1990) for A = A^T uses the unconjugated products
// rho = r^T r and mu = p^T A p.  One iteration here is three launches:
//   k_pupdate : p = r + beta p              (p = r after a restart)
//   k_apply   : q = A p, mu = p^T q, alpha = rho / mu      (apply.hip)
//   k_update  : x_lo += alpha p, r -= alpha q, rho' = r^T r, ||r||^2, beta = rho'/rho
// Residual replacement (Sleijpen & van der Vorst 1996 "reliable updates"; the mixed
// precision form of Clark et al. 2010): when the recursive ||r|| has dropped by
// `delta` since the last replacement, or reached the target, the system pauses and
//   k_accum   : x += x_lo (FP64), x_lo = 0
//   k_apply<RESID> : r = b - A x in FP64, ||r||^2 in FP64, rho' from the stored r
// replace the recursive residual with the true one and keep the search direction
// (beta = rho'/rho_old).  Every scalar is reduced in a fixed order by the last block
// to arrive, so results do not depend on block scheduling or on batch composition.
#include "common.hpp"

namespace mx {

constexpr int VB = 256;   // threads per vector block

__device__ __forceinline__ bool skip_sys(int f, int skip, int need) {
  return (f & skip) || (need && !(f & need));
}

// ---- p = r + beta p ------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(VB) void k_pupdate(const OpView<T>* __restrict__ ops, const SysDesc* __restrict__ systems,
                                              const BlockMap* __restrict__ bmap) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  const int f = S.st->flags;
  if (skip_sys(f, SF_PAUSE | SF_DONE | SF_INACTIVE, 0)) return;
  const cplx<T>* __restrict__ r = reinterpret_cast<const cplx<T>*>(S.vec[VS_R]);
  cplx<T>* __restrict__ p = reinterpret_cast<cplx<T>*>(S.vec[VS_P]);
  const long long n = S.n3;
  const long long stride = (long long)S.nblk_vec * VB;
  if (f & SF_MR) {   // breakdown rescue: p = hashed direction, zero on PEC entries
    const OpView<T>& op = ops[S.op];
    const long long ncell = n / 3;
    for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) {
      const int c = (int)(e / ncell);
      const bool pec = op.mask && ((op.mask[e - c * ncell] >> c) & 1);
      p[e] = pec ? zero<T>() : cast<T>(hashed_direction(e));
    }
  } else if (f & SF_RESTART) {
    for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) p[e] = r[e];
  } else {
    const cplx<T> beta = cast<T>(S.st->beta);
    for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) p[e] = r[e] + beta * p[e];
  }
}

// ---- x_lo += alpha p ; r -= alpha q ; rho, ||r||^2 ---------------------------------
template <typename T>
__global__ __launch_bounds__(VB) void k_update(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap) {
  __shared__ double s_red[3 * 16];
  __shared__ int s_flag;
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  SysState* st = S.st;
  if (skip_sys(st->flags, SF_PAUSE | SF_DONE | SF_INACTIVE, 0)) return;
  cplx<T>* __restrict__ x = reinterpret_cast<cplx<T>*>(S.vec[VS_XLO]);
  cplx<T>* __restrict__ r = reinterpret_cast<cplx<T>*>(S.vec[VS_R]);
  const cplx<T>* __restrict__ p = reinterpret_cast<const cplx<T>*>(S.vec[VS_P]);
  const cplx<T>* __restrict__ q = reinterpret_cast<const cplx<T>*>(S.vec[VS_Q]);
  const cplx<T> alpha = cast<T>(st->alpha);
  const long long n = S.n3;
  const long long stride = (long long)S.nblk_vec * VB;
  T rr_re = 0, rr_im = 0, nr = 0;
  for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) {
    x[e] = x[e] + alpha * p[e];
    const cplx<T> rv = r[e] - alpha * q[e];
    r[e] = rv;
    rr_re += rv.re * rv.re - rv.im * rv.im;
    rr_im += T(2) * rv.re * rv.im;
    nr += rv.re * rv.re + rv.im * rv.im;
  }
  double acc[3] = {(double)rr_re, (double)rr_im, (double)nr};
  block_sum<3>(acc, s_red);
  const int nblk = S.nblk_vec;
  if (!publish_partials<3>(acc, S.partial, bm.local, nblk, &st->counter[1], &s_flag)) return;
  double tot[3];
  final_sum<3>(tot, S.partial, nblk, s_red);
  if (threadIdx.x == 0) {
    st->counter[1] = 0;
    int f = st->flags;
    const cplx<double> rho{tot[0], tot[1]};
    const double rn2 = tot[2];
    if (f & SF_MR) {   // that was a minimal-residual step: COCG restarts from p = r
      f &= ~SF_MR;
      st->beta = cplx<double>{0.0, 0.0};
    } else if (norm2(st->rho) > 0.0) {
      st->beta = cdiv(rho, st->rho);
    } else {   // rho breakdown: restart from the replaced residual
      f |= SF_BREAKDOWN | SF_PAUSE | SF_RESTART;
      st->beta = cplx<double>{0.0, 0.0};
    }
    // r^T r ~ 0 with r != 0: the next step is a minimal-residual step from p = r
    if (norm2(rho) <= kRhoBreakdown * kRhoBreakdown * rn2 * rn2 && rn2 > 0.0) f |= SF_MR | SF_RESTART;
    st->rho_old = st->rho;
    st->rho = rho;
    st->rnorm2 = rn2;
    if (rn2 > st->rmax2) st->rmax2 = rn2;
    st->iters += 1;
    st->since += 1;
    if (!finite(rn2) || !finite(rho.re) || !finite(rho.im)) {
      f |= SF_NAN | SF_PAUSE | SF_LIMIT;
    } else if (rn2 <= st->stop2 || rn2 <= st->delta2 * st->rmax2 ||
               (st->stall_iters > 0 && st->since >= st->stall_iters)) {
      f |= SF_PAUSE;
    }
    if (st->iters >= st->maxiter) f |= SF_PAUSE | SF_LIMIT;
    st->flags = f;
  }
}

// ---- x (c128) += x_lo ; x_lo = 0   (paused systems) --------------------------------
template <typename T>
__global__ __launch_bounds__(VB) void k_accum(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  if (skip_sys(S.st->flags, SF_DONE | SF_INACTIVE, SF_PAUSE)) return;
  cplx<double>* __restrict__ x = reinterpret_cast<cplx<double>*>(S.vec[VS_X]);
  cplx<T>* __restrict__ xl = reinterpret_cast<cplx<T>*>(S.vec[VS_XLO]);
  const long long n = S.n3;
  const long long stride = (long long)S.nblk_vec * VB;
  for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) {
    x[e] = x[e] + cast<double>(xl[e]);
    xl[e] = zero<T>();
  }
}

// ---- ||b||^2 and the stopping target, once per solve --------------------------------
__global__ __launch_bounds__(VB) void k_bnorm(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap) {
  __shared__ double s_red[1 * 16];
  __shared__ int s_flag;
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  SysState* st = S.st;
  if (skip_sys(st->flags, SF_DONE | SF_INACTIVE, 0)) return;
  const cplx<double>* __restrict__ b = reinterpret_cast<const cplx<double>*>(S.vec[VS_B]);
  const long long n = S.n3;
  const long long stride = (long long)S.nblk_vec * VB;
  double acc[1] = {0.0};
  for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) acc[0] += norm2(b[e]);
  block_sum<1>(acc, s_red);
  const int nblk = S.nblk_vec;
  if (!publish_partials<1>(acc, S.partial, bm.local, nblk, &st->counter[3], &s_flag)) return;
  double tot[1];
  final_sum<1>(tot, S.partial, nblk, s_red);
  if (threadIdx.x == 0) {
    st->counter[3] = 0;
    st->bnorm2 = tot[0];
    const double a = st->rtol * st->rtol * tot[0], c = st->atol * st->atol;
    st->stop2 = a > c ? a : c;
  }
}

// ---- fill (graph-safe replacement for memset) ---------------------------------------
template <typename T>
__global__ __launch_bounds__(VB) void k_zero_slot(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap, int slot) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  if (skip_sys(S.st->flags, SF_DONE | SF_INACTIVE, 0)) return;
  cplx<T>* __restrict__ v = reinterpret_cast<cplx<T>*>(S.vec[slot]);
  const long long n = S.n3;
  const long long stride = (long long)S.nblk_vec * VB;
  for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) v[e] = zero<T>();
}

// ---- diagonal scaling p of the symmetrised system ------------------------------------
// p_x = se_x[i] sh_y[j] sh_z[k], p_y = sh_x[i] se_y[j] sh_z[k], p_z = sh_x[i] sh_y[j] se_z[k]
__device__ __forceinline__ cplx<double> pscale(const AxisCoef<double>* coef, int nx, int ny, int nz, long long e) {
  const long long ncell = (long long)nx * ny * nz;
  const int c = (int)(e / ncell);
  long long cell = e - c * ncell;
  const int k = (int)(cell % nz);
  cell /= nz;
  const int j = (int)(cell % ny);
  const int i = (int)(cell / ny);
  const AxisCoef<double> ax = coef[i], ay = coef[nx + j], az = coef[nx + ny + k];
  const cplx<double> sx = c == 0 ? ax.se : ax.sh;
  const cplx<double> sy = c == 1 ? ay.se : ay.sh;
  const cplx<double> sz = c == 2 ? az.se : az.sh;
  return sx * sy * sz;
}

// mode 0: out = p * (coef * in)            (forward b = Pl(-i w J), forward x0 = Pl E)
// mode 1: out = (1/p) * conj(coef * in)    (adjoint b = Pr^H(-i w J) conjugated, x0 likewise)
// mode 2: out = in / p                     (forward E = Pr x)
// mode 3: out = conj(p * in)               (adjoint E = Pl^H z with z = conj(x))
__global__ __launch_bounds__(VB) void k_diag(const AxisCoef<double>* __restrict__ coef, int nx, int ny, int nz,
                                            const cplx<double>* __restrict__ in, cplx<double>* __restrict__ out,
                                            double cre, double cim, int mode) {
  const long long n = 3LL * nx * ny * nz;
  const cplx<double> k{cre, cim};
  for (long long e = (long long)blockIdx.x * VB + threadIdx.x; e < n; e += (long long)gridDim.x * VB) {
    const cplx<double> p = pscale(coef, nx, ny, nz, e);
    const cplx<double> v = in[e];
    cplx<double> o;
    switch (mode) {
      case 0: o = p * (k * v); break;
      case 1: o = cdiv(conj(k * v), p); break;
      case 2: o = cdiv(v, p); break;
      default: o = conj(p * v); break;
    }
    out[e] = o;
  }
}

// ---- host launch helpers ----------------------------------------------------------
#define MX_VEC_LAUNCH(name, kern)                                                              \
  hipError_t name(const SysDesc* sys, const BlockMap* bmap, int nb, hipStream_t s) {           \
    if (nb <= 0) return hipSuccess;                                                            \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(VB), 0, s, sys, bmap);                             \
    return hipGetLastError();                                                                  \
  }
MX_VEC_LAUNCH(launch_update_f32, k_update<float>)
MX_VEC_LAUNCH(launch_update_f64, k_update<double>)
MX_VEC_LAUNCH(launch_accum_f32, k_accum<float>)
MX_VEC_LAUNCH(launch_accum_f64, k_accum<double>)
MX_VEC_LAUNCH(launch_bnorm, k_bnorm)
#undef MX_VEC_LAUNCH

hipError_t launch_pupdate_f32(const OpView<float>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pupdate<float>, dim3(nb), dim3(VB), 0, s, ops, sys, bmap);
  return hipGetLastError();
}
hipError_t launch_pupdate_f64(const OpView<double>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pupdate<double>, dim3(nb), dim3(VB), 0, s, ops, sys, bmap);
  return hipGetLastError();
}

hipError_t launch_zero_slot(int prec64, const SysDesc* sys, const BlockMap* bmap, int nb, int slot, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  if (prec64) hipLaunchKernelGGL(k_zero_slot<double>, dim3(nb), dim3(VB), 0, s, sys, bmap, slot);
  else hipLaunchKernelGGL(k_zero_slot<float>, dim3(nb), dim3(VB), 0, s, sys, bmap, slot);
  return hipGetLastError();
}

hipError_t launch_diag(const AxisCoef<double>* coef, int nx, int ny, int nz, const void* in, void* out,
                       double cre, double cim, int mode, hipStream_t s) {
  const long long n = 3LL * nx * ny * nz;
  long long nb = (n + VB - 1) / VB;
  if (nb > 4096) nb = 4096;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_diag, dim3((unsigned)nb), dim3(VB), 0, s, coef, nx, ny, nz,
                     reinterpret_cast<const cplx<double>*>(in), reinterpret_cast<cplx<double>*>(out), cre, cim, mode);
  return hipGetLastError();
}

}  // namespace mx
