// Fused Krylov vector kernels for complex-symmetric COCG with residual replacement,
// plus the diagonal maps J -> b and x -> E of meanas.fdfd.solvers.generic
// (fdfd/solvers.py:115,124-125,136-139).
//
// COCG (van der Vorst & Melissen 1990) for A = A^T uses the unconjugated products
// rho = r^T r and mu = p^T A p.  Two forms of the iteration are implemented; each
// system carries its form in SF_FUSED.
//
// Fused form (complex128 inner vectors start in it): two streaming launches, each
// followed by a one-block-per-system finishing kernel:
//   k_apply   : q = A p; mu = p^T q and q^T q                        (apply.hip)
//   k_fin<0>  : alpha = rho / mu; beta = (alpha^2 q^T q - rho) / rho
//   k_update  : x_lo += alpha p, r -= alpha q, p = r + beta p; rho' = r^T r, ||r||^2
//   k_fin<1>  : rho = rho', convergence and replacement checks
// x_lo is only needed at a replacement, so every other pass leaves it alone: a pass
// with beta away from zero defers its alpha p, which the next pass recovers as
// (alpha / beta) (p' - r') and adds together with its own term (8 vector streams per
// iteration on average instead of 9; k_accum flushes a pending term).
// beta is the value that makes the new p A-conjugate to the old one: with r^T A p =
// p^T A p (conjugacy of the previous direction), r'^T r' = rho - 2 alpha mu + alpha^2
// q^T q = alpha^2 q^T q - rho.  Knowing beta before r' is reduced lets the direction
// update ride in the same pass as x and r: 9 vector streams per iteration instead of
// 11 (Chronopoulos & Gear 1989 use the same identity).  The exact rho' still comes
// from the stored r' and is what alpha uses.  Batches whose systems all stay fused
// (mx_solve in complex128) launch k_fin<1> only before a replacement point: in between,
// k_fin<4> finishes the previous update and the apply together, so an iteration is
// k_apply, k_fin<4>, k_update.  The state evolves exactly as with separate launches; an
// apply that follows the update which paused its system is discarded.
//
// Classic form (complex64 inner vectors; complex128 systems whose recursion lost track):
// the prediction above cancels (|r^T r| is often 1e-3 ||r||^2 for these complex-symmetric
// operators), which complex64 cannot afford, and its accuracy floor is worse on nearly
// singular operators (near-static omega).  beta = rho' / rho from the stored r' and p
// is updated in its own pass:
//   k_pupdate : p = r + beta p              (p = r after a restart)
//   k_apply, k_fin<0> (alpha), k_update (x_lo, r, rho', ||r||^2), k_fin<1> (beta)
//
// Residual replacement (Sleijpen & van der Vorst 1996 "reliable updates"; the mixed
// precision form of Clark et al. 2010): when the recursive ||r|| has dropped by
// `delta` since the last replacement, reached the target, or `refresh_iters` passed,
// the system pauses and
//   k_accum   : x += x_lo (FP64), x_lo = 0
//   k_apply<RESID> : r = b - A x in FP64, ||r||^2 in FP64, rho from the stored r
// replace the recursive residual with the true one and keep the search direction:
// the fused form adds r_true - r_recursive to p in the same pass (k_pupdate then only
// runs for a restart or a rescue step); the classic form sets beta = rho'/rho_old for
// the next k_pupdate.  A fused system whose true residual at a replacement exceeds 10x
// the recursive one switches to the classic form (with a restart): in place, or (with
// SF_SWITCH_EXIT, as mx_solve runs it) by finishing so that the host continues it in a
// classic-form solve; then the fused iteration needs no direction-update launch.  Every
// scalar is reduced in a fixed order (per-block partials, then one finishing block per
// system), so results do not depend on block scheduling or on the batch.
#include "common.hpp"

namespace mx {

constexpr int VB = 256;   // threads per vector block
// x_lo and r are read and written with streaming (non-temporal) accesses in k_update:
// they are next touched one iteration later, after about 700 MB of other traffic.  That
// keeps them out of the 256 MB Infinity Cache, which then holds p, q and the materials
// (about 220 MB at 128^3) from the kernel that writes them to the kernels that read
// them: at 128^3 the apply drops from 73.5 to 56 us and the update from 119 to 106 us.
#ifndef MX_STREAM_XR
#define MX_STREAM_XR 1
#endif
constexpr bool kStreamXR = MX_STREAM_XR != 0;   // streaming accesses to x_lo and r in k_update
#ifndef MX_VU
#define MX_VU 4
#endif
constexpr int VU = MX_VU;   // 16-byte packs in flight per thread

__device__ __forceinline__ bool skip_sys(int f, int skip, int need) {
  return (f & skip) || (need && !(f & need));
}

// 16-byte pack of complex values: 2 complex64 or 1 complex128 (one dwordx4 access),
// loaded and stored through global-address-space pointers (see common.hpp).
template <typename T> struct alignas(16) Pack {
  static constexpr int K = 16 / sizeof(cplx<T>);
  cplx<T> v[K];
};
template <typename T, bool NT = false>
__device__ __forceinline__ Pack<T> ld(const void* base, long long i) {
  const typename PackV<T>::type w = NT ? pload_nt<T>(base, i) : pload<T>(base, i);
  Pack<T> p;
  if constexpr (Pack<T>::K == 2) {
    p.v[0] = cplx<T>{w[0], w[1]};
    p.v[1] = cplx<T>{w[2], w[3]};
  } else {
    p.v[0] = cplx<T>{w[0], w[1]};
  }
  return p;
}
template <typename T, bool NT = false>
__device__ __forceinline__ void stp(void* base, long long i, const Pack<T>& p) {
  typename PackV<T>::type w;
  if constexpr (Pack<T>::K == 2) {
    w[0] = p.v[0].re; w[1] = p.v[0].im; w[2] = p.v[1].re; w[3] = p.v[1].im;
  } else {
    w[0] = p.v[0].re; w[1] = p.v[0].im;
  }
  if constexpr (NT) pstore_nt<T>(base, i, w);
  else pstore<T>(base, i, w);
}

// Contiguous pack range [*b, *e) of this block: the vector is split into nblk equal runs.
__device__ __forceinline__ void block_range(long long npk, int local, int nblk, long long* b, long long* e) {
  const long long per = (npk + nblk - 1) / nblk;
  *b = min(npk, per * local);
  *e = min(npk, *b + per);
}

// Pack range of one loop trip of this thread: packs base + u*VB, u < VU, below pe.
// Every kernel below loads all its in-flight packs first, then computes and stores, so
// VU loads per vector are outstanding at once.
struct Trip {
  long long base, pe;
  __device__ __forceinline__ bool ok(int u) const { return base + (long long)u * VB < pe; }
  __device__ __forceinline__ long long at(int u) const { return base + (long long)u * VB; }
};

// ---- search direction -------------------------------------------------------------
// Fused form: only after a replacement (post = 1), for a restart (p = r) or a rescue
// step (p = g).  Classic form: at the start of every iteration (post = 0), p = r + beta p
// (or p = r, or p = g).
template <typename T>
__global__ __launch_bounds__(VB) void k_pupdate(const OpView<T>* __restrict__ ops, const SysDesc* __restrict__ systems,
                                              const BlockMap* __restrict__ bmap, int post) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  const int f = gload1<int>(&S.st->flags, 0);
  if (skip_sys(f, SF_PAUSE | SF_DONE | SF_INACTIVE, 0)) return;
  const bool fused = (f & SF_FUSED) != 0;
  if (fused ? (!post || !(f & (SF_MR | SF_RESTART))) : post) return;
  const cplx<T>* __restrict__ r = reinterpret_cast<const cplx<T>*>(S.vec[VS_R]);
  cplx<T>* __restrict__ p = reinterpret_cast<cplx<T>*>(S.vec[VS_P]);
  const long long n = S.n3;
  if (f & SF_MR) {   // breakdown rescue: p = hashed direction, zero on PEC entries
    const OpView<T>& op = ops[S.op];
    const long long ncell = n / 3;
    const long long stride = (long long)S.nblk_vec * VB;
    for (long long e = (long long)bm.local * VB + threadIdx.x; e < n; e += stride) {
      const int c = (int)(e / ncell);
      const bool pec = op.mask && ((op.mask[e - c * ncell] >> c) & 1);
      p[e] = pec ? zero<T>() : cast<T>(hashed_direction(e));
    }
    return;
  }
  constexpr int K = Pack<T>::K;
  long long pb, pe;
  block_range(n / K, bm.local, S.nblk_vec, &pb, &pe);
  const bool restart = (f & SF_RESTART) != 0;
  const cplx<T> beta = restart ? zero<T>() : cast<T>(gload<double>(&S.st->beta, 0));
  for (Trip t{pb + threadIdx.x, pe}; t.base < pe; t.base += (long long)VB * VU) {
    Pack<T> a[VU], c[VU];
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) {
        a[u] = ld<T>(r, t.at(u));
        if (!restart) c[u] = ld<T>(p, t.at(u));
      }
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) {
        if (!restart) {
#pragma unroll
          for (int k = 0; k < K; ++k) a[u].v[k] = a[u].v[k] + beta * c[u].v[k];
        }
        stp<T>(p, t.at(u), a[u]);
      }
  }
  if (n % K && bm.local == S.nblk_vec - 1 && threadIdx.x == 0)
    p[n - 1] = restart ? r[n - 1] : r[n - 1] + beta * p[n - 1];
}

// ---- x_lo += alpha p ; r -= alpha q ; [p = r + beta p] ; rho, ||r||^2 ----------------
// kFused: the direction update rides along; x_lo += xc1 p + xc2 r (r before this pass's
// update; xc1 = alpha, xc2 = 0 unless a deferred term is added), or with kSkipX x_lo is
// left alone (the fused form's deferred solution update).  Classic: x_lo += alpha p.
template <typename T, bool kFused, bool kSkipX = false>
__device__ __forceinline__ void update_body(const SysDesc& S, const BlockMap bm) {
  __shared__ double s_red[3 * 16];
  SysState* st = S.st;
  cplx<T>* __restrict__ x = reinterpret_cast<cplx<T>*>(S.vec[VS_XLO]);
  cplx<T>* __restrict__ r = reinterpret_cast<cplx<T>*>(S.vec[VS_R]);
  cplx<T>* __restrict__ p = reinterpret_cast<cplx<T>*>(S.vec[VS_P]);
  const cplx<T>* __restrict__ q = reinterpret_cast<const cplx<T>*>(S.vec[VS_Q]);
  const cplx<T> alpha = cast<T>(gload<double>(&st->alpha, 0));
  const cplx<T> beta = kFused ? cast<T>(gload<double>(&st->beta, 0)) : zero<T>();
  const cplx<T> xc1 = kFused ? cast<T>(gload<double>(&st->xc1, 0)) : alpha;
  const cplx<T> xc2 = kFused ? cast<T>(gload<double>(&st->xc2, 0)) : zero<T>();
  const long long n = S.n3;
  T rr_re = 0, rr_im = 0, nr = 0;
  auto one = [&](cplx<T>& xv, cplx<T>& rv, cplx<T>& pv, cplx<T> qv) {
    if constexpr (!kSkipX) {
      if constexpr (kFused) xv = xv + xc1 * pv + xc2 * rv;
      else xv = xv + alpha * pv;
    }
    rv = rv - alpha * qv;
    if constexpr (kFused) pv = rv + beta * pv;
    rr_re += rv.re * rv.re - rv.im * rv.im;
    rr_im += T(2) * rv.re * rv.im;
    nr += rv.re * rv.re + rv.im * rv.im;
  };
  constexpr int K = Pack<T>::K;
  long long pb, pe;
  block_range(n / K, bm.local, S.nblk_vec, &pb, &pe);
  for (Trip t{pb + threadIdx.x, pe}; t.base < pe; t.base += (long long)VB * VU) {
    Pack<T> xv[VU], rv[VU], pv[VU], qv[VU];
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) {
        if constexpr (!kSkipX) xv[u] = ld<T, kStreamXR>(x, t.at(u));
        rv[u] = ld<T, kStreamXR>(r, t.at(u));
        pv[u] = ld<T>(p, t.at(u));
        qv[u] = ld<T>(q, t.at(u));
      }
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) {
#pragma unroll
        for (int k = 0; k < K; ++k) one(xv[u].v[k], rv[u].v[k], pv[u].v[k], qv[u].v[k]);
        // x and r are next read one iteration later; p is read by the next kernel
        if constexpr (!kSkipX) stp<T, kStreamXR>(x, t.at(u), xv[u]);
        stp<T, kStreamXR>(r, t.at(u), rv[u]);
        if constexpr (kFused) stp<T>(p, t.at(u), pv[u]);
      }
  }
  if (n % K && bm.local == S.nblk_vec - 1 && threadIdx.x == 0) {
    cplx<T> xe = kSkipX ? zero<T>() : x[n - 1], re = r[n - 1], pe1 = p[n - 1];
    one(xe, re, pe1, q[n - 1]);
    if constexpr (!kSkipX) x[n - 1] = xe;
    r[n - 1] = re;
    if constexpr (kFused) p[n - 1] = pe1;
  }
  double acc[3] = {(double)rr_re, (double)rr_im, (double)nr};
  block_sum<3>(acc, s_red);
  store_partials<3>(acc, S.partial_vec, bm.local, S.nblk_vec);
}

template <typename T>
__global__ __launch_bounds__(VB) void k_update(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  const int f = gload1<int>(&S.st->flags, 0);
  if (skip_sys(f, SF_PAUSE | SF_DONE | SF_INACTIVE, 0)) return;
  if (sizeof(T) == 8 && (f & SF_FUSED)) {   // the direction update rides along
    if (f & SF_XSKIP) update_body<T, true, true>(S, bm);
    else update_body<T, true, false>(S, bm);
  } else {
    update_body<T, false>(S, bm);
  }
}

// ---- x (c128) += x_lo ; x_lo = 0   (paused systems) --------------------------------
template <typename T>
__global__ __launch_bounds__(VB) void k_accum(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  const int f = S.st->flags;
  if (skip_sys(f, SF_DONE | SF_INACTIVE, SF_PAUSE)) return;
  cplx<double>* __restrict__ x = reinterpret_cast<cplx<double>*>(S.vec[VS_X]);
  cplx<T>* __restrict__ xl = reinterpret_cast<cplx<T>*>(S.vec[VS_XLO]);
  const long long n = S.n3;
  if (sizeof(T) == 8 && (f & SF_XPEND)) {
    // the fused form's deferred alpha p = xa (p - r), recovered before the replacement
    // overwrites r and p: x += x_lo + xa (p - r), x_lo = 0
    const cplx<double> xa = gload<double>(&S.st->xa, 0);
    const cplx<double>* __restrict__ pp = reinterpret_cast<const cplx<double>*>(S.vec[VS_P]);
    const cplx<double>* __restrict__ rr = reinterpret_cast<const cplx<double>*>(S.vec[VS_R]);
    cplx<double>* __restrict__ xl64 = reinterpret_cast<cplx<double>*>(S.vec[VS_XLO]);
    long long pb, pe;
    block_range(n, bm.local, S.nblk_vec, &pb, &pe);
    for (long long e = pb + threadIdx.x; e < pe; e += VB) {
      const cplx<double> d = xa * (gload<double>(pp, e) - gload<double>(rr, e));
      gstore<double>(x, e, gload<double>(x, e) + (gload<double>(xl64, e) + d));
      gstore<double>(xl64, e, zero<double>());
    }
    return;
  }
  constexpr int K = Pack<T>::K;
  long long pb, pe;
  block_range(n / K, bm.local, S.nblk_vec, &pb, &pe);
  Pack<T> z;
#pragma unroll
  for (int k = 0; k < K; ++k) z.v[k] = zero<T>();
  for (Trip t{pb + threadIdx.x, pe}; t.base < pe; t.base += (long long)VB * VU) {
    Pack<T> a[VU];
    cplx<double> xv[VU][K];
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) {
        a[u] = ld<T>(xl, t.at(u));
#pragma unroll
        for (int k = 0; k < K; ++k) xv[u][k] = gload<double>(x, t.at(u) * K + k);
      }
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) {
#pragma unroll
        for (int k = 0; k < K; ++k) gstore<double>(x, t.at(u) * K + k, xv[u][k] + cast<double>(a[u].v[k]));
        stp<T>(xl, t.at(u), z);
      }
  }
  if (n % K && bm.local == S.nblk_vec - 1 && threadIdx.x == 0) {
    x[n - 1] = x[n - 1] + cast<double>(xl[n - 1]);
    xl[n - 1] = zero<T>();
  }
}

// ---- x_best = x after a replacement that found a new best true residual -----------------
__global__ __launch_bounds__(VB) void k_save_best(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  if (skip_sys(S.st->flags, SF_DONE | SF_INACTIVE, SF_SAVE)) return;
  const cplx<double>* __restrict__ x = reinterpret_cast<const cplx<double>*>(S.vec[VS_X]);
  cplx<double>* __restrict__ xb = reinterpret_cast<cplx<double>*>(S.vec[VS_XBEST]);
  const long long n = S.n3;
  long long pb, pe;
  block_range(n, bm.local, S.nblk_vec, &pb, &pe);
  for (Trip t{pb + threadIdx.x, pe}; t.base < pe; t.base += (long long)VB * VU) {
    cplx<double> v[VU];
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) v[u] = gload<double>(x, t.at(u));
#pragma unroll
    for (int u = 0; u < VU; ++u)
      if (t.ok(u)) gstore<double>(xb, t.at(u), v[u]);
  }
}

// ---- ||b||^2 and the stopping target, once per solve --------------------------------
__global__ __launch_bounds__(VB) void k_bnorm(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap) {
  __shared__ double s_red[1 * 16];
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  SysState* st = S.st;
  if (skip_sys(st->flags, SF_DONE | SF_INACTIVE, 0)) return;
  const cplx<double>* __restrict__ b = reinterpret_cast<const cplx<double>*>(S.vec[VS_B]);
  const long long n = S.n3;
  double acc[1] = {0.0};
  long long pb, pe;
  block_range(n, bm.local, S.nblk_vec, &pb, &pe);
  for (long long e = pb + threadIdx.x; e < pe; e += VB) acc[0] += norm2(gload<double>(b, e));
  block_sum<1>(acc, s_red);
  store_partials<1>(acc, S.partial, bm.local, S.nblk_vec);
}

// ---- fill (graph-safe replacement for memset) ---------------------------------------
template <typename T>
__global__ __launch_bounds__(VB) void k_zero_slot(const SysDesc* __restrict__ systems, const BlockMap* __restrict__ bmap, int slot) {
  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  if (skip_sys(S.st->flags, SF_DONE | SF_INACTIVE, 0)) return;
  cplx<T>* __restrict__ v = reinterpret_cast<cplx<T>*>(S.vec[slot]);
  const long long n = S.n3;
  constexpr int K = Pack<T>::K;
  long long pb, pe;
  block_range(n / K, bm.local, S.nblk_vec, &pb, &pe);
  Pack<T> z;
#pragma unroll
  for (int k = 0; k < K; ++k) z.v[k] = zero<T>();
  for (long long i = pb + threadIdx.x; i < pe; i += VB) stp<T>(v, i, z);
  if (n % K && bm.local == S.nblk_vec - 1 && threadIdx.x == 0) v[n - 1] = zero<T>();
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
// mode 4: out = s * (coef * in), s = +-1 the sign that turns the factored p into the
//         principal root sqrt(p^2) the reference uses (fdfd/operators.py:120): +1 when
//         p lies in the principal half-plane (Re p > 0, or Re p = 0 and Im p >= 0).
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
      case 3: o = conj(p * v); break;
      default: {
        const bool principal = p.re > 0.0 || (p.re == 0.0 && p.im >= 0.0);
        o = principal ? k * v : cplx<double>{-(k * v).re, -(k * v).im};
        break;
      }
    }
    out[e] = o;
  }
}


// ---- finishing kernels: one block per system sums that system's partials in a fixed
// order and advances its scalar state.  They follow the reducing kernel of the same
// kind and apply the same skip rule, so partials are never read stale.
// FIN_APPLY_MERGED (fused FP64 batches) first finishes the previous update, whose
// partials sit in their own array, then the apply: one launch per iteration instead of
// two.  The apply that ran in between is discarded when that update pauses the system.

// alpha (and the fused form's beta) from p^T A p and q^T q; tot = {mu re, im, qtq re, im}
// ({., ., ||A g||^2, .} for a rescue step)
__device__ __forceinline__ void fin_apply_state(SysState* st, int& f, const double (&tot)[4], bool fused) {
  if (f & SF_MR) {
    // rescue step along the hashed direction p = g: alpha = ||r|| / ||A g|| makes the
    // new residual r - alpha A g generic (r^T r no longer ~ 0); beta = 0 restarts
    // COCG from that residual.
    const double qq = tot[2];
    const cplx<double> alpha{sqrt(st->rnorm2 / qq), 0.0};
    if (!(qq > 0.0) || !finite(alpha.re)) f |= SF_BREAKDOWN | SF_PAUSE | SF_LIMIT;   // A g = 0
    else {
      st->alpha = alpha;
      if (fused) {   // beta = 0 restarts the fused recurrence (classic: k_pupdate does);
        st->beta = cplx<double>{0.0, 0.0};   // the step is taken, so the next apply is
        f &= ~(SF_RESTART | SF_MR);           // an ordinary one
        st->xc1 = alpha;                      // x_lo += alpha g now (a replacement has
        st->xc2 = cplx<double>{0.0, 0.0};     // just flushed any deferred term)
        f &= ~SF_XSKIP;
      }
    }
  } else {
    const cplx<double> mu{tot[0], tot[1]}, qtq{tot[2], tot[3]};
    st->mu = mu;
    const cplx<double> rho = st->rho;
    const cplx<double> alpha = cdiv(rho, mu);
    if (!(norm2(mu) > 0.0) || !finite(alpha.re) || !finite(alpha.im)) {
      // p^T A p vanished (or overflowed): replace the residual, then rescue with a
      // step along a hashed direction and restart from p = r
      f |= SF_BREAKDOWN | SF_PAUSE | SF_RESTART | SF_MR;
      if (!finite(mu.re) || !finite(mu.im)) f |= SF_NAN | SF_LIMIT;
    } else {
      if (fused) {   // predicted rho' = alpha^2 q^T q - rho; beta = rho' / rho
        cplx<double> beta = cdiv(alpha * alpha * qtq, rho) - cplx<double>{1.0, 0.0};
        if (!finite(beta.re) || !finite(beta.im)) beta = cplx<double>{0.0, 0.0};
        st->beta = beta;
        // Deferred solution update: x_lo needs alpha p of every pass, but only at a
        // replacement.  A pass may leave x_lo alone when the next one can recover that
        // term: the pass forms p' = r' + beta p, so alpha p = (alpha / beta) (p' - r').
        // The next pass then adds both terms at once, x_lo += (xa + alpha') p' - xa r'
        // (xa = alpha / beta), and x_lo is read and written every other pass.  A pending
        // term is flushed by k_accum before a replacement changes p and r.
        if (f & SF_XPEND) {
          st->xc1 = st->xa + alpha;
          st->xc2 = cplx<double>{-st->xa.re, -st->xa.im};
          f &= ~(SF_XPEND | SF_XSKIP);
        } else if (norm2(beta) >= kDeferBeta2) {
          st->xa = cdiv(alpha, beta);
          f |= SF_XSKIP | SF_XPEND;
        } else {
          st->xc1 = alpha;
          st->xc2 = cplx<double>{0.0, 0.0};
          f &= ~SF_XSKIP;
        }
      }
      st->alpha = alpha;
      f &= ~SF_RESTART;
    }
  }
  if (!(f & SF_PAUSE)) f |= SF_UPD_PENDING;   // the update runs next
}

// rho = r^T r, ||r||^2 of the updated residual and the checks; tot = {rho re, im, ||r||^2}
__device__ __forceinline__ void fin_update_state(SysState* st, int& f, const double (&tot)[3], bool fused) {
  const cplx<double> rho{tot[0], tot[1]};
  const double rn2 = tot[2];
  f &= ~(SF_SAVE | SF_UPD_PENDING);   // any saved best iterate was stored by k_save_best before this
  const bool rescued = (f & SF_MR) != 0;
  f &= ~SF_MR;       // a rescue step (if that was one) is done
  if (!fused) {      // classic: beta from the exact rho'; a rescue step restarts
    if (rescued) st->beta = cplx<double>{0.0, 0.0};
    else if (norm2(st->rho) > 0.0) st->beta = cdiv(rho, st->rho);
    else {           // rho breakdown: restart from the replaced residual
      f |= SF_BREAKDOWN | SF_PAUSE | SF_RESTART;
      st->beta = cplx<double>{0.0, 0.0};
    }
  }
  // r^T r ~ 0 with r != 0: the next step is a rescue step along a hashed direction.
  // The fused form has already built p from r, so it pauses for a replacement, after
  // which k_pupdate writes the rescue direction.
  if (norm2(rho) <= kRhoBreakdown * kRhoBreakdown * rn2 * rn2 && rn2 > 0.0)
    f |= SF_MR | SF_RESTART | (fused ? SF_PAUSE : 0);
  st->rho_old = st->rho;
  st->rho = rho;
  st->rnorm2 = rn2;
  if (rn2 > st->rmax2) st->rmax2 = rn2;
  st->iters += 1;
  st->since += 1;
  if (!finite(rn2) || !finite(rho.re) || !finite(rho.im)) {
    f |= SF_NAN | SF_PAUSE | SF_LIMIT;
  } else if (rn2 <= st->stop2 || rn2 <= st->delta2 * st->rmax2 ||
             (st->refresh_iters > 0 && st->since >= st->refresh_iters)) {
    f |= SF_PAUSE;
  }
  if (st->iters >= st->maxiter) f |= SF_PAUSE | SF_LIMIT;
}

// thread 0's part of k_fin<KIND>: the scalar state after the partial sums `tot`
template <int KIND>
__device__ __forceinline__ void finish_state(SysState* st, int f, const double (&tot)[4], bool fused) {
  if constexpr (KIND == FIN_APPLY) {
    fin_apply_state(st, f, tot, fused);
  } else if constexpr (KIND == FIN_UPDATE) {
    const double t3[3] = {tot[0], tot[1], tot[2]};
    fin_update_state(st, f, t3, fused);
  } else if constexpr (KIND == FIN_RESID) {
    // residual replacement / initial residual (k_accum has flushed any deferred term)
    f &= ~(SF_XPEND | SF_XSKIP);
    const double rn2 = tot[0];
    const cplx<double> rho{tot[1], tot[2]};
    st->true_rnorm2 = rn2;
    st->rnorm_rec2 = st->rnorm2;   // recursive residual the replacement overwrites
    st->rnorm2 = rn2;
    st->rmax2 = rn2;
    st->replacements += 1;
    st->since = 0;
    // stagnation: (a) the recursive residual had drifted >= 100x below the true one at
    // stall_limit consecutive replacements (the inner precision no longer tracks the
    // operator), or (b) the best true residual has not halved within stall_iters
    // iterations.  COCG's residual oscillates, so single replacements prove nothing.
    const bool first = st->replacements == 1;
    if (!first && st->rnorm_rec2 < 1e-4 * rn2) st->stalls += 1;
    else st->stalls = 0;
    if (first || rn2 < 0.25 * st->mark2) {
      st->mark2 = rn2;
      st->mark_iter = st->iters;
    }
    if (first || rn2 < st->best_true2) {
      st->best_true2 = rn2;
      f |= SF_SAVE;
    }
    // a fused recursion that lost track of the true residual continues in the classic
    // form, from a restart (p was built with the predicted beta)
    if (fused && !first && st->rnorm_rec2 < 1e-2 * rn2) {
      if (f & SF_SWITCH_EXIT) f |= SF_SWITCHED | SF_DONE;   // the host continues it in the classic form
      else f = (f & ~SF_FUSED) | SF_RESTART;
    }
    if (st->stall_limit > 0 && st->stalls >= st->stall_limit) f |= SF_STALL | SF_DONE;
    if (st->stall_iters > 0 && st->iters - st->mark_iter > st->stall_iters) f |= SF_STALL | SF_DONE;
    if (!finite(rn2)) f |= SF_NAN | SF_DONE;
    if (rn2 <= st->stop2 || (f & (SF_DONE | SF_LIMIT))) {
      f |= SF_DONE;
    } else {
      // fused: the replacement kernel already moved p onto the true residual (unless a
      // restart or a rescue step is pending: k_pupdate rebuilds p then).  Classic: the
      // next k_pupdate continues the direction with beta = rho_true / rho_old.
      if (!(f & SF_FUSED)) {
        const cplx<double> ro = st->rho_old;
        st->beta = (norm2(ro) > 0.0) ? cdiv(rho, ro) : cplx<double>{0.0, 0.0};
      }
      st->rho = rho;
      if (norm2(rho) <= kRhoBreakdown * kRhoBreakdown * rn2 * rn2) f |= SF_RESTART | SF_MR;
      f &= ~(SF_PAUSE | SF_BREAKDOWN);
    }
  } else {   // FIN_BNORM
    st->bnorm2 = tot[0];
    const double a = st->rtol * st->rtol * tot[0], c = st->atol * st->atol;
    st->stop2 = a > c ? a : c;
  }
  st->flags = f;
}

template <int KIND>
__global__ __launch_bounds__(kFinThreads) void k_fin(const SysDesc* __restrict__ systems, const int* __restrict__ syslist) {
  __shared__ double s_red[8 * 16];
  __shared__ SysState s_st;
  const SysDesc& S = systems[syslist[blockIdx.x]];
  SysState* const gst = S.st;
  // The state (one 8-byte word per lane of wave 0) and the partial sums are read in one
  // batch; thread 0 updates the state in LDS and it is written back in one batch.  The
  // skip rule is applied once both are in: a skipped system's sums are dropped.  (Read
  // first and updated in place through the global pointer, the flags and thread 0's
  // loads after its own stores formed a chain of about ten memory round trips.)
  constexpr int kWords = (int)(sizeof(SysState) / 8);
  static_assert(sizeof(SysState) % 8 == 0 && kWords <= 64, "state words: one per lane of a wave");
  const int tid = threadIdx.x;
  unsigned long long word = 0;
  if (tid < kWords) word = reinterpret_cast<const unsigned long long*>(gst)[tid];
  double tot[4] = {0.0, 0.0, 0.0, 0.0}, tu[3] = {0.0, 0.0, 0.0};
  if constexpr (KIND == FIN_APPLY_MERGED) {
    final_sum2<3, 4>(tu, S.partial_vec, S.nblk_vec, tot, S.partial, S.nblk_apply, s_red);
  } else if constexpr (KIND == FIN_BNORM) {
    double t1[1];
    final_sum<1>(t1, S.partial, S.nblk_vec, s_red);
    tot[0] = t1[0];
  } else if constexpr (KIND == FIN_APPLY) {
    final_sum<4>(tot, S.partial, S.nblk_apply, s_red);
  } else {
    double t3[3];
    final_sum<3>(t3, KIND == FIN_UPDATE ? S.partial_vec : S.partial, KIND == FIN_UPDATE ? S.nblk_vec : S.nblk_apply,
                 s_red);
    tot[0] = t3[0]; tot[1] = t3[1]; tot[2] = t3[2];
  }
  if (tid < kWords) reinterpret_cast<unsigned long long*>(&s_st)[tid] = word;
  __syncthreads();
  int f = s_st.flags;
  if (KIND == FIN_APPLY || KIND == FIN_APPLY_MERGED) {
    if (skip_sys(f, SF_PAUSE | SF_DONE | SF_INACTIVE, 0)) return;
  } else if (KIND == FIN_UPDATE) {
    if (skip_sys(f, SF_PAUSE | SF_DONE | SF_INACTIVE, SF_UPD_PENDING)) return;
  } else if (KIND == FIN_RESID) {
    if (skip_sys(f, SF_DONE | SF_INACTIVE, SF_PAUSE)) return;
  } else {
    if (skip_sys(f, SF_DONE | SF_INACTIVE, 0)) return;
  }
  const bool fused = (f & SF_FUSED) != 0;
  SysState* const st = &s_st;
  if (tid == 0) {
    if constexpr (KIND == FIN_APPLY_MERGED) {
      // the previous update's sums count only while its finishing is pending
      if (f & SF_UPD_PENDING) fin_update_state(st, f, tu, fused);
      if (!(f & SF_PAUSE)) fin_apply_state(st, f, tot, fused);
      st->flags = f;
    } else {
      finish_state<KIND>(st, f, tot, fused);
    }
  }
  __syncthreads();
  if (tid < kWords) reinterpret_cast<unsigned long long*>(gst)[tid] = reinterpret_cast<unsigned long long*>(&s_st)[tid];
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
MX_VEC_LAUNCH(launch_save_best, k_save_best)
#undef MX_VEC_LAUNCH

// post = 0: the iteration's direction update (classic form); post = 1: after a
// residual replacement (restart or rescue step of the fused form)
hipError_t launch_pupdate_f32(const OpView<float>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, int post,
                              hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pupdate<float>, dim3(nb), dim3(VB), 0, s, ops, sys, bmap, post);
  return hipGetLastError();
}
hipError_t launch_pupdate_f64(const OpView<double>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, int post,
                              hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pupdate<double>, dim3(nb), dim3(VB), 0, s, ops, sys, bmap, post);
  return hipGetLastError();
}

hipError_t launch_fin(int kind, const SysDesc* sys, const int* syslist, int nsys, hipStream_t s) {
  if (nsys <= 0) return hipSuccess;
  switch (kind) {
    case FIN_APPLY: hipLaunchKernelGGL(k_fin<FIN_APPLY>, dim3(nsys), dim3(kFinThreads), 0, s, sys, syslist); break;
    case FIN_UPDATE: hipLaunchKernelGGL(k_fin<FIN_UPDATE>, dim3(nsys), dim3(kFinThreads), 0, s, sys, syslist); break;
    case FIN_RESID: hipLaunchKernelGGL(k_fin<FIN_RESID>, dim3(nsys), dim3(kFinThreads), 0, s, sys, syslist); break;
    case FIN_APPLY_MERGED:
      hipLaunchKernelGGL(k_fin<FIN_APPLY_MERGED>, dim3(nsys), dim3(kFinThreads), 0, s, sys, syslist);
      break;
    default: hipLaunchKernelGGL(k_fin<FIN_BNORM>, dim3(nsys), dim3(kFinThreads), 0, s, sys, syslist); break;
  }
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
