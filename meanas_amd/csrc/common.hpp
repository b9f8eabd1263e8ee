// Shared device-side types for the FDFD solve library (gfx950).
//
// Vector layout (all field vectors): C order (3, nx, ny, nz), z fastest, complex
// interleaved (re, im) -- the layout of meanas.fdmath.vec (fdmath/vectorization.py:51-68).
//
// The operator applied is the symmetrised wave operator of meanas.fdfd.solvers.generic
// (fdfd/solvers.py:118-125):  A = Pl A0 Pr  with A0 = e_full (fdfd/operators.py:40-96)
// and Pl = diag(p), Pr = diag(1/p) (fdfd/operators.py:99-123).  With p factored per
// axis, p_x = se_x[i] sh_y[j] sh_z[k] (se = sqrt(dx_e), sh = sqrt(dx_h); p_y, p_z
// likewise), A takes the manifestly symmetric form
//     A = pe ( C^T M C - w^2 eps ) pe ,   M = pm diag(1/mu)
// where C is the forward curl built from the 1D scaled differences
//     (D_a f)[n] = fa_a[n] f[n+1] - fb_a[n] f[n],
//     fa_a[n] = 1 / (se_a[n] sh_a[n+1 mod N]),  fb_a[n] = 1 / (se_a[n] sh_a[n]),
// and C^T uses  (-D_a^T g)[n] = fb_a[n] g[n] - fa_a[n-1] g[n-1].  So the stencil reads
// the unknown directly; no per-cell scaling is needed.  p differs from the reference's
// sqrt(p^2) by at most a per-cell sign, a +-1 similarity that leaves E unchanged.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mx {

template <typename T> struct cplx { T re, im; };

template <typename T> __host__ __device__ __forceinline__ cplx<T> mk(T r, T i) { return {r, i}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return {s * a.re, s * a.im}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return {a.re, -a.im}; }
template <typename T> __host__ __device__ __forceinline__ T norm2(cplx<T> a) { return a.re * a.re + a.im * a.im; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> zero() { return {T(0), T(0)}; }
template <typename D, typename S> __host__ __device__ __forceinline__ cplx<D> cast(cplx<S> a) { return {D(a.re), D(a.im)}; }
// a*b - c*d  (complex), the shape of every scaled difference
template <typename T> __device__ __forceinline__ cplx<T> diff2(cplx<T> a, cplx<T> x1, cplx<T> b, cplx<T> x0) {
  return a * x1 - b * x0;
}
__host__ __device__ __forceinline__ cplx<double> cdiv(cplx<double> a, cplx<double> b) {
  double d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// ---- global-address-space access --------------------------------------------------
// Pointers read out of descriptor tables are generic (flat) to the compiler; flat loads
// count against lgkmcnt, so every LDS wait would also wait for in-flight prefetches, and
// they need 64-bit VGPR addresses.  These helpers access memory through address-space-1
// pointers (device pass) so the compiler emits global_load/store with an SGPR base and a
// 32-bit VGPR offset.  Complex values move as native 2-vectors.
#if defined(__HIP_DEVICE_COMPILE__)
#define MX_GLOBAL __attribute__((address_space(1)))
#else
#define MX_GLOBAL
#endif
// Read-only tables (PML coefficients) indexed by a value uniform over the wave or the
// block go through constant-address-space pointers, so the compiler can load them with
// scalar loads into SGPRs instead of holding a copy per lane.  (Loads only: nothing here
// writes through the scalar path.)
#if defined(__HIP_DEVICE_COMPILE__)
#define MX_CONST __attribute__((address_space(4)))
#else
#define MX_CONST
#endif
template <typename T> struct Vec2 { typedef T __attribute__((ext_vector_type(2))) type; };
template <typename T> struct Vec4 { typedef T __attribute__((ext_vector_type(4))) type; };

// Byte offsets are 32-bit unsigned (the host keeps every vector below 4 GiB), which
// lets the compiler use the SGPR-base + 32-bit VGPR-offset addressing form.
template <typename T>
__device__ __forceinline__ cplx<T> gload(const void* base, int idx) {
  typedef typename Vec2<T>::type V;
  const unsigned off = (unsigned)idx * (unsigned)sizeof(V);
  const V v = *(const MX_GLOBAL V*)((const MX_GLOBAL char*)base + off);
  return cplx<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void gstore(void* base, int idx, cplx<T> a) {
  typedef typename Vec2<T>::type V;
  const unsigned off = (unsigned)idx * (unsigned)sizeof(V);
  V v;
  v.x = a.re;
  v.y = a.im;
  *(MX_GLOBAL V*)((MX_GLOBAL char*)base + off) = v;
}
// Stores through a bounded buffer resource (raw, stride 0): the hardware drops a store
// whose byte offset lies at or beyond the resource's size, without touching memory.  A
// lane outside the grid can then issue its store with kDropOffset instead of branching
// around it, so every path through a loop issues the same stores.  That matters for the
// wait counter: vector loads and stores count on one in-order counter, and a store that
// may be missing on some path makes the compiler wait for every store of the previous
// step before it uses the next prefetched load (measured in the 128^3 solve's apply: the
// plane loop's wait went from vmcnt(0) to vmcnt(3), 8 fewer VGPRs, solve 0.2% faster).
// Resources here stay below 2 GiB.
constexpr unsigned kDropOffset = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bounded_rsrc(void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, 0x00020000);   // raw buffer, gfx9 word 3
}
template <typename T>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t rs, unsigned off, cplx<T> a) {
  typedef typename Vec2<T>::type V;
  V v;
  v.x = a.re;
  v.y = a.im;
  if constexpr (sizeof(T) == 8) {
    typedef unsigned __attribute__((ext_vector_type(4))) U;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U, v), rs, off, 0, 0);
  } else {
    typedef unsigned __attribute__((ext_vector_type(2))) U;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), rs, off, 0, 0);
  }
}
// complex value idx of a read-only table, through the constant address space; idx must
// be uniform over the wave for a scalar load
template <typename T>
__device__ __forceinline__ cplx<T> cload(const void* base, int idx) {
  typedef typename Vec2<T>::type V;
  const V v = ((const MX_CONST V*)base)[idx];
  return cplx<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ T gload1(const void* base, int idx) {
  const unsigned off = (unsigned)idx * (unsigned)sizeof(T);
  return *(const MX_GLOBAL T*)((const MX_GLOBAL char*)base + off);
}

// 16-byte pack i of a vector (2 complex64 or 1 complex128)
template <typename T> struct PackV;
template <> struct PackV<float> { typedef float __attribute__((ext_vector_type(4))) type; };
template <> struct PackV<double> { typedef double __attribute__((ext_vector_type(2))) type; };
template <typename T>
__device__ __forceinline__ typename PackV<T>::type pload(const void* base, long long i) {
  typedef typename PackV<T>::type V;
  const unsigned off = (unsigned)i * 16u;
  return *(const MX_GLOBAL V*)((const MX_GLOBAL char*)base + off);
}
// streaming load: for data not read again before much other traffic has passed
template <typename T>
__device__ __forceinline__ typename PackV<T>::type pload_nt(const void* base, long long i) {
  typedef typename PackV<T>::type V;
  const unsigned off = (unsigned)i * 16u;
  return __builtin_nontemporal_load((const MX_GLOBAL V*)((const MX_GLOBAL char*)base + off));
}
template <typename T>
__device__ __forceinline__ void pstore(void* base, long long i, typename PackV<T>::type v) {
  typedef typename PackV<T>::type V;
  const unsigned off = (unsigned)i * 16u;
  *(MX_GLOBAL V*)((MX_GLOBAL char*)base + off) = v;
}
// streaming store: for data that is not read again before much other traffic has
// passed (keeps it from displacing vectors the next kernel reads out of the caches)
template <typename T>
__device__ __forceinline__ void pstore_nt(void* base, long long i, typename PackV<T>::type v) {
  typedef typename PackV<T>::type V;
  const unsigned off = (unsigned)i * 16u;
  __builtin_nontemporal_store(v, (MX_GLOBAL V*)((MX_GLOBAL char*)base + off));
}

// Per-axis coefficients at index n: fa multiplies f[n+1], fb multiplies f[n] in the
// scaled forward difference; se = sqrt(dx_e[n]), sh = sqrt(dx_h[n]) give the diagonal
// scaling p used to map J -> b and x -> E.
template <typename T> struct AxisCoef { cplx<T> fa, fb, se, sh; };

// Material / mask format flags of an operator.
enum : int {
  OPF_EPS_COMPLEX = 1,   // eps interleaved complex, else real
  OPF_EPS_ANISO   = 2,   // eps has 3N values (per component), else N (isotropic)
  OPF_MU          = 4,   // inv_mu present
  OPF_MU_COMPLEX  = 8,
  OPF_MU_ANISO    = 16,
  OPF_MASK        = 32,  // uint8 per cell: bits 0-2 PEC x,y,z ; bits 3-5 PMC x,y,z
};

// Material kinds.  The apply has one instantiation per kind, in which the code and the
// registers of what the kind excludes drop out at compile time: SIMPLE = real isotropic
// eps only; ANISO = real per-component eps; MASK = real isotropic eps with a PEC/PMC
// mask; GENERAL = anything (complex or per-component eps with a mask, mu).
enum : int { MK_GENERAL = 0, MK_SIMPLE = 1, MK_ANISO = 2, MK_MASK = 3 };
__host__ __device__ constexpr int mat_kind(int flags) {
  return (flags & (OPF_MU | OPF_EPS_COMPLEX)) ? MK_GENERAL
         : (flags & OPF_MASK) ? ((flags & OPF_EPS_ANISO) ? MK_GENERAL : MK_MASK)
         : (flags & OPF_EPS_ANISO) ? MK_ANISO
         : MK_SIMPLE;
}

// Device view of one operator (one structure) in precision T.
template <typename T> struct OpView {
  int nx, ny, nz, flags;
  cplx<T> w2;                       // omega^2
  const AxisCoef<T>* coef;          // nx + ny + nz entries: x axis, then y, then z
  const T* eps;                     // real or interleaved complex, N or 3N entries
  const T* inv_mu;                  // same conventions, or null
  const uint8_t* mask;              // N bytes or null
};

// Solver state flags (per linear system).
enum : int {
  SF_PAUSE     = 1,    // kernels skip this system until the host clears it
  SF_BREAKDOWN = 2,    // p^T A p vanished (COCG breakdown)
  SF_DONE      = 4,    // converged / finished: kernels skip permanently
  SF_NAN       = 8,    // non-finite scalar seen
  SF_RESTART   = 16,   // next p-update uses beta = 0 (p = r)
  SF_SELECT    = 32,   // selected for a host-driven step (residual / accumulate)
  SF_INACTIVE  = 64,   // slot not in use
  SF_LIMIT     = 128,  // iteration limit reached: the next replacement finishes the system
  SF_MR        = 256,  // next step is a rescue step along a hashed direction (breakdown)
  SF_STALL     = 512,  // stopped: replacements no longer reduce the true residual
  SF_SAVE      = 1024, // the last replacement found the best true residual: save x
  SF_FUSED     = 2048, // fused iteration form (complex128 inner vectors; see vector.hip)
  SF_SWITCH_EXIT = 4096,  // on losing track, finish (SF_SWITCHED) instead of switching form in place
  SF_SWITCHED  = 8192, // finished to be continued in the classic form by the host
  SF_UPD_PENDING = 16384, // the update ran; its partials await the post-update finishing step
  SF_XPEND     = 32768, // fused form: x_lo still lacks the last alpha p, kept as xa (p - r)
  SF_XSKIP     = 65536, // fused form: the coming update pass leaves x_lo alone (it defers alpha p)
};

// Deferred solution update of the fused form: an update pass may leave x_lo alone when
// |beta|^2 is at least this (alpha p is then recovered one pass later from p' - r' = beta p)
constexpr double kDeferBeta2 = 1e-6;

// |r^T r| below this fraction of ||r||^2 counts as a COCG (Lanczos) breakdown
constexpr double kRhoBreakdown = 1e-10;

// Scalars of one linear system, device resident.  Written by the finishing kernel that
// follows each reducing kernel (deterministic fixed-order sums), read by later kernels.
struct SysState {
  cplx<double> rho;        // r^T r (unconjugated) of the current residual
  cplx<double> alpha, beta;
  cplx<double> mu;         // p^T A p
  cplx<double> rho_old;    // previous rho (classic form: beta after a residual replacement)
  double rnorm2;           // ||r||^2 of the (recursive) residual
  double rmax2;            // max recursive ||r||^2 since the last residual replacement
  double true_rnorm2;      // ||b - A x||^2 in FP64 at the last replacement
  double bnorm2;           // ||b||^2
  double stop2;            // recursive-residual target (absolute, squared)
  double delta2;           // reliable-update trigger: pause when rnorm2 < delta2 * rmax2
  double rtol, atol;       // convergence: ||b - A x|| <= max(rtol ||b||, atol)
  double best_true2;       // smallest true ||r||^2 seen at a replacement (x saved in VS_XBEST)
  double rnorm_rec2;       // recursive ||r||^2 just before the last replacement
  double mark2;            // true ||r||^2 at the last 2x improvement of the best residual
  int flags;
  int iters;               // inner iterations performed
  int replacements;        // residual replacements (reliable updates / refinements)
  int maxiter;
  int since;               // iterations since the last replacement
  int refresh_iters;       // a residual replacement is forced after this many iterations
                           // without one (keeps the best-iterate record current)
  int stall_iters;         // stall window (0 = off): the solve stops (SF_STALL) when the
                           // best true residual has not halved within this many iterations
  int stalls;              // consecutive replacements whose recursive residual had drifted
                           // >= 100x below the true one (precision exhausted)
  int stall_limit;         // stop (SF_STALL) after this many such replacements
  int mark_iter;           // iteration of the last 2x improvement
  int pad_;
  cplx<double> xa;         // fused form: the deferred alpha p equals xa (p - r) (SF_XPEND)
  cplx<double> xc1, xc2;   // fused form: the coming update pass adds xc1 p + xc2 r to x_lo
};

// Finishing-kernel kinds (k_fin<KIND>, vector.hip)
enum : int { FIN_APPLY = 0, FIN_UPDATE = 1, FIN_RESID = 2, FIN_BNORM = 3, FIN_APPLY_MERGED = 4 };

// One linear system of a batch.  Vector slots name the vectors a kernel reads/writes.
enum : int { VS_X = 0, VS_R, VS_P, VS_Q, VS_XLO, VS_B, VS_XBEST, VS_COUNT };

struct SysDesc {
  int op;                     // operator index
  int nblk_apply;             // blocks of the apply grid owned by this system
  int nblk_vec;               // blocks of the vector grid owned by this system
  int slab;                   // x-planes per apply block
  long long n3;               // 3 * nx * ny * nz
  void* vec[VS_COUNT];        // device vectors (precision depends on the slot and mode)
  SysState* st;
  double* partial;            // >= 4 * max(nblk_apply, nblk_vec) doubles
  double* partial_vec;        // >= 3 * nblk_vec doubles: the update's partials, kept apart
                              // so the next apply can run before they are summed
};

// Work assignment of one block: system index and block index within that system.
struct BlockMap { int sys, local; };

// ---- reductions ----------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum K doubles over the block; result valid in thread 0.  `scratch` >= K * nwaves.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* scratch) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthreads = blockDim.x * blockDim.y;
  const int lane = tid & 63, wave = tid >> 6, nwaves = (nthreads + 63) >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) scratch[q * 16 + wave] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) {
      double s = 0.0;
      for (int w = 0; w < nwaves; ++w) s += scratch[q * 16 + w];
      v[q] = s;
    }
  }
  __syncthreads();
}

// Store this block's K partial sums (plain stores by thread 0).  No fence: the partials
// are consumed by a separate finishing kernel, after the kernel boundary has made them
// visible on every XCD.  (A per-block device-scope release fence would write back the
// XCD's whole L2 on this part and costs more than the kernel itself.)
template <int K>
__device__ __forceinline__ void store_partials(const double (&v)[K], double* partial, int local, int nblk) {
  if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) ((MX_GLOBAL double*)partial)[q * nblk + local] = v[q];
  }
}

// Fixed-order sum of the K partial arrays of one system; result in thread 0.  Each
// thread first issues all of its loads (at most kFinLoads per array), then sums.  The
// loads are unconditional: a slot past the array reads entry 0 and is replaced by zero.
// (A load under `b < nblk` became a branch whose load the compiler waited for inside
// it: one memory round trip per array before the batch of the others.)
// 512 threads (8 waves): measured 0.5% faster per FP64 iteration at 128^3 than 1024 (the
// block reduction syncs fewer waves); 256 needs 16 loads per thread and array.
#ifndef MX_FIN_THREADS
#define MX_FIN_THREADS 512
#endif
constexpr int kFinThreads = MX_FIN_THREADS;
constexpr int kFinLoads = 4096 / kFinThreads;   // nblk <= 4096 partials per array
template <int K>
__device__ __forceinline__ void final_sum(double (&out)[K], const double* partial, int nblk, double* scratch) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthreads = blockDim.x * blockDim.y;
  double v[K][kFinLoads];
#pragma unroll
  for (int q = 0; q < K; ++q)
#pragma unroll
    for (int l = 0; l < kFinLoads; ++l) {
      const int b = tid + l * nthreads;
      const double x = gload1<double>(partial, b < nblk ? q * nblk + b : 0);
      v[q][l] = b < nblk ? x : 0.0;
    }
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double sacc = 0.0;
#pragma unroll
    for (int l = 0; l < kFinLoads; ++l) sacc += v[q][l];
    out[q] = sacc;
  }
  block_sum<K>(out, scratch);
}

// Fixed-order sums of two partial sets in one pass: every load of both is issued before
// any is summed, so the two reductions cost one memory round trip (unconditional loads,
// as in final_sum; both arrays hold at least one entry even when n is 0).
template <int K1, int K2>
__device__ __forceinline__ void final_sum2(double (&o1)[K1], const double* p1, int n1, double (&o2)[K2],
                                           const double* p2, int n2, double* scratch) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthreads = blockDim.x * blockDim.y;
  double v1[K1][kFinLoads], v2[K2][kFinLoads];
#pragma unroll
  for (int l = 0; l < kFinLoads; ++l) {
    const int b = tid + l * nthreads;
#pragma unroll
    for (int q = 0; q < K1; ++q) {
      const double x = gload1<double>(p1, b < n1 ? q * n1 + b : 0);
      v1[q][l] = b < n1 ? x : 0.0;
    }
#pragma unroll
    for (int q = 0; q < K2; ++q) {
      const double x = gload1<double>(p2, b < n2 ? q * n2 + b : 0);
      v2[q][l] = b < n2 ? x : 0.0;
    }
  }
  double all[K1 + K2];
#pragma unroll
  for (int q = 0; q < K1 + K2; ++q) {
    double sacc = 0.0;
#pragma unroll
    for (int l = 0; l < kFinLoads; ++l) sacc += q < K1 ? v1[q][l] : v2[q - K1][l];
    all[q] = sacc;
  }
  block_sum<K1 + K2>(all, scratch);
#pragma unroll
  for (int q = 0; q < K1; ++q) o1[q] = all[q];
#pragma unroll
  for (int q = 0; q < K2; ++q) o2[q] = all[K1 + q];
}

__device__ __forceinline__ bool finite(double v) { return __builtin_isfinite(v); }

// Deterministic pseudo-random complex direction (entries in [-1/2, 1/2)^2) of vector
// index e: a 64-bit mix (splitmix64 finaliser) of the index.  Used to restart COCG
// with a generic residual after a breakdown.
__device__ __forceinline__ cplx<double> hashed_direction(long long e) {
  unsigned long long z = (unsigned long long)e * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const double u = (double)(z >> 40) * (1.0 / 16777216.0) - 0.5;
  const double v = (double)((z >> 8) & 0xFFFFFFull) * (1.0 / 16777216.0) - 0.5;
  return cplx<double>{u, v};
}

}  // namespace mx
