// Matrix-free apply of the symmetrised FDFD wave operator (see common.hpp for the
// operator algebra), with fused epilogues for COCG and FP64 residual replacement.
//
// Reference operator: meanas/fdfd/operators.py:40-96 (e_full) scaled by
// meanas/fdfd/operators.py:99-123 (e_full_preconditioners) as in
// meanas/fdfd/solvers.py:118-125.  Curl conventions: meanas/fdmath/operators.py:99-171,
// 222-251 (forward differences on E, backward on H, periodic wrap).
//
// Kernel shape: a block owns a TY x 64 tile of (y, z) columns and marches along x over
// `slab` planes.  Per plane it reads the 3 components of the next x-plane of the unknown
// (tile + one-cell halo in y and z) once from HBM into registers; the H phase of plane i
// takes its y/z neighbours from plane i in LDS and the x-difference from plane i+1 in
// registers, then plane i+1 replaces plane i in the single LDS plane.  H on the tile plus
// a one-cell halo on the minus side goes to LDS, and the output is formed from H(i), its
// y-/z- neighbours in LDS and H(i-1) kept in registers.  z is the contiguous axis: a wave
// reads 64 consecutive complex values per row.
#include <cstdlib>
#include <type_traits>

#include "common.hpp"

namespace mx {

enum : int { EPI_PLAIN = 0, EPI_COCG = 1, EPI_RESID = 2 };

__device__ __forceinline__ int wrapi(int v, int n) {
  int r = v % n;
  return r < 0 ? r + n : r;
}

template <typename T>
__device__ __forceinline__ cplx<T> load_c(const T* base, int idx) {
  return gload<T>(base, idx);
}

// material value (eps or 1/mu) of component c at cell `cell`.  Ordinary loads: the
// materials are small enough to stay in the Infinity Cache from one apply to the next
// (streaming loads measured 24% slower in the solve).
template <typename T>
__device__ __forceinline__ cplx<T> load_mat(const T* base, bool is_complex, bool aniso, int c,
                                            int cell, int ncell) {
  const int idx = aniso ? c * ncell + cell : cell;
  if (is_complex) return gload<T>(base, idx);
  return cplx<T>{gload1<T>(base, idx), T(0)};
}

// fa x1 - fb x0 with complex coefficients.  (Real-coefficient variants for axes outside
// the PML, selected per plane and per wave, made the kernel 1.46x larger and measured 3%
// slower in the 128^3 solve, so there is one form.)
template <typename T>
__device__ __forceinline__ cplx<T> dif(const cplx<T>& fa, const cplx<T>& x1, const cplx<T>& fb, const cplx<T>& x0) {
  return fa * x1 - fb * x0;
}

// forward-difference coefficients (fa, fb) of entry n of a per-axis table; coef_u for
// an n that is uniform over the wave (scalar loads, SGPRs)
template <typename T> struct FaFb { cplx<T> fa, fb; };
template <typename T>
__device__ __forceinline__ FaFb<T> coef_at(const AxisCoef<T>* tab, int n) {
  return FaFb<T>{gload<T>(tab, 4 * n), gload<T>(tab, 4 * n + 1)};
}
template <typename T>
__device__ __forceinline__ FaFb<T> coef_u(const AxisCoef<T>* tab, int n) {
  return FaFb<T>{cload<T>(tab, 4 * n), cload<T>(tab, 4 * n + 1)};
}

// Occupancy target (the second launch-bounds argument is the minimum number of waves
// per SIMD): complex64 fits 4 waves (<= 128 VGPRs); complex128 needs 214 VGPRs, and
// holding it to 3 waves (168) spills 117 of them to scratch, so it runs 2.
// MX_APPLY_WAVES64 overrides the complex128 target for tuning builds.
//
// Halo work.  Besides its own row, a tile forms H on the halo row j0-1 (Hx, Hz) and the
// halo column k0-1 (Hx, Hy).  Hz of the row and Hy of the column need plane i+1 from
// the registers of the waves that loaded those sites (0 and TY-1); the two Hx parts read
// only LDS.  MX_APPLY_SPLIT_HALO = 1 gives the Hx parts to waves 1 and TY-2, so no wave
// forms more than one extra component and the others wait less at the barrier.
// MX_APPLY_HALO_PRIO > 0 instead raises the priority (s_setprio) of waves 0 and TY-1
// during the H phase.  Measured in the 128^3 solve (c128 apply): neither 56.1 us,
// priority 53.9-55.7 us, split 54.7 us; split plus priority no better than the split.
#ifndef MX_APPLY_SPLIT_HALO
#define MX_APPLY_SPLIT_HALO 1
#endif
#ifndef MX_APPLY_HALO_PRIO
#define MX_APPLY_HALO_PRIO 0
#endif
#ifndef MX_APPLY_WAVES64
#define MX_APPLY_WAVES64 2
#endif
// The material-kind kernels (MK other than MK_GENERAL) need 148-166 VGPRs in
// complex128, so they can hold 3 waves per SIMD; at the 128^3 tiling (512 blocks of 4
// waves) only 2 are used either way, and deeper occupancy measured no faster (slab sweep
// in the solve), so the bound stays 2.
#ifndef MX_APPLY_WAVES64_SM
#define MX_APPLY_WAVES64_SM 2
#endif
template <typename T, int MK>
constexpr int kWavesPerSimd = sizeof(T) == 4 ? 4 : (MK != MK_GENERAL ? MX_APPLY_WAVES64_SM : MX_APPLY_WAVES64);

template <typename T, int TY, int EPI, typename RT, int PD, int MK>
__global__ __launch_bounds__(64 * TY, (kWavesPerSimd<T, MK>)) void k_apply(const OpView<T>* __restrict__ ops,
                                                                 const SysDesc* __restrict__ systems,
                                                                 const BlockMap* __restrict__ bmap,
                                                                 int in_slot, int out_slot,
                                                                 int skip_flags, int need_flags,
                                                                 const void* in_ptr, void* out_ptr) {
  constexpr int TZ = 64;
  constexpr int LY = TY + 2, LZ = TZ + 2;
  static_assert(TY >= 2 && 2 * TY + 2 <= 64, "tile rows: the z-halo loads take 2 TY + 2 lanes of one wave");
  __shared__ cplx<T> sE[3][LY][LZ];   // plane i of the unknown, tile + one-cell halo
  __shared__ uint8_t sM[LY][LZ];      // its PEC/PMC mask bits
  __shared__ cplx<T> sH[3][TY + 1][TZ + 1];
  __shared__ double s_red[4 * 16];
  __shared__ FaFb<T> sCy[TY];         // y coefficients of the tile rows (the column halo's)

  const BlockMap bm = bmap[blockIdx.x];
  const SysDesc& S = systems[bm.sys];
  SysState* st = S.st;
  bool mr = false;     // rescue step: accumulate ||q||^2 instead of p^T q, q^T q
  bool keep_p = false;  // residual replacement that keeps the search direction
  if (EPI != EPI_PLAIN) {
    const int f = gload1<int>(&st->flags, 0);
    if ((f & skip_flags) || (need_flags && !(f & need_flags))) return;
    mr = (f & SF_MR) != 0;
    // fused iteration form only: the classic form rebuilds p in k_pupdate
    keep_p = EPI == EPI_RESID && (f & SF_FUSED) && !(f & (SF_MR | SF_RESTART)) && S.vec[VS_P] != nullptr;
  }
  const OpView<T> op = ops[S.op];
  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const long long plane = (long long)ny * nz;
  const long long ncell = plane * nx;
  const int ntj = (ny + TY - 1) / TY, ntk = (nz + TZ - 1) / TZ;
  const int tiles = ntj * ntk;
  const int slab_id = bm.local / tiles, tile = bm.local - slab_id * tiles;
  const int j0 = (tile / ntk) * TY, k0 = (tile % ntk) * TZ;
  const int ib = slab_id * S.slab;
  const int ie = min(ib + S.slab, nx);
  // threadIdx.y is uniform over a wave (blockDim.x = 64); readfirstlane tells the compiler
  const int tz = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int j = j0 + ty, k = k0 + tz;
  const bool valid = (j < ny) && (k < nz);
  const int jw = wrapi(j, ny), kw = wrapi(k, nz);

  // The material kind MK fixes the format flags at compile time (the launch picks the
  // kind of its operators), so the code and registers of what it excludes drop out;
  // MK_GENERAL reads them from the operator.  SM: real isotropic eps, one value for the
  // three components.
  constexpr bool SM = MK == MK_SIMPLE || MK == MK_MASK;
  const bool has_mask = MK == MK_MASK || (MK == MK_GENERAL && (op.flags & OPF_MASK) != 0);
  const bool has_mu = MK == MK_GENERAL && (op.flags & OPF_MU) != 0;
  const bool eps_cplx = MK == MK_GENERAL && (op.flags & OPF_EPS_COMPLEX) != 0;
  const bool eps_aniso = MK == MK_ANISO || (MK == MK_GENERAL && (op.flags & OPF_EPS_ANISO) != 0);
  const AxisCoef<T>* cx = op.coef;
  const AxisCoef<T>* cy = op.coef + nx;
  const AxisCoef<T>* cz = op.coef + nx + ny;

  const T* __restrict__ xin = reinterpret_cast<const T*>(in_ptr ? in_ptr : S.vec[in_slot]);

  // ---- load sites of this thread: own column, optional y-halo row, optional z-halo
  // cell.  Indices are 32-bit: the host guarantees 3N < 2^31.
  // Only plane i sits in LDS; the x-difference takes plane i+1 from the registers of the
  // thread that loaded it, so every value of plane i+1 that a thread needs is one of its
  // own loads: its own cell, its y-halo row (ty == 0 computes the halo H of row j0-1) and
  // its z-halo cell (lanes 0..TY-1 of wave TY-1 compute the halo H of column k0-1).
  const int own_off = jw * nz + kw;
  const int own_lds = (ty + 1) * LZ + (tz + 1);
  const bool has_row = (ty == 0) || (ty == TY - 1);
  const int row_off = wrapi(ty == 0 ? j0 - 1 : j0 + TY, ny) * nz + kw;
  const int row_lds = (ty == 0 ? 0 : TY + 1) * LZ + (tz + 1);
  // z-halo cells, wave TY-1: lanes 0..TY take column k0-1 at LDS rows 1..TY+1, lanes
  // TY+1..2TY+1 column k0+TZ at LDS rows 0..TY (the halo cells the stencil reads)
  const bool has_col = (ty == TY - 1) && (tz < 2 * TY + 2);
  const int col_r = tz < TY + 1 ? tz + 1 : tz - (TY + 1);
  const int col_c = tz < TY + 1 ? 0 : TZ + 1;
  const int col_off = has_col ? wrapi(j0 - 1 + col_r, ny) * nz + wrapi(col_c ? k0 + TZ : k0 - 1, nz) : 0;
  const int col_lds = col_r * LZ + col_c;
  const bool col_halo_h = (ty == TY - 1) && (tz < TY);   // halo Hy at column k0-1, LDS row tz+1
  // waves forming the LDS-only halo components (Hx of the row and of the column)
#if MX_APPLY_SPLIT_HALO
  constexpr int kRowHx = 1, kColHx = TY - 2;
#else
  constexpr int kRowHx = 0, kColHx = TY - 1;
#endif
  const int nc = (int)ncell;

  // One plane's worth of this thread's loads (own column, y-halo row, z-halo cell).
  struct Regs {
    cplx<T> o[3], r[3], c[3];
    uint8_t mo, mr, mc;
  };
  auto fetch = [&](Regs& R, int ip) {
    const int pb = ip * (int)plane;
#pragma unroll
    for (int c = 0; c < 3; ++c) R.o[c] = load_c<T>(xin, c * nc + pb + own_off);
    R.mo = has_mask ? gload1<uint8_t>(op.mask, pb + own_off) : 0;
    if (has_row) {
#pragma unroll
      for (int c = 0; c < 3; ++c) R.r[c] = load_c<T>(xin, c * nc + pb + row_off);
      R.mr = has_mask ? gload1<uint8_t>(op.mask, pb + row_off) : 0;
    }
    if (has_col) {
#pragma unroll
      for (int c = 0; c < 3; ++c) R.c[c] = load_c<T>(xin, c * nc + pb + col_off);
      R.mc = has_mask ? gload1<uint8_t>(op.mask, pb + col_off) : 0;
    }
  };
  // a loaded value with PEC applied (the unknown is zero on PEC components)
  auto pec = [](const cplx<T> (&v)[3], uint8_t m, int c) { return (m >> c) & 1 ? zero<T>() : v[c]; };
  auto put = [&](int lds, const cplx<T> (&v)[3], uint8_t m) {
    cplx<T>* base = &sE[0][0][0];
#pragma unroll
    for (int c = 0; c < 3; ++c) base[c * LY * LZ + lds] = pec(v, m, c);
    if (has_mask) (&sM[0][0])[lds] = m;
  };
  auto commit = [&](const Regs& R) {
    put(own_lds, R.o, R.mo);
    if (has_row) put(row_lds, R.r, R.mr);
    if (has_col) put(col_lds, R.c, R.mc);
  };
  // Register prefetch, PD register sets used in turn: the loads of plane i+1 are issued
  // in the middle of step i-PD and consumed by the H phase of step i.
  static_assert(PD == 1 || PD == 2, "prefetch depth 1 or 2");
  Regs RA, RB;
  Regs& R = RA;

  // per-thread y/z coefficients (wrapped indices)
  const FaFb<T> cyj = coef_u(cy, jw), czk = coef_at(cz, kw);
  const cplx<T> fa_ym = coef_u(cy, wrapi(j - 1, ny)).fa, fa_zm = coef_at(cz, wrapi(k - 1, nz)).fa;
  const FaFb<T> cyh = coef_u(cy, wrapi(j0 - 1, ny));   // halo row j0-1
  const FaFb<T> czh = coef_u(cz, wrapi(k0 - 1, nz));   // halo column k0-1
  // y coefficients of the rows whose column-halo H the lanes tz < TY form: each wave's
  // own row coefficients, shared through LDS (the prologue's barrier orders it), so the
  // plane loop reads them from LDS and no thread keeps them in registers
  if (tz == 0) sCy[ty] = cyj;

  // H multiplier at (cell, component): pmc keep * 1/mu
  auto hscale = [&](cplx<T> h, int c, uint8_t m, int cell) -> cplx<T> {
    if ((m >> (3 + c)) & 1) return zero<T>();
    if (has_mu) {
      const cplx<T> im = load_mat<T>(op.inv_mu, op.flags & OPF_MU_COMPLEX, op.flags & OPF_MU_ANISO, c, cell, nc);
      h = im * h;
    }
    return h;
  };
  auto e_ = [&](int c, int rr, int qq) { return sE[c][rr][qq]; };

  // eps of the next plane.  Simple materials hold one value for the three components in
  // one register: three copies of one load, carried round the plane loop, made the
  // compiler wait for that load right where it was issued, before the step's second half
  // (vmcnt counts in order, so the plane prefetch issued with it waited too).  Measured
  // at 128^3: COCG apply 43.0 -> 41.5 us, plain complex64 apply 35.5 -> 32.2 us.
  cplx<T> epsn[3];
  auto load_eps = [&](int cell) {
    if constexpr (SM) {
      epsn[0] = load_mat<T>(op.eps, false, false, 0, cell, nc);
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c)
        epsn[c] = load_mat<T>(op.eps, eps_cplx, eps_aniso, c, cell, nc);
    }
  };

  // Output stores (plain and COCG applies) go through one bounded resource per component
  // (kDropOffset for lanes outside the grid): the same three stores on every path, so the
  // wait for the next plane's prefetch in the H phase never waits for them.
  T* __restrict__ yout = reinterpret_cast<T*>(out_ptr ? out_ptr : S.vec[out_slot]);
  const unsigned comp_bytes = (unsigned)nc * (unsigned)sizeof(cplx<T>);
  __amdgpu_buffer_rsrc_t yrs[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) yrs[c] = bounded_rsrc(reinterpret_cast<char*>(yout) + (size_t)c * comp_bytes, comp_bytes);

  // ---- prologue: H(ib-1) at the own column (y, z components); plane ib into LDS
  cplx<T> hpy, hpz;
  {
    const int im1 = wrapi(ib - 1, nx);
    fetch(R, im1);
    commit(R);
    fetch(R, ib);
    __syncthreads();
    const FaFb<T> cxm = coef_u(cx, im1);
    const int r = ty + 1, q = tz + 1;
    const cplx<T> n1 = pec(R.o, R.mo, 1), n2 = pec(R.o, R.mo, 2);
    cplx<T> hy = (czk.fa * e_(0, r, q + 1) - czk.fb * e_(0, r, q)) - (cxm.fa * n2 - cxm.fb * e_(2, r, q));
    cplx<T> hz = (cxm.fa * n1 - cxm.fb * e_(1, r, q)) - (cyj.fa * e_(0, r + 1, q) - cyj.fb * e_(0, r, q));
    const uint8_t m = has_mask ? sM[r][q] : 0;
    const int cell = im1 * (int)plane + own_off;
    hpy = hscale(hy, 1, m, cell);
    hpz = hscale(hz, 2, m, cell);
    __syncthreads();
    commit(R);                      // plane ib
    load_eps(ib * (int)plane + own_off);
    fetch(RA, wrapi(ib + 1, nx));   // plane ib+1 (H(ie-1) needs plane ie, and ib < ie)
    if constexpr (PD == 2)
      if (ib + 2 <= ie) fetch(RB, wrapi(ib + 2, nx));
    // three dropped stores, so that the loop is entered as every later step is: with the
    // prefetch followed by three stores
    if constexpr (EPI != EPI_RESID) {
#pragma unroll
      for (int c = 0; c < 3; ++c) bstore<T>(yrs[c], kDropOffset, zero<T>());
    }
  }

  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const cplx<T> w2 = op.w2;
  const bool diag_real = !eps_cplx && w2.im == T(0);
  const double* __restrict__ bvec = reinterpret_cast<const double*>(S.vec[VS_B]);
  RT* __restrict__ rout = reinterpret_cast<RT*>(S.vec[out_slot]);
  RT* __restrict__ pdir = reinterpret_cast<RT*>(S.vec[VS_P]);

  // step i: Rn holds plane i+1; epsn holds eps of plane i, loaded during step i-1 ahead
  // of that step's plane prefetch (so waiting for it never waits for the prefetch)
  auto step = [&](const int i, Regs& Rn) {
    const int cell = i * (int)plane + own_off;
    cplx<T> epsv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) epsv[c] = epsn[SM ? 0 : c];
    const FaFb<T> cxi = coef_u(cx, i);
    const cplx<T> fa_xm = coef_u(cx, wrapi(i - 1, nx)).fa;
    __syncthreads();   // plane i is in LDS; the previous step's reads of sH are done

    const int r = ty + 1, q = tz + 1;
    cplx<T> ec[3] = {e_(0, r, q), e_(1, r, q), e_(2, r, q)};
    const uint8_t mown = has_mask ? sM[r][q] : 0;
#if MX_APPLY_HALO_PRIO
    if (ty == 0 || ty == TY - 1) __builtin_amdgcn_s_setprio(MX_APPLY_HALO_PRIO);   // the waves with halo H work
#endif
    {   // H phase
      const cplx<T> hx = dif(cyj.fa, e_(2, r + 1, q), cyj.fb, ec[2]) - dif(czk.fa, e_(1, r, q + 1), czk.fb, ec[1]);
      const cplx<T> hy = dif(czk.fa, e_(0, r, q + 1), czk.fb, ec[0]) - dif(cxi.fa, pec(Rn.o, Rn.mo, 2), cxi.fb, ec[2]);
      const cplx<T> hz = dif(cxi.fa, pec(Rn.o, Rn.mo, 1), cxi.fb, ec[1]) - dif(cyj.fa, e_(0, r + 1, q), cyj.fb, ec[0]);
      sH[0][r][q] = hscale(hx, 0, mown, cell);
      sH[1][r][q] = hscale(hy, 1, mown, cell);
      sH[2][r][q] = hscale(hz, 2, mown, cell);
      // Halo row j0-1 (LDS row 0) needs Hx and Hz, halo column k0-1 (LDS column 0, rows
      // 1..TY) Hx and Hy.  Hz of the row and Hy of the column use plane i+1, which is in
      // the registers of the waves that loaded those sites (0 and TY-1); the Hx parts
      // read only LDS and go to waves 1 and TY-2, so no wave forms more than one
      // extra component.
      if (ty == kRowHx) {   // halo row: Hx
        const uint8_t mh = has_mask ? sM[0][q] : 0;
        const int hcell = i * (int)plane + wrapi(j0 - 1, ny) * nz + kw;
        const cplx<T> gx = dif(cyh.fa, e_(2, 1, q), cyh.fb, e_(2, 0, q)) - dif(czk.fa, e_(1, 0, q + 1), czk.fb, e_(1, 0, q));
        sH[0][0][q] = hscale(gx, 0, mh, hcell);
      }
      if (ty == 0) {   // halo row: Hz
        const uint8_t mh = has_mask ? sM[0][q] : 0;
        const int hcell = i * (int)plane + wrapi(j0 - 1, ny) * nz + kw;
        const cplx<T> gz = dif(cxi.fa, pec(Rn.r, Rn.mr, 1), cxi.fb, e_(1, 0, q)) - dif(cyh.fa, e_(0, 1, q), cyh.fb, e_(0, 0, q));
        sH[2][0][q] = hscale(gz, 2, mh, hcell);
      }
      if (ty == kColHx && tz < TY) {   // halo column: Hx
        const int rr = tz + 1;
        const uint8_t mh = has_mask ? sM[rr][0] : 0;
        const int hcell = i * (int)plane + wrapi(j0 + tz, ny) * nz + wrapi(k0 - 1, nz);
        const FaFb<T> cyr = sCy[tz];
        const cplx<T> gx = dif(cyr.fa, e_(2, rr + 1, 0), cyr.fb, e_(2, rr, 0)) - dif(czh.fa, e_(1, rr, 1), czh.fb, e_(1, rr, 0));
        sH[0][rr][0] = hscale(gx, 0, mh, hcell);
      }
      if (col_halo_h) {   // halo column: Hy
        const int rr = tz + 1;
        const uint8_t mh = has_mask ? sM[rr][0] : 0;
        const int hcell = i * (int)plane + wrapi(j0 + tz, ny) * nz + wrapi(k0 - 1, nz);
        const cplx<T> gy = dif(czh.fa, e_(0, rr, 1), czh.fb, e_(0, rr, 0)) - dif(cxi.fa, pec(Rn.c, Rn.mc, 2), cxi.fb, e_(2, rr, 0));
        sH[1][rr][0] = hscale(gy, 1, mh, hcell);
      }
    }
#if MX_APPLY_HALO_PRIO
    __builtin_amdgcn_s_setprio(0);
#endif
    __syncthreads();   // H(i) is in LDS; every read of plane i is done
    if (i + 1 < ie) {
      commit(Rn);                                         // plane i+1 replaces plane i
      load_eps((i + 1) * (int)plane + own_off);           // eps of plane i+1, ahead of the prefetch
      if (i + 1 + PD <= ie) fetch(Rn, wrapi(i + 1 + PD, nx));   // plane i+1+PD into the freed set
    }

    const cplx<T> hx0 = sH[0][r][q], hy0 = sH[1][r][q], hz0 = sH[2][r][q];
    const cplx<T> yx = dif(cyj.fb, hz0, fa_ym, sH[2][r - 1][q]) - dif(czk.fb, hy0, fa_zm, sH[1][r][q - 1]);
    const cplx<T> yy = dif(czk.fb, hx0, fa_zm, sH[0][r][q - 1]) - dif(cxi.fb, hz0, fa_xm, hpz);
    const cplx<T> yz = dif(cxi.fb, hy0, fa_xm, hpy) - dif(cyj.fb, hx0, fa_ym, sH[0][r - 1][q]);
    hpy = hy0;
    hpz = hz0;
    cplx<T> out[3] = {yx, yy, yz};
    if (diag_real) {   // omega^2 eps real: one real scale of E
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T we = w2.re * epsv[c].re;
        out[c] = cplx<T>{out[c].re - we * ec[c].re, out[c].im - we * ec[c].im};
        if ((mown >> c) & 1) out[c] = zero<T>();
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        out[c] = out[c] - w2 * (epsv[c] * ec[c]);
        if ((mown >> c) & 1) out[c] = zero<T>();
      }
    }
    if constexpr (EPI != EPI_RESID) {
      const unsigned off = valid ? (unsigned)cell * (unsigned)sizeof(cplx<T>) : kDropOffset;
#pragma unroll
      for (int c = 0; c < 3; ++c) bstore<T>(yrs[c], off, out[c]);
    }
    if (valid) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int idx = c * nc + cell;
        if constexpr (EPI == EPI_RESID) {
          const cplx<double> bb = gload<double>(bvec, idx);
          const cplx<double> rr = bb - cast<double>(out[c]);
          const cplx<RT> rs = cast<RT>(rr);
          if (keep_p) {
            // p was built from the recursive residual: p += r_true - r_recursive keeps
            // its direction component and swaps in the true residual
            const cplx<RT> rold = gload<RT>(rout, idx);
            gstore<RT>(pdir, idx, gload<RT>(pdir, idx) + (rs - rold));
          }
          gstore<RT>(rout, idx, rs);
          const cplx<double> rd = cast<double>(rs);
          acc[0] += norm2(rr);
          acc[1] += rd.re * rd.re - rd.im * rd.im;
          acc[2] += 2.0 * rd.re * rd.im;
        } else if constexpr (EPI == EPI_COCG) {
          if (!mr) {
            const cplx<T> pq = ec[c] * out[c];
            const cplx<T> qq = out[c] * out[c];
            acc[0] += (double)pq.re;
            acc[1] += (double)pq.im;
            acc[2] += (double)qq.re;
            acc[3] += (double)qq.im;
          } else {
            acc[2] += norm2(cast<double>(out[c]));
          }
        }
      }
    }
  };
  if constexpr (PD == 1) {
#pragma unroll 1
    for (int i = ib; i < ie; ++i) step(i, RA);
  } else {
#pragma unroll 1
    for (int i = ib; i < ie; i += 2) {
      step(i, RA);
      if (i + 1 < ie) step(i + 1, RB);
    }
  }

  if constexpr (EPI == EPI_COCG) {
    block_sum<4>(acc, s_red);
    store_partials<4>(acc, S.partial, bm.local, S.nblk_apply);
  } else if constexpr (EPI == EPI_RESID) {
    double a3[3] = {acc[0], acc[1], acc[2]};
    block_sum<3>(a3, s_red);
    store_partials<3>(a3, S.partial, bm.local, S.nblk_apply);
  }
}

// ---- host-side launch helpers ------------------------------------------------------
// Prefetch depth (register sets of planes in flight): MEANAS_AMD_PREFETCH=1|2 overrides
// the default for tuning runs.
static int prefetch_depth(size_t elem) {
  static const int forced = [] {
    const char* v = getenv("MEANAS_AMD_PREFETCH");
    return v ? atoi(v) : 0;
  }();
  if (forced == 1 || forced == 2) return forced;
  (void)elem;
  return 1;   // depth 2 (complex128 only: complex64 would spill) measured no faster in a solve
}

template <typename T, int EPI, typename RT>
static hipError_t launch_apply_ty(int ty, int kind, const OpView<T>* ops, const SysDesc* sys, const BlockMap* bmap,
                                  int nblocks, int in_slot, int out_slot, int skip, int need, hipStream_t s,
                                  const void* in_ptr = nullptr, void* out_ptr = nullptr) {
  const int pd = prefetch_depth(sizeof(T));
  // material-kind kernels exist for prefetch depth 1 only (depth 2 is a tuning override)
  const int mk = pd == 1 ? kind : MK_GENERAL;
#define MX_APPLY_CASE(TY_, PD_, MK_)                                                                        \
  case (TY_ * 10 + PD_) * 4 + MK_:                                                                         \
    hipLaunchKernelGGL((k_apply<T, TY_, EPI, RT, PD_, MK_>), dim3(nblocks), dim3(64, TY_), 0, s, ops, sys,   \
                       bmap, in_slot, out_slot, skip, need, in_ptr, out_ptr);                                 \
    break;
#define MX_APPLY_KINDS(TY_)                                                                                 \
  MX_APPLY_CASE(TY_, 1, MK_GENERAL)                                                                         \
  MX_APPLY_CASE(TY_, 1, MK_SIMPLE)                                                                          \
  MX_APPLY_CASE(TY_, 1, MK_ANISO)                                                                           \
  MX_APPLY_CASE(TY_, 1, MK_MASK)                                                                            \
  MX_APPLY_CASE(TY_, 2, MK_GENERAL)
  switch ((ty * 10 + pd) * 4 + mk) {
    MX_APPLY_KINDS(2)
    MX_APPLY_KINDS(4)
    MX_APPLY_KINDS(8)
    default: return hipErrorInvalidValue;
  }
#undef MX_APPLY_KINDS
#undef MX_APPLY_CASE
  return hipGetLastError();
}

hipError_t launch_apply_plain_f32(int ty, int kind, const OpView<float>* ops, const SysDesc* sys, const BlockMap* bmap, int nb,
                                  const void* x, void* y, hipStream_t s) {
  return launch_apply_ty<float, EPI_PLAIN, float>(ty, kind, ops, sys, bmap, nb, 0, 0, 0, 0, s, x, y);
}
hipError_t launch_apply_plain_f64(int ty, int kind, const OpView<double>* ops, const SysDesc* sys, const BlockMap* bmap, int nb,
                                  const void* x, void* y, hipStream_t s) {
  return launch_apply_ty<double, EPI_PLAIN, double>(ty, kind, ops, sys, bmap, nb, 0, 0, 0, 0, s, x, y);
}
hipError_t launch_apply_cocg_f32(int ty, int kind, const OpView<float>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, hipStream_t s) {
  return launch_apply_ty<float, EPI_COCG, float>(ty, kind, ops, sys, bmap, nb, VS_P, VS_Q,
                                                 SF_PAUSE | SF_DONE | SF_INACTIVE, 0, s);
}
hipError_t launch_apply_cocg_f64(int ty, int kind, const OpView<double>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, hipStream_t s) {
  return launch_apply_ty<double, EPI_COCG, double>(ty, kind, ops, sys, bmap, nb, VS_P, VS_Q,
                                                   SF_PAUSE | SF_DONE | SF_INACTIVE, 0, s);
}
// residual r = b - A x in FP64; r stored as float (mixed mode) or double (FP64 mode)
hipError_t launch_apply_resid_r32(int ty, int kind, const OpView<double>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, hipStream_t s) {
  return launch_apply_ty<double, EPI_RESID, float>(ty, kind, ops, sys, bmap, nb, VS_X, VS_R, SF_DONE | SF_INACTIVE, SF_PAUSE, s);
}
hipError_t launch_apply_resid_r64(int ty, int kind, const OpView<double>* ops, const SysDesc* sys, const BlockMap* bmap, int nb, hipStream_t s) {
  return launch_apply_ty<double, EPI_RESID, double>(ty, kind, ops, sys, bmap, nb, VS_X, VS_R, SF_DONE | SF_INACTIVE, SF_PAUSE, s);
}

}  // namespace mx
