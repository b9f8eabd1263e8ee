// Host runtime of the FDFD solve library: C ABI (include/meanas_fdfd.h), work-table
// construction, hipGraph-captured COCG iteration chunks and the pipelined host loop.
//
// Replaces meanas/fdfd/solvers.py:62 (scipy qmr on the assembled CSR matrix) with a
// batched, matrix-free, complex-symmetric Krylov solve on the GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <limits>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "../../include/meanas_fdfd.h"

namespace mx {
// apply.hip
hipError_t launch_apply_plain_f32(int ty, int kind, const OpView<float>*, const SysDesc*, const BlockMap*, int, const void*, void*, hipStream_t);
hipError_t launch_apply_plain_f64(int ty, int kind, const OpView<double>*, const SysDesc*, const BlockMap*, int, const void*, void*, hipStream_t);
hipError_t launch_apply_cocg_f32(int ty, int kind, const OpView<float>*, const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_apply_cocg_f64(int ty, int kind, const OpView<double>*, const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_apply_resid_r32(int ty, int kind, const OpView<double>*, const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_apply_resid_r64(int ty, int kind, const OpView<double>*, const SysDesc*, const BlockMap*, int, hipStream_t);
// vector.hip
hipError_t launch_pupdate_f32(const OpView<float>*, const SysDesc*, const BlockMap*, int, int, hipStream_t);
hipError_t launch_pupdate_f64(const OpView<double>*, const SysDesc*, const BlockMap*, int, int, hipStream_t);
hipError_t launch_update_f32(const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_update_f64(const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_accum_f32(const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_accum_f64(const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_bnorm(const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_save_best(const SysDesc*, const BlockMap*, int, hipStream_t);
hipError_t launch_zero_slot(int prec64, const SysDesc*, const BlockMap*, int, int, hipStream_t);
hipError_t launch_fin(int kind, const SysDesc*, const int* syslist, int nsys, hipStream_t);
hipError_t launch_diag(const AxisCoef<double>*, int, int, int, const void*, void*, double, double, int, hipStream_t);
}  // namespace mx

using namespace mx;
static_assert(SF_SWITCHED == MX_FLAG_SWITCHED, "the switched flag is part of the C ABI");

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

#define MX_TRY(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// A copy the host waits for, through a non-blocking stream of the calling thread and
// device instead of the legacy stream that hipMemcpy uses.  While another thread
// captures a graph, HIP refuses a legacy-stream operation ("would make the legacy stream
// depend on a capturing blocking stream"), and the refusal also invalidates that
// thread's capture, so no entry point issues one.  The streams live until the process
// ends.  Callers order the copy after device work that produces its source themselves
// (a stream synchronisation); every destination here is read by work queued afterwards.
hipError_t copy_now(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  thread_local std::vector<std::pair<int, hipStream_t>> streams;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipStream_t s = nullptr;
  for (const auto& ds : streams)
    if (ds.first == dev) s = ds.second;
  if (!s) {
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    streams.push_back({dev, s});
  }
  e = hipMemcpyAsync(dst, src, bytes, kind, s);
  return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

constexpr int kVecBlock = 256;
constexpr int kVecElemsPerThread = 8;
#ifndef MX_VEC_MAX_BLOCKS
#define MX_VEC_MAX_BLOCKS 2048
#endif
constexpr int kVecMaxBlocks = MX_VEC_MAX_BLOCKS;
constexpr size_t kAlign = 256;
constexpr int kMaxPartials = kFinThreads * kFinLoads;   // per-system partials one finishing block sums

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

long long ncell_of(const mx_operator& op) { return (long long)op.nx * op.ny * op.nz; }

bool valid_op(const mx_operator& op, int mode, std::string* why) {
  if (op.nx < 1 || op.ny < 1 || op.nz < 1) { *why = "grid dimensions must be >= 1"; return false; }
  // every complex128 vector must stay below 4 GiB: kernels address it with 32-bit byte offsets
  if (ncell_of(op) * 3 * 16 >= (1LL << 32)) { *why = "grid too large (a complex128 field must be < 4 GiB, i.e. < 89M cells)"; return false; }
  const bool need32 = (mode == MX_MIXED);
  if (!op.coef64 || !op.eps64) { *why = "coef64/eps64 are required"; return false; }
  if (need32 && (!op.coef32 || !op.eps32)) { *why = "coef32/eps32 are required for the mixed mode"; return false; }
  if ((op.flags & MX_MU) && (!op.inv_mu64 || (need32 && !op.inv_mu32))) { *why = "MX_MU set but inv_mu is NULL"; return false; }
  if ((op.flags & MX_MASK) && !op.mask) { *why = "MX_MASK set but mask is NULL"; return false; }
  return true;
}

template <typename T>
OpView<T> view_of(const mx_operator& op) {
  OpView<T> v;
  v.nx = op.nx; v.ny = op.ny; v.nz = op.nz; v.flags = op.flags;
  v.w2 = cplx<T>{(T)op.omega2_re, (T)op.omega2_im};
  if constexpr (sizeof(T) == 4) {
    v.coef = reinterpret_cast<const AxisCoef<T>*>(op.coef32);
    v.eps = reinterpret_cast<const T*>(op.eps32);
    v.inv_mu = reinterpret_cast<const T*>(op.inv_mu32);
  } else {
    v.coef = reinterpret_cast<const AxisCoef<T>*>(op.coef64);
    v.eps = reinterpret_cast<const T*>(op.eps64);
    v.inv_mu = reinterpret_cast<const T*>(op.inv_mu64);
  }
  v.mask = op.mask;
  return v;
}

// The automatic tile height is fixed: a system's block partition (and with it the order
// of every reduction) must depend on that system alone, never on the rest of a batch.
constexpr int kAutoTileRows = 4;
constexpr int kStallWindow = 2000;    // default stall window of MX_MIXED
constexpr int kRefreshIters = 2000;   // forced FP64 residual replacement cadence

// Replacement cadence: the stall rule is evaluated at replacements, so they come at
// least once per stall window.
int refresh_iters(int stall_iters) {
  return stall_iters > 0 ? std::min(stall_iters, kRefreshIters) : kRefreshIters;
}

// Material kind of an operator (common.hpp): the apply kernel its launches use.
int op_kind(const mx_operator& op) { return mat_kind(view_of<double>(op).flags); }

// One finishing launch per fused iteration (default); MEANAS_AMD_MERGED_FIN=0 launches
// the post-apply and post-update finishing kernels separately (same results).
bool merged_fin_enabled() {
  const char* v = getenv("MEANAS_AMD_MERGED_FIN");
  return !(v && v[0] == '0');
}

// Automatic slab depth: enough blocks to fill 256 CUs several times over, but deep slabs
// to amortise the two extra planes a slab reads.  With `reduce` (kernels that leave
// per-block partial sums) the block count must also stay within what one finishing
// block sums; returns -1 when even a full-depth slab (one block per tile) exceeds it,
// i.e. when the (y, z) cross-section alone has more than kMaxPartials tiles.
int auto_slab(const mx_operator& op, int ty, bool reduce) {
  const long long tiles = (long long)((op.ny + ty - 1) / ty) * ((op.nz + 63) / 64);
  int slab = 16;
  while (slab > 4 && tiles * ((op.nx + slab - 1) / slab) < 2048) slab /= 2;
  if (reduce)
    while (slab < op.nx && tiles * ((op.nx + slab - 1) / slab) > kMaxPartials) slab *= 2;
  slab = std::min(slab, op.nx);
  if (reduce && tiles * ((op.nx + slab - 1) / slab) > kMaxPartials) return -1;
  return slab;
}

std::string too_many_tiles(const mx_operator& op, int ty) {
  const long long tiles = (long long)((op.ny + ty - 1) / ty) * ((op.nz + 63) / 64);
  return "the " + std::to_string(op.ny) + " x " + std::to_string(op.nz) + " cross-section needs " +
         std::to_string(tiles) + " apply tiles of " + std::to_string(ty) + " x 64 cells, more than the " +
         std::to_string(kMaxPartials) + " per-system partial sums one finishing block reduces";
}

int apply_blocks(const mx_operator& op, int ty, int slab) {
  const long long tiles = (long long)((op.ny + ty - 1) / ty) * ((op.nz + 63) / 64);
  return (int)(tiles * ((op.nx + slab - 1) / slab));
}

int apply_blocks_upper(const mx_operator& op) {   // ty = 2, slab = 1
  return (int)((long long)((op.ny + 1) / 2) * ((op.nz + 63) / 64) * op.nx);
}

int vec_blocks(long long n3) {
  long long nb = (n3 + (long long)kVecBlock * kVecElemsPerThread - 1) / ((long long)kVecBlock * kVecElemsPerThread);
  return (int)std::max(1LL, std::min<long long>(nb, kVecMaxBlocks));
}

size_t elem_bytes(int mode) { return mode == MX_MIXED ? 8 : 16; }

// workspace layout per system: r, p, q, x_lo (inner precision), x_best (FP64), state,
// partials (apply and reductions), the update's partials
struct WorkLayout {
  size_t off_r, off_p, off_q, off_xlo, off_xbest, off_state, off_partial, off_partial_vec, total;
};

WorkLayout work_layout(const mx_operator& op, int mode) {
  const long long n3 = 3 * ncell_of(op);
  const size_t vb = align_up((size_t)n3 * elem_bytes(mode));
  WorkLayout w;
  w.off_r = 0;
  w.off_p = w.off_r + vb;
  w.off_q = w.off_p + vb;
  w.off_xlo = w.off_q + vb;
  w.off_xbest = w.off_xlo + vb;
  w.off_state = w.off_xbest + align_up((size_t)n3 * 16);
  w.off_partial = w.off_state + align_up(sizeof(SysState));
  const size_t nbmax = (size_t)std::max(apply_blocks_upper(op), kVecMaxBlocks);
  w.off_partial_vec = w.off_partial + align_up(4 * nbmax * sizeof(double));
  w.total = w.off_partial_vec + align_up(3 * (size_t)kVecMaxBlocks * sizeof(double));
  return w;
}

// Order the blocks of a grid so that each XCD (blocks are dispatched round-robin over
// 8 XCDs) receives a contiguous run of neighbouring tiles: better L2 reuse of halos and
// of shared materials.  `natural` is the work list in spatial order.
std::vector<BlockMap> xcd_order(const std::vector<BlockMap>& natural) {
  const size_t n = natural.size();
  constexpr size_t kXcd = 8;
  std::vector<BlockMap> out(n);
  const size_t per = n / kXcd, rem = n % kXcd;
  // blockIdx w -> XCD (w % 8), position (w / 8) within that XCD's run
  std::vector<size_t> start(kXcd + 1, 0);
  for (size_t x = 0; x < kXcd; ++x) start[x + 1] = start[x] + per + (x < rem ? 1 : 0);
  for (size_t w = 0; w < n; ++w) {
    const size_t x = w % kXcd, pos = w / kXcd;
    out[w] = natural[start[x] + pos];
  }
  return out;
}

// ---- a batch of systems with device-resident tables ---------------------------------
// Launch groups.  Systems are launched in groups whose hot data -- the vectors the apply
// and the update pass between each other (p, q) and the materials -- fits the 256 MB
// Infinity Cache (MALL); a captured chunk runs all of its iterations on one group before
// the next, so each group's p, q and eps stay cache-resident from kernel to kernel as
// in a single solve (x_lo and r are streamed past the cache by the update).  A group of
// many small systems also fills the GPU that one of them would leave idle.  A system's
// block partition does not depend on its group, so results do not either.
// reps: times a chunk repeats this group's iterations (sessions: equal time per group)
struct Group { int a_off, na, v_off, nv, s_off, ns, ty, kind, reps = 1; };

// MALL bytes a launch group may fill (MEANAS_AMD_CACHE_MB overrides; 0: one group)
size_t cache_budget() {
  static const size_t b = [] {
    const char* v = getenv("MEANAS_AMD_CACHE_MB");
    const double mb = v ? atof(v) : 224.0;
    return mb > 0.0 ? (size_t)(mb * (1u << 20)) : (size_t)0;
  }();
  return b;
}

// bytes of one system's p and q, and of its operator's materials, in the inner precision
size_t vec_hot_bytes(const mx_operator& op, int mode) { return 2 * (size_t)(3 * ncell_of(op)) * elem_bytes(mode); }
size_t mat_hot_bytes(const mx_operator& op, int mode) {
  const size_t n = (size_t)ncell_of(op), real = mode == MX_MIXED ? 4 : 8;
  auto mat = [&](bool cplx, bool aniso) { return (aniso ? 3 : 1) * n * real * (cplx ? 2 : 1); };
  size_t b = mat(op.flags & MX_EPS_COMPLEX, op.flags & MX_EPS_ANISO);
  if (op.flags & MX_MU) b += mat(op.flags & MX_MU_COMPLEX, op.flags & MX_MU_ANISO);
  if (op.flags & MX_MASK) b += n;
  return b;
}

struct Batch {
  int mode = MX_MIXED;
  int ty = 4;
  int nops = 0, nsys = 0;
  int nb_apply = 0, nb_vec = 0;
  char* d_mem = nullptr;
  OpView<float>* d_ops32 = nullptr;
  OpView<double>* d_ops64 = nullptr;
  SysDesc* d_sys = nullptr;
  BlockMap* d_bmap_apply = nullptr;
  BlockMap* d_bmap_vec = nullptr;
  int* d_syslist = nullptr;   // systems the finishing kernels visit (one block each)
  int nsl = 0;
  // the whole (all groups) maps; d_bmap_apply etc. point at one group while it is enqueued
  BlockMap* bm_apply0 = nullptr;
  BlockMap* bm_vec0 = nullptr;
  int* syslist0 = nullptr;
  int nb_apply0 = 0, nb_vec0 = 0, nsl0 = 0;
  std::vector<Group> groups;
  std::vector<SysDesc> h_sys;
  std::vector<int> h_op;        // operator index of each system
  std::vector<int> h_ty;        // apply tile height of each system
  const mx_operator* ops = nullptr;
  SysState* h_state[2] = {nullptr, nullptr};   // pinned, double-buffered polls
  hipStream_t stream = nullptr;                  // library-owned capture stream
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_chunk = 0;
  bool classic_launch = true;   // enqueue the classic-form direction update every iteration
  bool merged_fin = false;      // every system stays fused: one finishing launch per iteration
  int kind = MK_GENERAL;        // material kind of every operator (MK_GENERAL when they differ)

  void use_group(const Group& g) {
    ty = g.ty; kind = g.kind;
    d_bmap_apply = bm_apply0 + g.a_off; nb_apply = g.na;
    d_bmap_vec = bm_vec0 + g.v_off; nb_vec = g.nv;
    d_syslist = syslist0 + g.s_off; nsl = g.ns;
  }
  void use_all() {
    d_bmap_apply = bm_apply0; nb_apply = nb_apply0;
    d_bmap_vec = bm_vec0; nb_vec = nb_vec0;
    d_syslist = syslist0; nsl = nsl0;
  }

  ~Batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    for (auto& h : h_state) if (h) (void)hipHostFree(h);
    if (d_mem) (void)hipFree(d_mem);
  }
};

// One system as a member of a launch group.
struct Member {
  int sys;                  // system (batch) or slot (session) index
  const mx_operator* op;
  int op_id;                // systems with the same op_id share materials
  int ty;                   // apply tile height
  int kind;                 // material kind (op_kind)
  int na, nv;               // apply and vector blocks
};

// Lay `members` out as consecutive launch groups (see Group): one tile height and one
// material kind per group, and within those, consecutive members
// while their hot bytes fit the cache budget.  Each group's maps are XCD-ordered on their
// own; apply, vec and sl receive the concatenated maps and system lists.
void make_groups(const std::vector<Member>& members, int mode, std::vector<BlockMap>& apply,
                 std::vector<BlockMap>& vec, std::vector<int>& sl, std::vector<Group>& groups) {
  const size_t budget = cache_budget();
  apply.clear(); vec.clear(); sl.clear(); groups.clear();
  for (int ty : {2, 4, 8})
    for (int kind : {MK_SIMPLE, MK_ANISO, MK_MASK, MK_GENERAL}) {
      std::vector<BlockMap> ga, gv;
      std::vector<int> gs, gops;
      size_t bytes = 0;
      auto close = [&]() {
        if (gs.empty()) return;
        Group g;
        g.a_off = (int)apply.size(); g.v_off = (int)vec.size(); g.s_off = (int)sl.size();
        const std::vector<BlockMap> oa = xcd_order(ga), ov = xcd_order(gv);
        apply.insert(apply.end(), oa.begin(), oa.end());
        vec.insert(vec.end(), ov.begin(), ov.end());
        sl.insert(sl.end(), gs.begin(), gs.end());
        g.na = (int)oa.size(); g.nv = (int)ov.size(); g.ns = (int)gs.size();
        g.ty = ty; g.kind = kind;
        groups.push_back(g);
        ga.clear(); gv.clear(); gs.clear(); gops.clear();
        bytes = 0;
      };
      for (const Member& m : members) {
        if (m.ty != ty || m.kind != kind) continue;
        const bool new_op = std::find(gops.begin(), gops.end(), m.op_id) == gops.end();
        const size_t add = vec_hot_bytes(*m.op, mode) + (new_op ? mat_hot_bytes(*m.op, mode) : 0);
        if (budget > 0 && !gs.empty() && bytes + add > budget) close();
        if (std::find(gops.begin(), gops.end(), m.op_id) == gops.end()) gops.push_back(m.op_id);
        bytes += add;
        gs.push_back(m.sys);
        for (int b = 0; b < m.na; ++b) ga.push_back({m.sys, b});
        for (int b = 0; b < m.nv; ++b) gv.push_back({m.sys, b});
      }
      close();
    }
}

// Lay out the launch groups of the `active` systems of a batch and copy the maps to the
// device (synchronous; the caller has drained the stream).
int layout_groups(Batch& B, const std::vector<char>& active) {
  std::vector<Member> members;
  for (int s = 0; s < B.nsys; ++s)
    if (active[s])
      members.push_back({s, &B.ops[B.h_op[s]], B.h_op[s], B.h_ty[s], op_kind(B.ops[B.h_op[s]]),
                         B.h_sys[s].nblk_apply, B.h_sys[s].nblk_vec});
  std::vector<BlockMap> apply, vec;
  std::vector<int> sl;
  make_groups(members, B.mode, apply, vec, sl, B.groups);
  B.nb_apply0 = (int)apply.size();
  B.nb_vec0 = (int)vec.size();
  B.nsl0 = (int)sl.size();
  if (!apply.empty()) MX_TRY(copy_now(B.bm_apply0, apply.data(), sizeof(BlockMap) * apply.size(), hipMemcpyHostToDevice));
  if (!vec.empty()) MX_TRY(copy_now(B.bm_vec0, vec.data(), sizeof(BlockMap) * vec.size(), hipMemcpyHostToDevice));
  if (!sl.empty()) MX_TRY(copy_now(B.syslist0, sl.data(), sizeof(int) * sl.size(), hipMemcpyHostToDevice));
  B.use_all();
  return 0;
}

int build_batch(Batch& B, const mx_operator* ops, int nops, const mx_system* sys, int nsys, const mx_options& o) {
  std::string why;
  B.mode = o.mode;
  B.nops = nops;
  B.nsys = nsys;
  B.ops = ops;
  const int ty_default = o.tile_rows ? o.tile_rows : kAutoTileRows;
  if (ty_default != 2 && ty_default != 4 && ty_default != 8) return fail("tile_rows must be 2, 4 or 8");
  B.ty = ty_default;
  for (int i = 0; i < nops; ++i)
    if (!valid_op(ops[i], o.mode, &why)) return fail("operator " + std::to_string(i) + ": " + why);

  size_t total_apply = 0, total_vec = 0;
  B.h_sys.resize(nsys);
  B.h_op.resize(nsys);
  B.h_ty.resize(nsys);
  for (int s = 0; s < nsys; ++s) {
    const mx_system& y = sys[s];
    if (y.op < 0 || y.op >= nops) return fail("system " + std::to_string(s) + ": bad operator index");
    const int ty = y.tile_rows ? y.tile_rows : ty_default;
    if (ty != 2 && ty != 4 && ty != 8) return fail("system " + std::to_string(s) + ": tile_rows must be 2, 4 or 8");
    if (!y.b || !y.x || !y.work) return fail("system " + std::to_string(s) + ": b, x and work are required");
    const mx_operator& op = ops[y.op];
    const WorkLayout w = work_layout(op, o.mode);
    char* wk = reinterpret_cast<char*>(y.work);
    SysDesc d;
    std::memset(&d, 0, sizeof(d));
    d.op = y.op;
    d.slab = y.slab > 0 ? std::min(y.slab, op.nx) : o.slab > 0 ? std::min(o.slab, op.nx) : auto_slab(op, ty, true);
    if (d.slab < 1) return fail("system " + std::to_string(s) + ": " + too_many_tiles(op, ty));
    d.nblk_apply = apply_blocks(op, ty, d.slab);
    d.n3 = 3 * ncell_of(op);
    d.nblk_vec = vec_blocks(d.n3);
    if (d.nblk_apply > kMaxPartials)
      return fail("system " + std::to_string(s) + ": " + std::to_string(d.nblk_apply) +
                  " apply blocks exceed the reduction limit; use a deeper slab");
    d.vec[VS_X] = y.x;
    d.vec[VS_B] = const_cast<void*>(y.b);
    d.vec[VS_R] = wk + w.off_r;
    d.vec[VS_P] = wk + w.off_p;
    d.vec[VS_Q] = wk + w.off_q;
    d.vec[VS_XLO] = wk + w.off_xlo;
    d.vec[VS_XBEST] = wk + w.off_xbest;
    d.st = reinterpret_cast<SysState*>(wk + w.off_state);
    d.partial = reinterpret_cast<double*>(wk + w.off_partial);
    d.partial_vec = reinterpret_cast<double*>(wk + w.off_partial_vec);
    B.h_sys[s] = d;
    B.h_op[s] = y.op;
    B.h_ty[s] = ty;
    total_apply += d.nblk_apply;
    total_vec += d.nblk_vec;
  }

  const size_t s_ops32 = align_up(sizeof(OpView<float>) * nops);
  const size_t s_ops64 = align_up(sizeof(OpView<double>) * nops);
  const size_t s_sys = align_up(sizeof(SysDesc) * nsys);
  const size_t s_ba = align_up(sizeof(BlockMap) * total_apply);
  const size_t s_bv = align_up(sizeof(BlockMap) * total_vec);
  const size_t s_sl = align_up(sizeof(int) * nsys);
  MX_TRY(hipMalloc(&B.d_mem, s_ops32 + s_ops64 + s_sys + s_ba + s_bv + s_sl));
  char* p = B.d_mem;
  B.d_ops32 = reinterpret_cast<OpView<float>*>(p); p += s_ops32;
  B.d_ops64 = reinterpret_cast<OpView<double>*>(p); p += s_ops64;
  B.d_sys = reinterpret_cast<SysDesc*>(p); p += s_sys;
  B.bm_apply0 = reinterpret_cast<BlockMap*>(p); p += s_ba;
  B.bm_vec0 = reinterpret_cast<BlockMap*>(p); p += s_bv;
  B.syslist0 = reinterpret_cast<int*>(p);

  std::vector<OpView<float>> v32(nops);
  std::vector<OpView<double>> v64(nops);
  B.kind = nops > 0 ? op_kind(ops[0]) : MK_GENERAL;
  for (int i = 0; i < nops; ++i) {
    if (op_kind(ops[i]) != B.kind) B.kind = MK_GENERAL;
    v64[i] = view_of<double>(ops[i]);
    if (o.mode == MX_MIXED) v32[i] = view_of<float>(ops[i]);
    else std::memset(&v32[i], 0, sizeof(v32[i]));
  }
  MX_TRY(copy_now(B.d_ops32, v32.data(), sizeof(OpView<float>) * nops, hipMemcpyHostToDevice));
  MX_TRY(copy_now(B.d_ops64, v64.data(), sizeof(OpView<double>) * nops, hipMemcpyHostToDevice));
  MX_TRY(copy_now(B.d_sys, B.h_sys.data(), sizeof(SysDesc) * nsys, hipMemcpyHostToDevice));
  if (layout_groups(B, std::vector<char>(nsys, 1))) return -1;

  for (auto& h : B.h_state) MX_TRY(hipHostMalloc(reinterpret_cast<void**>(&h), sizeof(SysState) * nsys, hipHostMallocDefault));
  MX_TRY(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
  for (auto& e : B.ev) MX_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return 0;
}

// one COCG iteration over every runnable system of the batch (the direction update
// launch only acts on systems in the classic form; see vector.hip)
int enqueue_iteration(Batch& B, hipStream_t s) {
  if (B.mode == MX_MIXED) {
    MX_TRY(launch_pupdate_f32(B.d_ops32, B.d_sys, B.d_bmap_vec, B.nb_vec, 0, s));
    MX_TRY(launch_apply_cocg_f32(B.ty, B.kind, B.d_ops32, B.d_sys, B.d_bmap_apply, B.nb_apply, s));
    MX_TRY(launch_fin(FIN_APPLY, B.d_sys, B.d_syslist, B.nsl, s));
    MX_TRY(launch_update_f32(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
  } else {
    if (B.merged_fin) {   // the previous update is finished together with this apply
      MX_TRY(launch_apply_cocg_f64(B.ty, B.kind, B.d_ops64, B.d_sys, B.d_bmap_apply, B.nb_apply, s));
      MX_TRY(launch_fin(FIN_APPLY_MERGED, B.d_sys, B.d_syslist, B.nsl, s));
      MX_TRY(launch_update_f64(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
      return 0;
    }
    if (B.classic_launch) MX_TRY(launch_pupdate_f64(B.d_ops64, B.d_sys, B.d_bmap_vec, B.nb_vec, 0, s));
    MX_TRY(launch_apply_cocg_f64(B.ty, B.kind, B.d_ops64, B.d_sys, B.d_bmap_apply, B.nb_apply, s));
    MX_TRY(launch_fin(FIN_APPLY, B.d_sys, B.d_syslist, B.nsl, s));
    MX_TRY(launch_update_f64(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
  }
  MX_TRY(launch_fin(FIN_UPDATE, B.d_sys, B.d_syslist, B.nsl, s));
  return 0;
}

// residual replacement on paused systems: x += x_lo ; r = b - A x (FP64); then the
// search direction of fused-form systems that restart or take a rescue step
int enqueue_replacement(Batch& B, hipStream_t s) {
  if (B.mode == MX_MIXED) {
    MX_TRY(launch_accum_f32(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
    MX_TRY(launch_apply_resid_r32(B.ty, B.kind, B.d_ops64, B.d_sys, B.d_bmap_apply, B.nb_apply, s));
  } else {
    MX_TRY(launch_accum_f64(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
    MX_TRY(launch_apply_resid_r64(B.ty, B.kind, B.d_ops64, B.d_sys, B.d_bmap_apply, B.nb_apply, s));
  }
  MX_TRY(launch_fin(FIN_RESID, B.d_sys, B.d_syslist, B.nsl, s));
  MX_TRY(launch_save_best(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
  if (B.mode == MX_FP64) MX_TRY(launch_pupdate_f64(B.d_ops64, B.d_sys, B.d_bmap_vec, B.nb_vec, 1, s));
  return 0;
}

int enqueue_group_chunk(Batch& B, hipStream_t s, int iters, int check_every) {
  for (int it = 0; it < iters; ++it) {
    if (enqueue_iteration(B, s)) return -1;
    if ((it + 1) % check_every == 0 || it + 1 == iters) {
      // merged launches: finish the last update, so that the state the replacement and
      // the host read is current
      if (B.merged_fin) MX_TRY(launch_fin(FIN_UPDATE, B.d_sys, B.d_syslist, B.nsl, s));
      if (enqueue_replacement(B, s)) return -1;
    }
  }
  return 0;
}

// residual replacement on the paused systems of every launch group
int enqueue_replacement_groups(Batch& B, hipStream_t s) {
  if (B.groups.empty()) return enqueue_replacement(B, s);
  int rc = 0;
  for (const Group& g : B.groups) {
    B.use_group(g);
    rc = enqueue_replacement(B, s);
    if (rc) break;
  }
  B.use_all();
  return rc;
}

// `iters` iterations of every launch group (`reps` times that for a group with reps > 1),
// one group after the other (a batch view without groups is one group).  A repetition
// is the same chunk with the same check points, so a system's iterates do not depend on
// how often its group repeats.
int enqueue_chunk(Batch& B, hipStream_t s, int iters, int check_every) {
  if (B.groups.empty()) return enqueue_group_chunk(B, s, iters, check_every);
  int rc = 0;
  for (const Group& g : B.groups) {
    B.use_group(g);
    for (int r = 0; r < g.reps && !rc; ++r) rc = enqueue_group_chunk(B, s, iters, check_every);
    if (rc) break;
  }
  B.use_all();
  return rc;
}

// ---- solve plans ---------------------------------------------------------------
// A batch's device tables, block maps, stream, events, pinned poll buffers and captured
// chunk graph depend only on its operators, its systems' vectors and tilings and the
// options -- not on the stopping rules, which enter the per-call initial state.  The
// last mx_solve of a thread keeps them, and a call with the same inputs reuses them:
// no allocation, table upload or graph capture (MEANAS_AMD_PLAN_CACHE=0 builds every
// batch afresh).  The tables hold the caller's device pointers; a kept plan is only
// used again by a call that passes the same pointers, which then are valid.
bool plan_cache_enabled() {
  const char* v = getenv("MEANAS_AMD_PLAN_CACHE");
  return !(v && v[0] == '0');
}

std::string plan_key(const mx_operator* ops, int nops, const mx_system* sys, int nsys, const mx_options& o, int dev) {
  std::string k;
  auto put = [&](const void* p, size_t n) { k.append(reinterpret_cast<const char*>(p), n); };
  put(&dev, sizeof dev);
  put(&nops, sizeof nops);
  put(&nsys, sizeof nsys);
  put(ops, sizeof(mx_operator) * (size_t)nops);
  for (int i = 0; i < nsys; ++i) {
    mx_system y = sys[i];
    y.rtol = 0.0;
    y.atol = 0.0;
    y.maxiter = 0;
    put(&y, sizeof y);
  }
  put(&o, sizeof o);
  const char sw = (char)merged_fin_enabled();
  put(&sw, sizeof sw);
  return k;
}

struct SolvePlan {
  std::string key;
  Batch* B = nullptr;
  bool busy = false;       // in use by a call of this thread (the classic re-solve nests)
  bool relayout = false;   // its block maps were compacted by that call
};
// Never destroyed: at process exit the HIP runtime may already be gone.
thread_local SolvePlan* tl_plan = nullptr;
thread_local int64_t tl_plan_reuses = 0;

// The batch one mx_solve call runs on: the thread's kept plan when the inputs match,
// else a new one, which is kept afterwards if the call succeeds.  An outer call's plan
// is never touched by a nested call.  A failed call drops the plan it used.
class PlanUse {
 public:
  Batch* B = nullptr;

  int acquire(const mx_operator* ops, int nops, const mx_system* sys, int nsys, const mx_options& o) {
    if (!tl_plan) tl_plan = new SolvePlan;
    if (!plan_cache_enabled() || tl_plan->busy) {
      owned_.reset(new Batch);
      B = owned_.get();
      return build_batch(*B, ops, nops, sys, nsys, o);
    }
    tl_plan->busy = true;
    owner_ = true;
    int dev = 0;
    if (tl_plan->B) {   // (the arguments are checked by build_batch before any HIP call)
      MX_TRY(hipGetDevice(&dev));
      key_ = plan_key(ops, nops, sys, nsys, o, dev);
    }
    if (tl_plan->B && tl_plan->key == key_) {
      B = tl_plan->B;
      reused_ = true;
      ++tl_plan_reuses;
      B->ops = ops;
      if (tl_plan->relayout) {   // back to every system's blocks, and a new graph
        tl_plan->relayout = false;
        if (B->gexec) { MX_TRY(hipGraphExecDestroy(B->gexec)); B->gexec = nullptr; }
        if (B->graph) { MX_TRY(hipGraphDestroy(B->graph)); B->graph = nullptr; }
        if (layout_groups(*B, std::vector<char>(nsys, 1))) return -1;
      }
      return 0;
    }
    drop();
    owned_.reset(new Batch);
    B = owned_.get();
    if (build_batch(*B, ops, nops, sys, nsys, o)) return -1;
    if (key_.empty()) {
      MX_TRY(hipGetDevice(&dev));
      key_ = plan_key(ops, nops, sys, nsys, o, dev);
    }
    return 0;
  }
  // the call succeeded; `compacted`: its block maps no longer cover every system
  void keep(bool compacted) { ok_ = true; compacted_ = compacted; }

  ~PlanUse() {
    if (!owner_) return;
    tl_plan->busy = false;
    if (!ok_) {
      if (reused_) drop();
      return;
    }
    if (!reused_) {
      tl_plan->B = owned_.release();
      tl_plan->key = key_;
    }
    tl_plan->relayout = compacted_;
  }

 private:
  static void drop() {
    delete tl_plan->B;
    tl_plan->B = nullptr;
    tl_plan->key.clear();
    tl_plan->relayout = false;
  }
  std::unique_ptr<Batch> owned_;
  std::string key_;
  bool owner_ = false, reused_ = false, ok_ = false, compacted_ = false;
};

// A finished system that did not converge returns its best iterate (restored by
// restore_best); its reported residual is that iterate's.
bool wants_best(const SysState& st) {
  return (st.flags & SF_DONE) && !(st.true_rnorm2 <= st.stop2) && st.best_true2 < st.true_rnorm2 &&
         !(st.flags & SF_NAN);
}

void fill_stats(const SysState& st, mx_stats* out) {
  out->iterations = st.iters;
  out->replacements = st.replacements;
  out->flags = st.flags;
  out->bnorm = std::sqrt(st.bnorm2);
  const double r2 = wants_best(st) ? st.best_true2 : st.true_rnorm2;
  out->rel_residual = st.bnorm2 > 0 ? std::sqrt(r2 / st.bnorm2) : std::sqrt(r2);
  if (st.flags & SF_NAN) out->status = MX_NONFINITE;
  else if (st.true_rnorm2 <= st.stop2) out->status = MX_CONVERGED;
  else if (st.flags & SF_STALL) out->status = MX_STALLED;
  else if (st.flags & SF_BREAKDOWN) out->status = MX_BREAKDOWN;
  else out->status = MX_MAXITER;
}

bool valid_options(const mx_options& o, std::string* why) {
  if (o.mode != MX_MIXED && o.mode != MX_FP64) { *why = "mode must be MX_MIXED or MX_FP64"; return false; }
  if (o.form != MX_FORM_FUSED && o.form != MX_FORM_CLASSIC) { *why = "form must be MX_FORM_AUTO, MX_FORM_FUSED or MX_FORM_CLASSIC"; return false; }
  if (o.form == MX_FORM_FUSED && o.mode != MX_FP64) { *why = "the fused form needs complex128 inner vectors (MX_FP64)"; return false; }
  return true;
}

mx_options defaults(const mx_options* opt) {
  mx_options o;
  std::memset(&o, 0, sizeof(o));
  if (opt) o = *opt;
  if (o.chunk <= 0) o.chunk = 16;
  if (o.check_every <= 0) o.check_every = (o.mode == MX_MIXED) ? 1 : o.chunk;
  if (o.delta <= 0) o.delta = (o.mode == MX_MIXED) ? 0.1 : 0.0;
  // The stall window only pays where a stall hands the system to a stronger solver
  // (mixed -> FP64).  FP64 COCG on PML problems crosses plateaus of several thousand
  // iterations, so FP64 runs to maxiter unless the caller sets a window.
  if (o.stall_iters == 0) o.stall_iters = (o.mode == MX_MIXED) ? kStallWindow : -1;
  if (o.stall_limit == 0) o.stall_limit = 3;
  if (o.form == MX_FORM_AUTO) o.form = (o.mode == MX_FP64) ? MX_FORM_FUSED : MX_FORM_CLASSIC;
  return o;
}

}  // namespace

// =====================================================================================
namespace mx {
// error reporting for the other translation units of the library (mx_last_error)
int set_error(const std::string& msg) { return fail(msg); }
}  // namespace mx

extern "C" {

int mx_abi_version(void) { return MX_ABI_VERSION; }

const char* mx_last_error(void) { return g_err.c_str(); }

size_t mx_work_bytes(const mx_operator* op, int32_t mode) {
  if (!op || (mode != MX_MIXED && mode != MX_FP64)) return 0;
  return work_layout(*op, mode).total;
}

// ---- apply plans -----------------------------------------------------------------
struct mx_plan {
  int mode, ty, nb;
  char* d_mem;
  int kind;
};

mx_plan* mx_plan_create(const mx_operator* op, int32_t mode, int32_t tile_rows, int32_t slab) {
  std::string why;
  if (!op) { fail("null operator"); return nullptr; }
  if (mode != MX_MIXED && mode != MX_FP64) { fail("mode must be MX_MIXED or MX_FP64"); return nullptr; }
  if (!valid_op(*op, mode, &why)) { fail(why); return nullptr; }
  const int ty = tile_rows ? tile_rows : kAutoTileRows;
  if (ty != 2 && ty != 4 && ty != 8) { fail("tile_rows must be 2, 4 or 8"); return nullptr; }
  SysDesc d;
  std::memset(&d, 0, sizeof(d));
  d.op = 0;
  // a plan's apply leaves no partial sums, so its block count has no reduction limit
  d.slab = slab > 0 ? std::min<int>(slab, op->nx) : auto_slab(*op, ty, false);
  d.nblk_apply = apply_blocks(*op, ty, d.slab);
  d.n3 = 3 * ncell_of(*op);
  std::vector<BlockMap> nat;
  for (int b = 0; b < d.nblk_apply; ++b) nat.push_back({0, b});
  std::vector<BlockMap> bm = xcd_order(nat);
  const size_t s_op = align_up(std::max(sizeof(OpView<float>), sizeof(OpView<double>)));
  const size_t s_sys = align_up(sizeof(SysDesc));
  const size_t s_bm = align_up(sizeof(BlockMap) * bm.size());
  mx_plan* P = new (std::nothrow) mx_plan{mode, ty, d.nblk_apply, nullptr, op_kind(*op)};
  if (!P) { fail("out of host memory"); return nullptr; }
  hipError_t e = hipMalloc(&P->d_mem, s_op + s_sys + s_bm);
  if (e != hipSuccess) { fail(std::string("hipMalloc: ") + hipGetErrorString(e)); delete P; return nullptr; }
  if (mode == MX_MIXED) {
    OpView<float> v = view_of<float>(*op);
    e = copy_now(P->d_mem, &v, sizeof(v), hipMemcpyHostToDevice);
  } else {
    OpView<double> v = view_of<double>(*op);
    e = copy_now(P->d_mem, &v, sizeof(v), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = copy_now(P->d_mem + s_op, &d, sizeof(d), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = copy_now(P->d_mem + s_op + s_sys, bm.data(), sizeof(BlockMap) * bm.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) { fail(std::string("hipMemcpy: ") + hipGetErrorString(e)); (void)hipFree(P->d_mem); delete P; return nullptr; }
  return P;
}

int mx_plan_apply(mx_plan* P, const void* x, void* y, void* stream) {
  if (!P || !x || !y) return fail("null argument");
  const size_t s_op = align_up(std::max(sizeof(OpView<float>), sizeof(OpView<double>)));
  const size_t s_sys = align_up(sizeof(SysDesc));
  const SysDesc* sys = reinterpret_cast<const SysDesc*>(P->d_mem + s_op);
  const BlockMap* bm = reinterpret_cast<const BlockMap*>(P->d_mem + s_op + s_sys);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (P->mode == MX_MIXED)
    MX_TRY(launch_apply_plain_f32(P->ty, P->kind, reinterpret_cast<const OpView<float>*>(P->d_mem), sys, bm, P->nb, x, y, s));
  else
    MX_TRY(launch_apply_plain_f64(P->ty, P->kind, reinterpret_cast<const OpView<double>*>(P->d_mem), sys, bm, P->nb, x, y, s));
  return 0;
}

int mx_plan_blocks(const mx_plan* P) { return P ? P->nb : -1; }

void mx_plan_destroy(mx_plan* P) {
  if (!P) return;
  if (P->d_mem) (void)hipFree(P->d_mem);
  delete P;
}

int mx_apply(const mx_operator* op, int32_t mode, const void* x, void* y, void* stream) {
  mx_plan* P = mx_plan_create(op, mode, 0, 0);
  if (!P) return -1;
  int rc = mx_plan_apply(P, x, y, stream);
  hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
  mx_plan_destroy(P);
  if (rc) return rc;
  if (e != hipSuccess) return fail(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  return 0;
}

int mx_map(const mx_operator* op, int32_t map, double c_re, double c_im, const void* in, void* out, void* stream) {
  if (!op || !in || !out || !op->coef64) return fail("null argument");
  if (map < 0 || map > 4) return fail("unknown map");
  MX_TRY(launch_diag(reinterpret_cast<const AxisCoef<double>*>(op->coef64), op->nx, op->ny, op->nz, in, out, c_re, c_im,
                     map, reinterpret_cast<hipStream_t>(stream)));
  return 0;
}

}  // extern "C"

namespace {

// Device memory of one FP64 residual r = b - A x: the operator and system descriptors,
// the system state, the per-block partial sums (3 per apply block), the block map and a
// one-entry system list, each aligned.
struct ResidLayout { size_t s_op, s_sys, s_st, s_part, s_bm, total; };
ResidLayout resid_layout(int nblk) {
  ResidLayout L;
  L.s_op = align_up(sizeof(OpView<double>));
  L.s_sys = align_up(sizeof(SysDesc));
  L.s_st = align_up(sizeof(SysState));
  L.s_part = align_up(sizeof(double) * 3 * (size_t)nblk);
  L.s_bm = align_up(sizeof(BlockMap) * (size_t)nblk);
  L.total = L.s_op + L.s_sys + L.s_st + L.s_part + L.s_bm + align_up(sizeof(int));
  return L;
}

// Checks and the tiling of a residual; fills d (vectors, tiling) or fails.
int resid_setup(const mx_operator* op, const void* b, const void* x, void* r, const char* who, SysDesc* d) {
  if (!op || !b || !x || !r) return fail(std::string(who) + ": null argument");
  std::string why;
  if (!valid_op(*op, MX_FP64, &why)) return fail(std::string(who) + ": " + why);
  std::memset(d, 0, sizeof(*d));
  d->slab = auto_slab(*op, kAutoTileRows, true);
  if (d->slab < 1) return fail(std::string(who) + ": " + too_many_tiles(*op, kAutoTileRows));
  d->nblk_apply = apply_blocks(*op, kAutoTileRows, d->slab);
  d->n3 = 3 * ncell_of(*op);
  d->nblk_vec = vec_blocks(d->n3);
  d->vec[VS_X] = const_cast<void*>(x);
  d->vec[VS_B] = const_cast<void*>(b);
  d->vec[VS_R] = r;
  return 0;
}

// r = b - A x and ||r||^2 with all device-side state in `mem` (resid_layout bytes):
// tables staged in stream order, the residual kernel and its finishing kernel, then the
// state read back (the call returns when the stream has reached that point).
int resid_run(const mx_operator& op, SysDesc d, char* mem, double* norm2_out, hipStream_t s, const char* who) {
  const ResidLayout L = resid_layout(d.nblk_apply);
  char* p_sys = mem + L.s_op;
  d.st = reinterpret_cast<SysState*>(p_sys + L.s_sys);
  d.partial = reinterpret_cast<double*>(p_sys + L.s_sys + L.s_st);
  BlockMap* bmd = reinterpret_cast<BlockMap*>(p_sys + L.s_sys + L.s_st + L.s_part);
  int* d_one = reinterpret_cast<int*>(reinterpret_cast<char*>(bmd) + L.s_bm);
  std::vector<BlockMap> bm((size_t)d.nblk_apply);
  for (int i = 0; i < d.nblk_apply; ++i) bm[i] = {0, i};
  SysState st;
  std::memset(&st, 0, sizeof(st));
  st.flags = SF_PAUSE | SF_RESTART;   // RESTART: no search direction to carry over
  st.stop2 = -1.0;
  const OpView<double> v = view_of<double>(op);
  const int zero_idx = 0;
  hipError_t e = hipMemcpyAsync(mem, &v, sizeof(v), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(p_sys, &d, sizeof(d), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d.st, &st, sizeof(st), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(bmd, bm.data(), sizeof(BlockMap) * bm.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_one, &zero_idx, sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = launch_apply_resid_r64(kAutoTileRows, op_kind(op), reinterpret_cast<const OpView<double>*>(mem),
                               reinterpret_cast<const SysDesc*>(p_sys), bmd, d.nblk_apply, s);
  if (e == hipSuccess) e = launch_fin(FIN_RESID, reinterpret_cast<const SysDesc*>(p_sys), d_one, 1, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&st, d.st, sizeof(st), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);   // also keeps the host sources alive until copied
  if (e != hipSuccess) return fail(std::string(who) + ": " + hipGetErrorString(e));
  if (norm2_out) *norm2_out = st.true_rnorm2;
  return 0;
}

}  // namespace

// Fixed device workspace of mx_residual_noalloc, one copy per GPU (each device that loads
// the library holds it until the process ends).  Host code never uses this name as a
// device address: it takes the address hipGetSymbolAddress reports for the current device.
constexpr size_t kResidScratchBytes = 192u << 10;
static __device__ __align__(256) char scratch[kResidScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch`, looked up once per device (with that device current).
int scratch_for_current_device(char** out) {
  int dev = -1;
  MX_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("mx_residual_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("mx_residual_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = reinterpret_cast<char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

int mx_residual(const mx_operator* op, const void* b, const void* x, void* r, double* norm2_out, void* stream) {
  SysDesc d;
  if (resid_setup(op, b, x, r, "mx_residual", &d)) return -1;
  char* mem = nullptr;
  MX_TRY(hipMalloc(&mem, resid_layout(d.nblk_apply).total));
  const int rc = resid_run(*op, d, mem, norm2_out, reinterpret_cast<hipStream_t>(stream), "mx_residual");
  (void)hipFree(mem);
  return rc;
}

int mx_residual_noalloc(const mx_operator* op, const void* b, const void* x, void* r, double* norm2_out,
                        void* stream) {
  SysDesc d;
  if (resid_setup(op, b, x, r, "mx_residual_noalloc", &d)) return -1;
  const size_t need = resid_layout(d.nblk_apply).total;
  if (need > kResidScratchBytes)
    return fail("mx_residual_noalloc: the grid needs " + std::to_string(need) + " bytes of workspace, more than the " +
                std::to_string(kResidScratchBytes) + "-byte fixed workspace");
  char* mem = nullptr;
  if (scratch_for_current_device(&mem)) return -1;
  return resid_run(*op, d, mem, norm2_out, reinterpret_cast<hipStream_t>(stream), "mx_residual_noalloc");
}

int mx_time_kernels(const mx_operator* op, const mx_system* sys, const mx_options* opt, int32_t reps,
                    double* us_out, int32_t n_out, void* stream) {
  if (!op || !sys || !us_out || n_out < 1 || reps < 1) return fail("mx_time_kernels: bad argument");
  const mx_options o = defaults(opt);
  std::string why;
  if (!valid_options(o, &why)) return fail(why);
  mx_system y = *sys;
  y.op = 0;
  Batch B;
  if (build_batch(B, op, 1, &y, 1, o)) return -1;
  B.use_group(B.groups[0]);   // one system: one launch group
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const bool fp64 = o.mode == MX_FP64;
  // set-up as in mx_solve, with every stopping and pausing rule off, so each kernel
  // below runs its full work on every launch
  SysState st;
  std::memset(&st, 0, sizeof(st));
  st.flags = SF_PAUSE | SF_RESTART | (o.form == MX_FORM_FUSED ? SF_FUSED : 0);
  st.maxiter = std::numeric_limits<int>::max();
  SysState* dst = B.h_sys[0].st;
  MX_TRY(hipMemcpyAsync(dst, &st, sizeof(st), hipMemcpyHostToDevice, s));
  if (launch_zero_slot(fp64, B.d_sys, B.d_bmap_vec, B.nb_vec, VS_XLO, s) != hipSuccess) return fail("zero x_lo");
  MX_TRY(launch_bnorm(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
  MX_TRY(launch_fin(FIN_BNORM, B.d_sys, B.d_syslist, B.nsl, s));
  if (enqueue_replacement(B, s)) return -1;
  SysState snap;
  MX_TRY(hipMemcpyAsync(&snap, dst, sizeof(snap), hipMemcpyDeviceToHost, s));
  MX_TRY(hipStreamSynchronize(s));
  if (snap.flags & (SF_DONE | SF_PAUSE | SF_MR)) return fail("mx_time_kernels: the system needs a nonzero, generic residual");
  // The iteration runs as in a solve (each kernel right after the one it follows, so
  // caches hold what they would), with an event between consecutive kernels: kernel k's
  // time is the span from the event before it to the event after it.
  constexpr int kKernels = 5;
  auto launch = [&](int k) -> hipError_t {
    switch (k) {
      case 0: return fp64 ? launch_pupdate_f64(B.d_ops64, B.d_sys, B.d_bmap_vec, B.nb_vec, 0, s)
                          : launch_pupdate_f32(B.d_ops32, B.d_sys, B.d_bmap_vec, B.nb_vec, 0, s);
      case 1: return fp64 ? launch_apply_cocg_f64(B.ty, B.kind, B.d_ops64, B.d_sys, B.d_bmap_apply, B.nb_apply, s)
                          : launch_apply_cocg_f32(B.ty, B.kind, B.d_ops32, B.d_sys, B.d_bmap_apply, B.nb_apply, s);
      case 2: return launch_fin(FIN_APPLY, B.d_sys, B.d_syslist, B.nsl, s);
      case 3: return fp64 ? launch_update_f64(B.d_sys, B.d_bmap_vec, B.nb_vec, s)
                          : launch_update_f32(B.d_sys, B.d_bmap_vec, B.nb_vec, s);
      default: return launch_fin(FIN_UPDATE, B.d_sys, B.d_syslist, B.nsl, s);
    }
  };
  std::vector<hipEvent_t> ev((size_t)reps * (kKernels + 1), nullptr);
  int rc = 0;
  hipError_t e = hipSuccess;
  for (auto& x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  for (int w = 0; w < 2 && e == hipSuccess; ++w)
    for (int k = 0; k < kKernels && e == hipSuccess; ++k) e = launch(k);
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    hipEvent_t* er = &ev[(size_t)r * (kKernels + 1)];
    e = hipEventRecord(er[0], s);
    for (int k = 0; k < kKernels && e == hipSuccess; ++k) {
      e = launch(k);
      if (e == hipSuccess) e = hipEventRecord(er[k + 1], s);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  double tot[kKernels] = {0, 0, 0, 0, 0};
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    for (int k = 0; k < kKernels && e == hipSuccess; ++k) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, ev[(size_t)r * (kKernels + 1) + k], ev[(size_t)r * (kKernels + 1) + k + 1]);
      tot[k] += ms;
    }
  if (e != hipSuccess) rc = fail(std::string("mx_time_kernels: ") + hipGetErrorString(e));
  // output order: apply, its finishing kernel, update, its finishing kernel, direction update
  const int order[kKernels] = {1, 2, 3, 4, 0};
  for (int k = 0; k < std::min<int>(n_out, kKernels) && !rc; ++k) us_out[k] = 1e3 * tot[order[k]] / reps;
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  return rc;
}

int64_t mx_solve_plan_reuses(void) { return tl_plan_reuses; }

int mx_solve_plan_release(void) {
  if (!tl_plan) return 0;
  if (tl_plan->busy) return fail("mx_solve_plan_release: the plan is in use by a solve on this thread");
  delete tl_plan->B;
  tl_plan->B = nullptr;
  tl_plan->key.clear();
  tl_plan->relayout = false;
  return 0;
}

int mx_solve(const mx_operator* ops, int32_t nops, const mx_system* systems, int32_t nsys, const mx_options* opt,
             mx_stats* stats, void* stream) {
  if (!ops || nops < 1 || !systems || nsys < 1) return fail("mx_solve: empty batch");
  const mx_options o = defaults(opt);
  std::string why;
  if (!valid_options(o, &why)) return fail(why);
  PlanUse plan;
  if (plan.acquire(ops, nops, systems, nsys, o)) return -1;
  Batch& B = *plan.B;
  B.classic_launch = o.form != MX_FORM_FUSED;
  B.merged_fin = !B.classic_launch && merged_fin_enabled();
  hipStream_t user = reinterpret_cast<hipStream_t>(stream);
  hipStream_t s = B.stream;
  // order after the caller's pending work (inputs are produced on the caller's stream)
  MX_TRY(hipEventRecord(B.ev[0], user));
  MX_TRY(hipStreamWaitEvent(s, B.ev[0], 0));

  // initial state: paused + restart, so the first replacement computes r = b - A x0
  std::vector<SysState> init(nsys);
  int maxiter_all = 0;
  for (int i = 0; i < nsys; ++i) {
    SysState& st = init[i];
    std::memset(&st, 0, sizeof(st));
    // fused systems that lose track finish and continue below in the classic form, so
    // the fused iteration needs no direction-update launch
    st.flags = SF_PAUSE | SF_RESTART | (o.form == MX_FORM_FUSED ? SF_FUSED | SF_SWITCH_EXIT : 0);
    st.rtol = systems[i].rtol;
    st.atol = systems[i].atol;
    st.maxiter = std::max(1, systems[i].maxiter);
    st.delta2 = o.delta * o.delta;
    st.stall_iters = std::max(0, o.stall_iters);
    st.refresh_iters = refresh_iters(o.stall_iters);
    st.stall_limit = std::max(0, o.stall_limit);
    maxiter_all = std::max(maxiter_all, st.maxiter);
  }
  for (int i = 0; i < nsys; ++i) MX_TRY(hipMemcpyAsync(B.h_sys[i].st, &init[i], sizeof(SysState), hipMemcpyHostToDevice, s));
  if (launch_zero_slot(B.mode == MX_FP64, B.d_sys, B.d_bmap_vec, B.nb_vec, VS_XLO, s) != hipSuccess) return fail("zero x_lo");
  MX_TRY(launch_bnorm(B.d_sys, B.d_bmap_vec, B.nb_vec, s));
  MX_TRY(launch_fin(FIN_BNORM, B.d_sys, B.d_syslist, B.nsl, s));
  if (enqueue_replacement_groups(B, s)) return -1;

  const int chunk = o.chunk;
  const int check = std::min(o.check_every, chunk);
  auto capture = [&]() -> int {
    if (B.gexec) { MX_TRY(hipGraphExecDestroy(B.gexec)); B.gexec = nullptr; }
    if (B.graph) { MX_TRY(hipGraphDestroy(B.graph)); B.graph = nullptr; }
    MX_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_chunk(B, s, chunk, check);
    hipError_t ce = hipStreamEndCapture(s, &B.graph);
    if (rc) return -1;
    MX_TRY(ce);
    MX_TRY(hipGraphInstantiate(&B.gexec, B.graph, nullptr, nullptr, 0));
    return 0;
  };
  if (o.use_graph && !B.gexec && capture()) return -1;   // a kept plan has its graph
  auto launch_chunk = [&]() -> int {
    if (o.use_graph) {
      MX_TRY(hipGraphLaunch(B.gexec, s));
      return 0;
    }
    return enqueue_chunk(B, s, chunk, check);
  };
  auto poll = [&](int slot) -> int {
    for (int i = 0; i < nsys; ++i)
      MX_TRY(hipMemcpyAsync(&B.h_state[slot][i], B.h_sys[i].st, sizeof(SysState), hipMemcpyDeviceToHost, s));
    MX_TRY(hipEventRecord(B.ev[slot], s));
    return 0;
  };
  auto all_done = [&](int slot) {
    for (int i = 0; i < nsys; ++i)
      if (!(B.h_state[slot][i].flags & SF_DONE)) return false;
    return true;
  };
  // Finished systems leave the launch grids: once the apply blocks of unfinished systems
  // fall below 90% of those launched, the stream is drained, the maps are rebuilt for
  // the unfinished systems only and the chunk is captured again.  (A finished system's
  // blocks would otherwise still be dispatched by every launch, only to exit.)
  int launched_blocks = B.nb_apply0;
  bool compacted = false;
  auto compact = [&](int slot) -> int {   // returns 1 if rebuilt
    if (nsys == 1) return 0;
    std::vector<char> active(nsys, 0);
    int blocks = 0;
    for (int i = 0; i < nsys; ++i)
      if (!(B.h_state[slot][i].flags & SF_DONE)) { active[i] = 1; blocks += B.h_sys[i].nblk_apply; }
    if (blocks == 0 || blocks > 0.9 * launched_blocks) return 0;
    MX_TRY(hipStreamSynchronize(s));
    // a system may have finished in the chunk that was still queued: take the newest states
    for (int i = 0; i < nsys; ++i) {
      MX_TRY(copy_now(&B.h_state[slot][i], B.h_sys[i].st, sizeof(SysState), hipMemcpyDeviceToHost));
      active[i] = !(B.h_state[slot][i].flags & SF_DONE);
    }
    compacted = true;
    if (layout_groups(B, active)) return -1;
    launched_blocks = B.nb_apply0;
    if (B.nb_apply0 > 0 && o.use_graph && capture()) return -1;
    return 1;
  };
  // pipelined: keep one chunk queued behind the one whose state is being inspected
  // Safety bound only: every chunk advances each unfinished system by at least one
  // iteration or finishes it (a system that pauses for a replacement idles for the rest
  // of that chunk, so a chunk can be worth a single iteration).
  const long long max_chunks = (long long)maxiter_all + 4;
  if (poll(0)) return -1;
  if (launch_chunk() || poll(1)) return -1;
  int cur = 0;
  for (long long c = 0;; ++c) {
    MX_TRY(hipEventSynchronize(B.ev[cur]));
    if (all_done(cur)) break;
    if (c > max_chunks) return fail("mx_solve: iteration limit exceeded without completion (internal error)");
    const int rebuilt = compact(cur);
    if (rebuilt < 0) return -1;
    if (rebuilt) {   // the pipeline restarts from the drained stream
      if (B.nb_apply0 == 0) break;
      if (poll(0) || launch_chunk() || poll(1)) return -1;
      cur = 0;
      continue;
    }
    if (launch_chunk() || poll(cur)) return -1;
    cur ^= 1;
  }
  MX_TRY(hipStreamSynchronize(s));
  // final states (the queued chunk, if any, was a no-op on finished systems)
  std::vector<SysState> fin(nsys);
  for (int i = 0; i < nsys; ++i) MX_TRY(copy_now(&fin[i], B.h_sys[i].st, sizeof(SysState), hipMemcpyDeviceToHost));
  for (int i = 0; i < nsys; ++i)
    if (wants_best(fin[i]))
      MX_TRY(hipMemcpyAsync(B.h_sys[i].vec[VS_X], B.h_sys[i].vec[VS_XBEST], (size_t)B.h_sys[i].n3 * 16,
                            hipMemcpyDeviceToDevice, s));
  MX_TRY(hipStreamSynchronize(s));
  std::vector<mx_stats> st_out(nsys);
  for (int i = 0; i < nsys; ++i) fill_stats(fin[i], &st_out[i]);
  // systems whose fused recursion lost track continue in the classic form from their
  // (best) iterate, with the iterations they have left
  std::vector<int> again;
  for (int i = 0; i < nsys; ++i)
    if (fin[i].flags & SF_SWITCHED) again.push_back(i);
  if (!again.empty()) {
    std::vector<mx_system> sub;
    for (int i : again) {
      mx_system y = systems[i];
      y.maxiter = std::max(1, systems[i].maxiter - fin[i].iters);
      sub.push_back(y);
    }
    mx_options oc = o;
    oc.form = MX_FORM_CLASSIC;
    std::vector<mx_stats> more(sub.size());
    if (mx_solve(ops, nops, sub.data(), (int32_t)sub.size(), &oc, more.data(), s)) return -1;
    for (size_t k = 0; k < again.size(); ++k) {
      mx_stats& t = st_out[again[k]];
      const int it = t.iterations, rep = t.replacements;
      t = more[k];
      t.iterations += it;
      t.replacements += rep;
    }
  }
  if (stats)
    for (int i = 0; i < nsys; ++i) stats[i] = st_out[i];
  MX_TRY(hipEventRecord(B.ev[0], s));
  MX_TRY(hipStreamWaitEvent(user, B.ev[0], 0));
  plan.keep(compacted);
  return 0;
}


// ---- continuous-batching session --------------------------------------------------
// Slots hold admitted systems.  Device tables (operators, system descriptors) have one
// entry per slot; the block maps of every step cover the active slots only and are
// rebuilt when the active set changes.  Iteration chunks are captured as hipGraphs and
// cached per active-set version.
struct mx_session {
  mx_options o;
  int capacity = 0;
  char* d_tab = nullptr;                 // ops32[cap], ops64[cap], sys[cap]
  OpView<float>* d_ops32 = nullptr;
  OpView<double>* d_ops64 = nullptr;
  SysDesc* d_sys = nullptr;
  BlockMap* d_bm = nullptr;              // combined maps: apply blocks, then vector blocks
  size_t bm_cap = 0;
  BlockMap* d_adm = nullptr;             // maps of the slot being admitted
  size_t adm_cap = 0;
  int* d_syslist = nullptr;              // capacity + 1 ints: active slots, then the admitted slot
  int nsl = 0;
  std::vector<SysDesc> h_sys;
  std::vector<mx_operator> h_ops;
  std::vector<char> active;
  std::vector<char> restored;            // best iterate already restored for a finished slot
  std::vector<int> h_ty;                 // apply tile height of each slot
  std::vector<Group> groups;             // launch groups of the active slots (offsets into d_bm)
  int ty = 4;
  int nb_apply = 0, nb_vec = 0;
  bool dirty = true;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  SysState* h_state = nullptr;           // pinned, capacity entries
  BlockMap* h_bm = nullptr;              // pinned staging for maps
  size_t h_bm_cap = 0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_iters = 0;

  ~mx_session() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
    if (h_state) (void)hipHostFree(h_state);
    if (h_bm) (void)hipHostFree(h_bm);
    if (d_bm) (void)hipFree(d_bm);
    if (d_adm) (void)hipFree(d_adm);
    if (d_tab) (void)hipFree(d_tab);
  }
};

namespace {

int ensure_device(BlockMap** d, size_t* cap, size_t need, hipStream_t s) {
  if (need <= *cap) return 0;
  MX_TRY(hipStreamSynchronize(s));
  if (*d) MX_TRY(hipFree(*d));
  *d = nullptr;
  size_t c = std::max<size_t>(need, 2 * *cap);
  MX_TRY(hipMalloc(reinterpret_cast<void**>(d), c * sizeof(BlockMap)));
  *cap = c;
  return 0;
}

int ensure_staging(mx_session* S, size_t need) {
  if (need <= S->h_bm_cap) return 0;
  MX_TRY(hipStreamSynchronize(S->stream));
  if (S->h_bm) MX_TRY(hipHostFree(S->h_bm));
  S->h_bm = nullptr;
  size_t c = std::max<size_t>(need, 2 * S->h_bm_cap);
  MX_TRY(hipHostMalloc(reinterpret_cast<void**>(&S->h_bm), c * sizeof(BlockMap), hipHostMallocDefault));
  S->h_bm_cap = c;
  return 0;
}

// Batch view of a session (shares its device tables) for the enqueue helpers: the
// maps of the admitted slot (one group of tile height ty), or with `groups` the launch
// groups of every active slot.
Batch session_batch(mx_session* S, const BlockMap* bm_apply, int nb_apply, const BlockMap* bm_vec, int nb_vec,
                    const int* syslist, int nsl, int ty, int kind, const std::vector<Group>* groups = nullptr) {
  Batch B;
  B.mode = S->o.mode;
  B.d_ops32 = S->d_ops32;
  B.d_ops64 = S->d_ops64;
  B.d_sys = S->d_sys;
  B.bm_apply0 = const_cast<BlockMap*>(bm_apply);
  B.bm_vec0 = const_cast<BlockMap*>(bm_vec);
  B.syslist0 = const_cast<int*>(syslist);
  B.nb_apply0 = nb_apply;
  B.nb_vec0 = nb_vec;
  B.nsl0 = nsl;
  if (groups) B.groups = *groups;
  B.use_all();
  B.ty = ty;
  B.kind = kind;
  return B;
}

void release_view(Batch& B) {   // the view owns nothing
  B.d_mem = nullptr;
  B.d_syslist = nullptr;
  B.syslist0 = nullptr;
  B.stream = nullptr;
  B.ev[0] = B.ev[1] = nullptr;
  B.h_state[0] = B.h_state[1] = nullptr;
}

// Equal time per launch group in a session step.  Groups run one after the other, so
// with the same iteration count a group of small systems would advance at the pace of
// the costliest group, and a small request's latency would grow with what it shares the
// session with.  A group repeats its chunk about cost_max / cost times (at most
// kMaxGroupReps), with the cost of one of its iterations modelled as its cells plus a
// fixed launch cost of kLaunchCells (a fit of single-system FP64 iteration times: about
// 28 us at 48^3 and 140 us at 128^3 on MI355X).  One group: no repetition.
constexpr int kMaxGroupReps = 16;
constexpr double kLaunchCells = 4.0e5;
void session_group_reps(mx_session* S, const std::vector<int>& sl) {
  std::vector<double> cost(S->groups.size(), kLaunchCells);
  double cmax = 0.0;
  for (size_t g = 0; g < S->groups.size(); ++g) {
    const Group& G = S->groups[g];
    for (int i = 0; i < G.ns; ++i) cost[g] += (double)(S->h_sys[sl[G.s_off + i]].n3 / 3);
    cmax = std::max(cmax, cost[g]);
  }
  for (size_t g = 0; g < S->groups.size(); ++g)
    S->groups[g].reps = std::max(1, std::min(kMaxGroupReps, (int)(cmax / cost[g])));
}

int session_rebuild(mx_session* S) {
  std::vector<Member> members;
  for (int i = 0; i < S->capacity; ++i)
    if (S->active[i])
      members.push_back({i, &S->h_ops[i], i, S->h_ty[i], op_kind(S->h_ops[i]), S->h_sys[i].nblk_apply,
                         S->h_sys[i].nblk_vec});
  std::vector<BlockMap> a, v;
  std::vector<int> sl;
  make_groups(members, S->o.mode, a, v, sl, S->groups);
  session_group_reps(S, sl);
  S->nsl = (int)sl.size();
  S->nb_apply = (int)a.size();
  S->nb_vec = (int)v.size();
  const size_t need = a.size() + v.size();
  if (ensure_device(&S->d_bm, &S->bm_cap, std::max<size_t>(need, 1), S->stream)) return -1;
  if (ensure_staging(S, std::max<size_t>(need, 1))) return -1;
  MX_TRY(hipStreamSynchronize(S->stream));   // staging buffer reuse
  std::copy(a.begin(), a.end(), S->h_bm);
  std::copy(v.begin(), v.end(), S->h_bm + a.size());
  if (need) MX_TRY(hipMemcpyAsync(S->d_bm, S->h_bm, need * sizeof(BlockMap), hipMemcpyHostToDevice, S->stream));
  if (!sl.empty()) MX_TRY(copy_now(S->d_syslist, sl.data(), sizeof(int) * sl.size(), hipMemcpyHostToDevice));
  if (S->gexec) { MX_TRY(hipGraphExecDestroy(S->gexec)); S->gexec = nullptr; }
  if (S->graph) { MX_TRY(hipGraphDestroy(S->graph)); S->graph = nullptr; }
  S->dirty = false;
  return 0;
}

}  // namespace

mx_session* mx_session_create(int32_t capacity, const mx_options* opt) {
  if (capacity < 1) { fail("capacity must be >= 1"); return nullptr; }
  mx_session* S = new (std::nothrow) mx_session();
  if (!S) { fail("out of host memory"); return nullptr; }
  S->o = defaults(opt);
  {
    std::string why;
    if (!valid_options(S->o, &why)) { fail(why); delete S; return nullptr; }
  }
  S->capacity = capacity;
  S->ty = S->o.tile_rows ? S->o.tile_rows : kAutoTileRows;
  S->h_sys.assign(capacity, SysDesc{});
  S->h_ops.assign(capacity, mx_operator{});
  S->active.assign(capacity, 0);
  S->restored.assign(capacity, 0);
  S->h_ty.assign(capacity, S->ty);
  const size_t s32 = align_up(sizeof(OpView<float>) * capacity);
  const size_t s64 = align_up(sizeof(OpView<double>) * capacity);
  const size_t ss = align_up(sizeof(SysDesc) * capacity);
  const size_t ssl = align_up(sizeof(int) * (capacity + 1));
  hipError_t e = hipMalloc(&S->d_tab, s32 + s64 + ss + ssl);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&S->ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&S->h_state), sizeof(SysState) * capacity, hipHostMallocDefault);
  if (e != hipSuccess) { fail(std::string("mx_session_create: ") + hipGetErrorString(e)); delete S; return nullptr; }
  S->d_ops32 = reinterpret_cast<OpView<float>*>(S->d_tab);
  S->d_ops64 = reinterpret_cast<OpView<double>*>(S->d_tab + s32);
  S->d_sys = reinterpret_cast<SysDesc*>(S->d_tab + s32 + s64);
  S->d_syslist = reinterpret_cast<int*>(S->d_tab + s32 + s64 + ss);
  return S;
}

int mx_session_admit(mx_session* S, const mx_operator* op, const mx_system* y, void* stream) {
  if (!S || !op || !y) return fail("null argument");
  std::string why;
  if (!valid_op(*op, S->o.mode, &why)) return fail(why);
  if (!y->b || !y->x || !y->work) return fail("b, x and work are required");
  int slot = -1;
  for (int i = 0; i < S->capacity; ++i)
    if (!S->active[i]) { slot = i; break; }
  if (slot < 0) return fail("session is full");
  hipStream_t s = S->stream;
  // caller's inputs (b, x) are produced on the caller's stream
  MX_TRY(hipEventRecord(S->ev, reinterpret_cast<hipStream_t>(stream)));
  MX_TRY(hipStreamWaitEvent(s, S->ev, 0));

  const WorkLayout w = work_layout(*op, S->o.mode);
  char* wk = reinterpret_cast<char*>(y->work);
  SysDesc d;
  std::memset(&d, 0, sizeof(d));
  d.op = slot;   // one operator per slot
  const int ty = y->tile_rows ? y->tile_rows : S->ty;
  if (ty != 2 && ty != 4 && ty != 8) return fail("tile_rows must be 2, 4 or 8");
  d.slab = y->slab > 0 ? std::min<int>(y->slab, op->nx)
         : S->o.slab > 0 ? std::min<int>(S->o.slab, op->nx) : auto_slab(*op, ty, true);
  if (d.slab < 1) return fail(too_many_tiles(*op, ty));
  d.nblk_apply = apply_blocks(*op, ty, d.slab);
  d.n3 = 3 * ncell_of(*op);
  d.nblk_vec = vec_blocks(d.n3);
  if (d.nblk_apply > kMaxPartials) return fail("too many apply blocks for one system; use a deeper slab");
  d.vec[VS_X] = y->x;
  d.vec[VS_B] = const_cast<void*>(y->b);
  d.vec[VS_R] = wk + w.off_r;
  d.vec[VS_P] = wk + w.off_p;
  d.vec[VS_Q] = wk + w.off_q;
  d.vec[VS_XLO] = wk + w.off_xlo;
  d.vec[VS_XBEST] = wk + w.off_xbest;
  d.st = reinterpret_cast<SysState*>(wk + w.off_state);
  d.partial = reinterpret_cast<double*>(wk + w.off_partial);
  d.partial_vec = reinterpret_cast<double*>(wk + w.off_partial_vec);
  S->h_sys[slot] = d;
  S->h_ops[slot] = *op;
  S->h_ty[slot] = ty;

  // tables and initial state, staged through pinned memory owned by the session
  MX_TRY(hipStreamSynchronize(s));
  SysState* st = &S->h_state[slot];
  std::memset(st, 0, sizeof(*st));
  // fused systems that lose track finish (MX_FLAG_SWITCHED) for the host to continue in
  // the classic form, so every slot stays fused and the three-launch fused iteration applies
  st->flags = SF_PAUSE | SF_RESTART | (S->o.form == MX_FORM_FUSED ? SF_FUSED | SF_SWITCH_EXIT : 0);
  st->rtol = y->rtol;
  st->atol = y->atol;
  st->maxiter = std::max(1, y->maxiter);
  st->delta2 = S->o.delta * S->o.delta;
  st->stall_iters = std::max(0, S->o.stall_iters);
  st->refresh_iters = refresh_iters(S->o.stall_iters);
  st->stall_limit = std::max(0, S->o.stall_limit);
  OpView<double> v64 = view_of<double>(*op);
  OpView<float> v32;
  if (S->o.mode == MX_MIXED) v32 = view_of<float>(*op);
  else std::memset(&v32, 0, sizeof(v32));
  MX_TRY(copy_now(S->d_ops32 + slot, &v32, sizeof(v32), hipMemcpyHostToDevice));
  MX_TRY(copy_now(S->d_ops64 + slot, &v64, sizeof(v64), hipMemcpyHostToDevice));
  MX_TRY(copy_now(S->d_sys + slot, &d, sizeof(d), hipMemcpyHostToDevice));
  MX_TRY(copy_now(d.st, st, sizeof(*st), hipMemcpyHostToDevice));

  // admission kernels on this slot only: x_lo = 0, ||b||, r = b - A x0
  std::vector<BlockMap> a, v;
  for (int b = 0; b < d.nblk_apply; ++b) a.push_back({slot, b});
  for (int b = 0; b < d.nblk_vec; ++b) v.push_back({slot, b});
  a = xcd_order(a);
  v = xcd_order(v);
  const size_t need = a.size() + v.size();
  if (ensure_device(&S->d_adm, &S->adm_cap, need, s)) return -1;
  if (ensure_staging(S, need)) return -1;
  std::copy(a.begin(), a.end(), S->h_bm);
  std::copy(v.begin(), v.end(), S->h_bm + a.size());
  MX_TRY(copy_now(S->d_adm, S->h_bm, need * sizeof(BlockMap), hipMemcpyHostToDevice));
  int* adm_sl = S->d_syslist + S->capacity;
  MX_TRY(copy_now(adm_sl, &slot, sizeof(int), hipMemcpyHostToDevice));
  Batch B = session_batch(S, S->d_adm, (int)a.size(), S->d_adm + a.size(), (int)v.size(), adm_sl, 1, ty, op_kind(*op));
  int rc = 0;
  if (launch_zero_slot(S->o.mode == MX_FP64, B.d_sys, B.d_bmap_vec, B.nb_vec, VS_XLO, s) != hipSuccess) rc = fail("zero x_lo");
  if (!rc && launch_bnorm(B.d_sys, B.d_bmap_vec, B.nb_vec, s) != hipSuccess) rc = fail("bnorm");
  if (!rc && launch_fin(FIN_BNORM, B.d_sys, B.d_syslist, B.nsl, s) != hipSuccess) rc = fail("bnorm finish");
  if (!rc) rc = enqueue_replacement(B, s);
  release_view(B);
  if (rc) return rc;
  S->active[slot] = 1;
  S->restored[slot] = 0;
  S->dirty = true;
  return slot;
}

int mx_session_step(mx_session* S, int32_t iterations, void* stream) {
  if (!S) return fail("null session");
  (void)stream;
  if (S->dirty && session_rebuild(S)) return -1;
  if (S->nb_apply == 0 || iterations <= 0) return 0;
  Batch B = session_batch(S, S->d_bm, S->nb_apply, S->d_bm + S->nb_apply, S->nb_vec, S->d_syslist, S->nsl, S->ty,
                          MK_GENERAL, &S->groups);   // each launch group carries its own kind
  B.classic_launch = S->o.form != MX_FORM_FUSED;
  B.merged_fin = !B.classic_launch && merged_fin_enabled();
  const int check = std::min<int>(S->o.check_every, iterations);
  int rc = 0;
  if (S->o.use_graph) {
    if (!S->gexec || S->graph_iters != iterations) {
      if (S->gexec) { (void)hipGraphExecDestroy(S->gexec); S->gexec = nullptr; }
      if (S->graph) { (void)hipGraphDestroy(S->graph); S->graph = nullptr; }
      hipError_t e = hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) rc = fail(std::string("capture: ") + hipGetErrorString(e));
      if (!rc) {
        rc = enqueue_chunk(B, S->stream, iterations, check);
        e = hipStreamEndCapture(S->stream, &S->graph);
        if (!rc && e != hipSuccess) rc = fail(std::string("end capture: ") + hipGetErrorString(e));
      }
      if (!rc) {
        e = hipGraphInstantiate(&S->gexec, S->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(std::string("instantiate: ") + hipGetErrorString(e));
        S->graph_iters = iterations;
      }
    }
    if (!rc) {
      hipError_t e = hipGraphLaunch(S->gexec, S->stream);
      if (e != hipSuccess) rc = fail(std::string("graph launch: ") + hipGetErrorString(e));
    }
  } else {
    rc = enqueue_chunk(B, S->stream, iterations, check);
  }
  release_view(B);
  return rc;
}

int mx_session_poll(mx_session* S, mx_stats* stats, int32_t* done, void* stream) {
  if (!S) return fail("null session");
  for (int i = 0; i < S->capacity; ++i)
    if (S->active[i]) MX_TRY(hipMemcpyAsync(&S->h_state[i], S->h_sys[i].st, sizeof(SysState), hipMemcpyDeviceToHost, S->stream));
  MX_TRY(hipStreamSynchronize(S->stream));
  for (int i = 0; i < S->capacity; ++i)
    if (S->active[i] && !S->restored[i] && (S->h_state[i].flags & SF_DONE)) {
      if (wants_best(S->h_state[i]))
        MX_TRY(hipMemcpyAsync(S->h_sys[i].vec[VS_X], S->h_sys[i].vec[VS_XBEST], (size_t)S->h_sys[i].n3 * 16,
                              hipMemcpyDeviceToDevice, S->stream));
      S->restored[i] = 1;
    }
  for (int i = 0; i < S->capacity; ++i) {
    if (stats) {
      if (S->active[i]) fill_stats(S->h_state[i], &stats[i]);
      else std::memset(&stats[i], 0, sizeof(mx_stats));
    }
    if (done) done[i] = S->active[i] ? ((S->h_state[i].flags & SF_DONE) ? 1 : 0) : -1;
  }
  // results become visible to work the caller queues afterwards on its stream
  MX_TRY(hipEventRecord(S->ev, S->stream));
  MX_TRY(hipStreamWaitEvent(reinterpret_cast<hipStream_t>(stream), S->ev, 0));
  return 0;
}

int mx_session_retire(mx_session* S, int32_t slot) {
  if (!S || slot < 0 || slot >= S->capacity || !S->active[slot]) return fail("bad slot");
  MX_TRY(hipStreamSynchronize(S->stream));
  S->active[slot] = 0;
  S->dirty = true;
  return 0;
}

void mx_session_destroy(mx_session* S) {
  if (!S) return;
  (void)hipStreamSynchronize(S->stream);
  delete S;
}

}  // extern "C"
