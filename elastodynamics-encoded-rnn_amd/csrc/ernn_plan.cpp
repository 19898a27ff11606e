// Host side of libernn_hip.so: descriptor validation, workspace layout, tile
// and event tables, the per-pass launch schedule (forward, checkpointed
// recompute, reverse-time adjoint), optional hipGraph capture, and the C ABI
// declared in include/ernn.h.
//
// Schedule (K = checkpoint interval, S = ceil(nt_max / K) checkpoints).  A
// time level of either recurrence is a state (u, c, v) in a state slot
// (ernn_internal.h); every launch reads one level and writes its last level to
// a state slot other than the one it reads.
//  forward   steps n = 0 .. nt-1 take level n to level n + 1.  Levels sK (s <
//            S) are checkpoints and go to their own state slots; checkpoint 0
//            is the zero initial state of PI:256-259 (u^0 = u^-1 = 0).  Other
//            levels alternate between two ring slots; the time-blocked forward
//            runs KS steps per launch and writes only its last level and the
//            checkpoints inside it.  Levels of the last checkpoint segment also
//            go to the bank (u only), as its recompute would leave them.
//  adjoint   for s = S-1 .. 0: recompute levels sK+1 .. top (top = min((s+1)K,
//            nt)) from checkpoint s, writing every u^m to the segment's bank,
//            then run the adjoint steps m = top .. sK+1 (lam level m + 1 ->
//            m), which read u^m from the bank; lam's levels alternate between
//            two slots.  The forward has no atomics and a fixed order, so the
//            recomputed states equal the forward's bit for bit.
//  fused     small grids: one launch runs adjoint work of segment s and
//            recompute work of segment s - 1 (two banks).
//  fixed_c0  (damage field over known constants, kernel mode kDOnly) the bank
//            slot of level m holds L^m = (dt^2/rho) A(C0) u^m instead of u^m:
//            step n banks L^n of the level it leaves, so a recompute of
//            segment s runs steps sK .. top (one past the segment, for
//            L^top; the last segment stops at nt - 1, and the adjoint skips
//            the term of level nt, whose lam^{nt+1} is zero).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ernn.h"
#include "ernn_internal.h"

namespace ernn {
hipError_t launch_fwd_step(const StepArgs& a, int V, int blocks, bool loop, hipStream_t st);
hipError_t launch_adj_step(const StepArgs& a, int V, int blocks, bool loop, hipStream_t st);
hipError_t launch_fused_step(const FusedArgs& fa, int V, int blocks, hipStream_t st);
hipError_t launch_fwd_block(const BlockArgs& a, int V, int grid, hipStream_t st);
hipError_t launch_adj_block(const BlockAdjArgs& a, int V, int grid, hipStream_t st);
hipError_t launch_rev_block(const BlockAdjArgs& a, int V, int grid, hipStream_t st);
hipError_t launch_fused_block(const FusedBlockArgs& fa, int V, int grid, hipStream_t st);
hipError_t launch_set_call(CallDev* dst, const CallDev& v, hipStream_t st);
hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t st);
hipError_t launch_fill_ws(const CallDev* call, int64_t off, int64_t n, hipStream_t st);
hipError_t launch_pack_fields(const CallDev* call, const InstDev* inst, const int* set_inst,
                              const int64_t* set_ext_off, int n_coef, int set, int64_t ncell,
                              hipStream_t st);
hipError_t launch_finalize_const(const CallDev* call, int64_t partial_off, const int* set_inst_list,
                                 const int* set_inst_begin, const int* rec_first, int per_rec, int n_sets,
                                 hipStream_t st);
hipError_t launch_finalize_field(const CallDev* call, const InstDev* inst, const int* set_inst_list,
                                 const int* set_inst_begin, const int64_t* set_grad_off, int n_acc,
                                 int set, int64_t ncell, hipStream_t st);
}  // namespace ernn

using namespace ernn;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define ERNN_HIP(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      return fail(ERNN_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define ERNN_TRY(x)             \
  do {                          \
    int r_ = (x);               \
    if (r_ != ERNN_OK) return r_; \
  } while (0)

// Tile shapes with a compiled kernel (ERNN_TILE_CASES in ernn_kernels.hip):
// V in {1, 2, 4} columns per lane, R a power of two rows up to 8, V * R <= 16.
bool tile_ok(int V, int R) {
  if (!(V == 1 || V == 2 || V == 4)) return false;
  if (R < 1 || R > 8 || (R & (R - 1)) != 0) return false;
  return V * R <= 16;
}
// Adjoint shapes that would spill registers are not built (ERNN_TILE_CASES_*
// in ernn_kernels.hip): the damage-field adjoint needs V * R <= 8 and
// R <= 4, the six-field adjoint V * R <= 4 and R <= 2.
bool adj_tile_ok(int mode, int V, int R) {
  if (!tile_ok(V, R)) return false;
  if (mode == ERNN_PARAM_DFIELD) return V * R <= 8 && R <= 4;
  if (mode == ERNN_PARAM_FIELD6) return V * R <= 4 && R <= 2;
  return true;
}
// The six-field forward is built for V * R <= 8 (larger shapes would spill).
bool fwd_tile_ok(int mode, int V, int R) {
  if (!tile_ok(V, R)) return false;
  return !(mode == ERNN_PARAM_FIELD6 && V * R > 8);
}

// Time-blocked forward shapes (ERNN_BLOCK_CASES in ernn_kernels.hip): V in
// {1, 2}, RW in {4, 8} rows per wavefront, KS in {2, 4, 8} steps with an
// output tile of 4 RW - 2 KS > 0 rows, and (1, 8, 5); (2, 8) would spill for
// field modes.
bool block_ok(int mode, int V, int RW, int KS) {
  if (V == 1 && RW == 8 && KS == 5) return true;
  if (!(V == 1 || V == 2) || !(RW == 4 || RW == 8) || !(KS == 2 || KS == 4 || KS == 8)) return false;
  if (blk_waves(RW) * RW - 2 * KS <= 0) return false;
  return !(mode != ERNN_PARAM_CONST6 && V == 2 && RW == 8);
}
// Time-blocked adjoint shapes (ERNN_ADJ_BLOCK_CASES): constant and damage
// modes: the forward's V = 1 shapes and (1, 6, 4); V = 2 (RW = 4) for
// constant stiffness only.  Six fields (ERNN_ADJ_BLOCK_CASES_FIELD6): V = 1,
// RW in {2, 4} (16-row regions), KS in {2, 4}.
bool adj_block_ok(int mode, int V, int RW, int KS) {
  if (mode == ERNN_PARAM_FIELD6) return V == 1 && (RW == 2 || RW == 4) && (KS == 2 || KS == 4);
  if (RW == 6) return V == 1 && KS == 4;
  if (KS == 5) return mode == ERNN_PARAM_CONST6 && V == 1 && RW == 8;
  if (!block_ok(mode, V, RW, KS)) return false;
  if (V == 2) return mode == ERNN_PARAM_CONST6 && RW == 4;
  return true;
}
// Fused time-blocked backward shapes (ERNN_FUSED_BLOCK_CASES): on a fuse
// plan the adjoint blocking shape runs both halves of each fused launch.
bool fused_block_ok(int mode, int V, int RW, int KS) {
  if (mode == ERNN_PARAM_FIELD6 || V != 1) return false;
  return (RW == 4 && (KS == 2 || KS == 4)) || (mode == ERNN_PARAM_CONST6 && RW == 8 && KS == 4);
}
// Rows per wavefront of the fused kernel's recompute half (ERNN_FUSED_BLOCK_CASES):
// the adjoint's, or 8 beside the 4-row adjoint shapes.
bool fused_ff_ok(int V, int RW, int KS, int RWF) {
  return RWF == RW || (V == 1 && RW == 4 && RWF == 8 && (KS == 2 || KS == 4));
}
// Reversed-u adjoint shapes (ERNN_REV_BLOCK_CASES): V = 1, RW in {4, 6, 8}, KS in {2, 4}, and
// RW in {8, 10}, KS in {5, 6}; the joint gradient leaves out RW = 10.
bool rev_block_ok(int V, int RW, int KS, bool joint = false) {
  if (joint && RW == 10) return false;   // ERNN_REV_BLOCK_CASES_JOINT
  return V == 1 && (((RW == 4 || RW == 6 || RW == 8) && (KS == 2 || KS == 4)) ||
                    ((RW == 8 || RW == 10) && (KS == 5 || KS == 6)));
}
bool ablk_ok(int mode, bool fuse, int V, int RW, int KS) {
  return fuse ? fused_block_ok(mode, V, RW, KS) : adj_block_ok(mode, V, RW, KS);
}
const char* const kAblkShapes =
    "time-blocked adjoint shape (V, RW, KS): unfused: V = 1 with RW in {4,8}, KS in {2,4,8}, 4 RW > 2 KS, or "
    "(1, 6, 4), or (constant) (1, 8, 5) or V = 2, RW = 4, KS in {2,4}; six fields: V = 1, RW in {2,4}, KS in "
    "{2,4}; fused (constant or damage mode): (1, 4, 2), (1, 4, 4) or (constant) (1, 8, 4)";

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct Tiling {
  int V = 0, R = 0, bpc = 0;
  int n_tiles = 0, blocks = 0;
  bool loop = false;      // grid capped below the tile count: wavefronts loop over tiles
  int fused_blocks = 0;   // grid of the fused kernel (2 x tiles)
  TileDev* d_tiles = nullptr;
  SrcEv* d_src = nullptr;
  RecEv* d_rec = nullptr;
  int* d_inst_first = nullptr;   // [n_inst + 1]: first tile of each instance
};

struct BlockTiling {
  int V = 0, RW = 0, KS = 0;   // KS = 0: the forward runs one step per launch
  int n_blocks = 0, grid = 0;
  BlockDev* d_blocks = nullptr;
  SrcEv* d_src = nullptr;
  RecEv* d_rec = nullptr;
  int* d_inst_first = nullptr;   // [n_inst + 1]: first block of each instance
};

struct Plan {
  int device = 0;
  int n_cu = 0;
  int mode = 0, n_inst = 0, n_sets = 0, P = 1, snapshot = 0, K = 2, S = 1, nt_max = 0;
  int n_coef = 0, n_acc = 0;
  bool use_graph = false;
  bool fuse = false;
  int ff_rw = 0;   // rows per wavefront of the fused recompute half (0: no fused time-blocked backward)
  // fixed_c0 (damage field over known constants): the bank holds the operator
  // value L^m = (dt^2/rho) A(C0) u^m of level m instead of u^m, written by the
  // step that leaves level m (a recompute runs one step past its segment for
  // L^top), and the kernels run in mode kDOnly
  bool lbank = false;
  int kmode = 0;   // kernel mode (ParamMode)
  // reverse_u (with fixed_c0): no bank; the adjoint runs u backward from the
  // checkpoints (rev_block_kernel).  Slots: checkpoint S holds the last level,
  // and each instance's own last level goes to checkpoint ceil(nt_b / K);
  // two more ring slots carry the reversed u between launches.
  bool rtr = false;
  // ortho (with fixed_c0): the caller promises C16 = C26 = 0 in every cell, so
  // the damage kernels' operator sums drop those terms (the same bits)
  bool ortho = false;
  std::vector<ernn_instance_t> inst_h;
  std::vector<ernn_source_t> src_h;
  std::vector<ernn_receiver_t> rec_h;
  std::vector<InstDev> inst;
  InstDev* d_inst = nullptr;
  CallDev* d_call = nullptr;
  Tiling fwd, adj;
  // blk: time-blocked forward; blk_adj: time-blocked adjoint; blk_ff: the
  // forward-style table (sources over the region) with blk_adj's shape, for
  // the recompute half of the fused time-blocked backward
  BlockTiling blk, blk_adj, blk_ff;
  int64_t ws_floats = 0;
  int64_t partial_off = 0;  // in doubles
  int64_t partial_cap = 0;  // tiles
  std::vector<int64_t> trace_off;
  std::vector<int64_t> set_ext_off, set_grad_off;
  std::vector<int> set_first, set_list, set_begin;
  int* d_set_first = nullptr;
  int* d_set_list = nullptr;
  int* d_set_begin = nullptr;
  int64_t* d_set_ext_off = nullptr;
  int64_t* d_set_grad_off = nullptr;
  int64_t field_floats = 0, grad_field_doubles = 0;
  // g_adj_skip: the adjoint that starts on the last segment's stored states
  hipGraphExec_t g_fwd = nullptr, g_adj = nullptr, g_adj_skip = nullptr;
  // The workspace whose last forward pass left every state of the last
  // checkpoint segment in that segment's bank, not yet used up by an adjoint
  // (its recompute of earlier segments reuses the bank), or null.
  const void* seg_ws = nullptr;
  hipStream_t cap_stream = nullptr;

  // An instance's state slots (kStateFields fields each): two forward ring
  // slots, the checkpoints ck(s) = level sK, two lam ring slots; its bank
  // slots (u only, or L with lbank): banks x K, bank(b, i) holding u^{sK+1+i}
  // of a segment s with bank_of(s) = b.
  int banks() const { return fuse ? 2 : 1; }
  int ring(int r) const { return r; }
  int ck(int s) const { return 2 + s; }   // s <= S with rtr (checkpoint S: the last level)
  int lam(int r) const { return 2 + S + (rtr ? 1 : 0) + r; }
  int ur(int r) const { return 4 + S + 1 + r; }   // rtr only
  int n_state() const { return 4 + S + (rtr ? 3 : 0); }
  int n_bank() const { return rtr ? 0 : banks() * K; }
  int bank_of(int s) const { return fuse ? (s & 1) : 0; }
  int bank(int s, int m) const { return bank_of(s) * K + (m - s * K - 1); }   // u^m, sK < m <= top
  bool is_ck(int m) const { return m % K == 0 && m / K < S; }
  int other_ring(int q) const { return q == ring(0) ? ring(1) : ring(0); }
  int other_lam(int q) const { return q == lam(0) ? lam(1) : lam(0); }
  int other_ur(int q) const { return q == ur(0) ? ur(1) : ur(0); }
  int seg_top(int s) const { return std::min((s + 1) * K, nt_max); }
};

template <typename T>
int dev_upload(T** dst, const std::vector<T>& v) {
  *dst = nullptr;
  if (v.empty()) return ERNN_OK;
  ERNN_HIP(hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * v.size()));
  ERNN_HIP(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return ERNN_OK;
}

void free_tiling(Tiling& t) {
  if (t.d_tiles) (void)hipFree(t.d_tiles);
  if (t.d_src) (void)hipFree(t.d_src);
  if (t.d_rec) (void)hipFree(t.d_rec);
  if (t.d_inst_first) (void)hipFree(t.d_inst_first);
  t = Tiling{};
}

void free_blocking(BlockTiling& b) {
  if (b.d_blocks) (void)hipFree(b.d_blocks);
  if (b.d_src) (void)hipFree(b.d_src);
  if (b.d_rec) (void)hipFree(b.d_rec);
  if (b.d_inst_first) (void)hipFree(b.d_inst_first);
  b = BlockTiling{};
}

void drop_graphs(Plan* p) {
  if (p->g_fwd) (void)hipGraphExecDestroy(p->g_fwd);
  if (p->g_adj) (void)hipGraphExecDestroy(p->g_adj);
  if (p->g_adj_skip) (void)hipGraphExecDestroy(p->g_adj_skip);
  p->g_fwd = p->g_adj = p->g_adj_skip = nullptr;
}

int64_t tiles_for(const Plan& p, int V, int R) {
  int64_t n = 0;
  for (const auto& I : p.inst_h)
    n += ((I.nx + R - 1) / R) * (int64_t)((I.ny + kWave * V - 1) / (kWave * V));
  return n;
}

int build_tiling(Plan* p, Tiling* t, int V, int R, int bpc) {
  if (!tile_ok(V, R))
    return fail(ERNN_E_INVALID, "tile (V, R) must have V in {1,2,4}, R a power of two up to 8, V*R <= 16");
  if (t == &p->fwd && !fwd_tile_ok(p->mode, V, R))
    return fail(ERNN_E_INVALID, "forward tile (V, R) with V*R > 8 is not built for six fields");
  if (t == &p->adj && !adj_tile_ok(p->mode, V, R))
    return fail(ERNN_E_INVALID, "adjoint tile (V, R) not built for this mode (damage field: V*R <= 8 "
                                "and R <= 4, six fields: V*R <= 4 and R <= 2)");
  if (bpc < 0 || bpc > 64) return fail(ERNN_E_INVALID, "blocks per CU must be in [0, 64]");
  const int64_t nt64 = tiles_for(*p, V, R);
  if (nt64 > (int64_t)1 << 30) return fail(ERNN_E_CAPACITY, "too many tiles");
  std::vector<TileDev> tiles;
  tiles.reserve((size_t)nt64);
  std::vector<SrcEv> sev;
  std::vector<RecEv> rev;
  std::vector<int> first(p->n_inst + 1, 0);
  const int W = kWave * V;
  for (int b = 0; b < p->n_inst; ++b) {
    const ernn_instance_t& I = p->inst_h[b];
    first[b] = (int)tiles.size();
    const int nti = (I.nx + R - 1) / R, ntj = (I.ny + W - 1) / W;
    // bucket this instance's sources and receivers by local tile, sorted by (i, j)
    // key: (local tile, row, column); value: index into the source / receiver slice
    struct Key { int64_t tile; int i, j, idx; };
    const auto by_key = [](const Key& a, const Key& b) {
      if (a.tile != b.tile) return a.tile < b.tile;
      if (a.i != b.i) return a.i < b.i;
      return a.j < b.j;
    };
    std::vector<Key> sk, rk;
    for (int e = 0; e < I.src_count; ++e) {
      const ernn_source_t& s = p->src_h[I.src_begin + e];
      sk.push_back(Key{(int64_t)(s.i / R) * ntj + s.j / W, s.i, s.j, I.src_begin + e});
    }
    if (!p->snapshot) {
      for (int e = 0; e < I.rec_count; ++e) {
        const ernn_receiver_t& r = p->rec_h[I.rec_begin + e];
        rk.push_back(Key{(int64_t)(r.i / R) * ntj + r.j / W, r.i, r.j, e});
      }
    }
    std::stable_sort(sk.begin(), sk.end(), by_key);
    std::stable_sort(rk.begin(), rk.end(), by_key);
    size_t si = 0, ri = 0;
    for (int ti = 0; ti < nti; ++ti) {
      for (int tj = 0; tj < ntj; ++tj) {
        const int64_t tl = (int64_t)ti * ntj + tj;
        TileDev td{};
        td.inst = b;
        td.i0 = ti * R;
        td.j0 = tj * W;
        td.set = I.set;
        td.src_begin = (int)sev.size();
        while (si < sk.size() && sk[si].tile == tl) {
          const ernn_source_t& s = p->src_h[sk[si].idx];
          sev.push_back(SrcEv{s.i, s.j, s.w, s.comp});
          ++si;
        }
        td.src_end = (int)sev.size();
        td.rec_begin = (int)rev.size();
        while (ri < rk.size() && rk[ri].tile == tl) {
          const ernn_receiver_t& r = p->rec_h[I.rec_begin + rk[ri].idx];
          rev.push_back(RecEv{r.i, r.j, rk[ri].idx, 0});
          ++ri;
        }
        td.rec_end = (int)rev.size();
        tiles.push_back(td);
      }
    }
  }
  first[p->n_inst] = (int)tiles.size();
  // The new tables are uploaded into a temporary and swapped in only when
  // every upload succeeded: a failure leaves the plan's tiling as it was.
  Tiling nt;
  nt.V = V;
  nt.R = R;
  nt.bpc = bpc;
  nt.n_tiles = (int)tiles.size();
  // 8 XCDs, each owning ceil(n_tiles / 8) tiles; a wavefront per tile unless capped
  const int64_t chunk = (nt.n_tiles + 7) / 8;
  int64_t per_xcd = std::max<int64_t>((chunk + kWavesPerBlock - 1) / kWavesPerBlock, 1);
  const int64_t cap = std::max<int64_t>((int64_t)bpc * p->n_cu / 8, 1);
  nt.loop = bpc > 0 && cap < per_xcd;
  if (nt.loop) per_xcd = cap;
  nt.blocks = (int)(8 * per_xcd);
  const int64_t fchunk = (2 * (int64_t)nt.n_tiles + 7) / 8;
  int64_t fper = std::max<int64_t>((fchunk + kWavesPerBlock - 1) / kWavesPerBlock, 1);
  if (bpc > 0) fper = std::min<int64_t>(fper, cap);
  nt.fused_blocks = (int)(8 * fper);
  int r = dev_upload(&nt.d_tiles, tiles);
  if (r == ERNN_OK) r = dev_upload(&nt.d_src, sev);
  if (r == ERNN_OK) r = dev_upload(&nt.d_rec, rev);
  if (r == ERNN_OK) r = dev_upload(&nt.d_inst_first, first);
  if (r != ERNN_OK) {
    free_tiling(nt);
    return r;
  }
  free_tiling(*t);
  *t = nt;
  return ERNN_OK;
}

// Block table of a time-blocked shape, built into *out (empty on entry; left
// empty on failure).  Output tiles of (blk_waves(RW) RW - 2 KS) x (64 V - 2 KS)
// cells per instance, row-major; each block's source list covers its whole
// region (the halo cells' states feed the later steps).  Receivers: a
// forward-style table (adjoint = false) lists those inside the output tile
// (recording), an adjoint table those inside the whole region (injection).
int make_blocking(const Plan* p, bool adjoint, int V, int RW, int KS, BlockTiling* out) {
  const int OR = blk_waves(RW) * RW - 2 * KS, OC = kWave * V - 2 * KS;
  std::vector<BlockDev> blocks;
  std::vector<SrcEv> sev;
  std::vector<RecEv> rev;
  struct Ev { int64_t blk; int i, j, idx; };
  const auto by_key = [](const Ev& a, const Ev& b) {
    if (a.blk != b.blk) return a.blk < b.blk;
    if (a.i != b.i) return a.i < b.i;
    return a.j < b.j;
  };
  std::vector<int> first(p->n_inst + 1, 0);
  for (int b = 0; b < p->n_inst; ++b) {
    const ernn_instance_t& I = p->inst_h[b];
    first[b] = (int)blocks.size();
    const int nbi = (I.nx + OR - 1) / OR, nbj = (I.ny + OC - 1) / OC;
    std::vector<Ev> se, re;
    for (int e = 0; e < I.src_count; ++e) {
      const ernn_source_t& s = p->src_h[I.src_begin + e];
      // regions span [bi OR - KS, (bi + 1) OR + KS): at most three blocks per axis
      for (int bi = std::max(0, (s.i + KS) / OR - 2); bi <= std::min(nbi - 1, (s.i + KS) / OR); ++bi) {
        if (s.i < bi * OR - KS || s.i >= (bi + 1) * OR + KS) continue;
        for (int bj = std::max(0, (s.j + KS) / OC - 2); bj <= std::min(nbj - 1, (s.j + KS) / OC); ++bj) {
          if (s.j < bj * OC - KS || s.j >= (bj + 1) * OC + KS) continue;
          se.push_back(Ev{(int64_t)bi * nbj + bj, s.i, s.j, I.src_begin + e});
        }
      }
    }
    if (!p->snapshot) {
      for (int e = 0; e < I.rec_count; ++e) {
        const ernn_receiver_t& r = p->rec_h[I.rec_begin + e];
        if (!adjoint) {   // recording: the output tile holding the receiver
          re.push_back(Ev{(int64_t)(r.i / OR) * nbj + r.j / OC, r.i, r.j, e});
          continue;
        }
        // injection: every block whose region holds the receiver
        for (int bi = std::max(0, (r.i + KS) / OR - 2); bi <= std::min(nbi - 1, (r.i + KS) / OR); ++bi) {
          if (r.i < bi * OR - KS || r.i >= (bi + 1) * OR + KS) continue;
          for (int bj = std::max(0, (r.j + KS) / OC - 2); bj <= std::min(nbj - 1, (r.j + KS) / OC); ++bj) {
            if (r.j < bj * OC - KS || r.j >= (bj + 1) * OC + KS) continue;
            re.push_back(Ev{(int64_t)bi * nbj + bj, r.i, r.j, e});
          }
        }
      }
    }
    std::stable_sort(se.begin(), se.end(), by_key);
    std::stable_sort(re.begin(), re.end(), by_key);
    size_t si = 0, ri = 0;
    for (int bi = 0; bi < nbi; ++bi) {
      for (int bj = 0; bj < nbj; ++bj) {
        const int64_t id = (int64_t)bi * nbj + bj;
        BlockDev bd{};
        bd.I = p->inst[b];
        bd.inst = b;
        bd.oi0 = bi * OR;
        bd.oj0 = bj * OC;
        bd.set = I.set;
        bd.src_begin = (int)sev.size();
        for (; si < se.size() && se[si].blk == id; ++si) {
          const ernn_source_t& s = p->src_h[se[si].idx];
          sev.push_back(SrcEv{s.i, s.j, s.w, s.comp});
        }
        bd.src_end = (int)sev.size();
        bd.rec_begin = (int)rev.size();
        for (; ri < re.size() && re[ri].blk == id; ++ri) {
          const ernn_receiver_t& r = p->rec_h[I.rec_begin + re[ri].idx];
          if ((int)rev.size() > bd.rec_begin && rev.back().i == r.i && rev.back().j == r.j) bd.rec_shared = 1;
          rev.push_back(RecEv{r.i, r.j, re[ri].idx, 0});
        }
        bd.rec_end = (int)rev.size();
        blocks.push_back(bd);
      }
    }
  }
  first[p->n_inst] = (int)blocks.size();
  if (blocks.size() > ((size_t)1 << 30)) return fail(ERNN_E_CAPACITY, "too many blocks");
  if (adjoint && (int64_t)blocks.size() * blk_waves(RW) > p->partial_cap)
    return fail(ERNN_E_CAPACITY, "time-blocked adjoint needs more partial slots than the workspace holds");
  BlockTiling& nb = *out;
  nb.V = V;
  nb.RW = RW;
  nb.KS = KS;
  nb.n_blocks = (int)blocks.size();
  nb.grid = 8 * ((nb.n_blocks + 7) / 8);
  int r = dev_upload(&nb.d_blocks, blocks);
  if (r == ERNN_OK) r = dev_upload(&nb.d_src, sev);
  if (r == ERNN_OK) r = dev_upload(&nb.d_rec, rev);
  if (r == ERNN_OK) r = dev_upload(&nb.d_inst_first, first);
  if (r != ERNN_OK) free_blocking(nb);
  return r;
}

const char* const kBlkShapes =
    "time-blocked forward shape (V, RW, KS) must have V in {1,2}, RW in {4,8}, KS in {2,4,8}, 4 RW > 2 KS "
    "(damage field / six fields: not V = 2 with RW = 8), or be (1, 8, 5)";

// The time-blocked forward (KS <= 1: one step per launch).  The plan keeps
// its old table if the new one cannot be built.
int set_fwd_blocking(Plan* p, int V, int RW, int KS) {
  if (KS <= 1) {
    free_blocking(p->blk);
    return ERNN_OK;
  }
  if (!block_ok(p->mode, V, RW, KS)) return fail(ERNN_E_INVALID, kBlkShapes);
  BlockTiling nb;
  ERNN_TRY(make_blocking(p, false, V, RW, KS, &nb));
  free_blocking(p->blk);
  p->blk = nb;
  return ERNN_OK;
}

// The time-blocked adjoint (KS <= 1: one step per launch) and, on fuse plans,
// the block table of the fused kernel's recompute half with RWF rows per
// wavefront (0: the adjoint's RW).  Both tables are built first and replace
// the plan's together, so a failure leaves the plan as it was.
int set_adj_blocking(Plan* p, int V, int RW, int KS, int RWF) {
  if (p->rtr && (KS <= 1 || !rev_block_ok(V, RW, KS, !p->lbank)))
    return fail(ERNN_E_INVALID, "a reverse_u plan runs the reversed adjoint: (1, RW, KS), RW in {4, 6, 8}, KS in {2, 4}, "
                                "or RW in {8, 10}, KS in {5, 6} (joint gradient: not RW = 10)");
  if (KS <= 1) {
    free_blocking(p->blk_adj);
    free_blocking(p->blk_ff);
    p->ff_rw = 0;
    return ERNN_OK;
  }
  if (!p->rtr && !ablk_ok(p->mode, p->fuse, V, RW, KS)) return fail(ERNN_E_INVALID, kAblkShapes);
  const int rwf = RWF ? RWF : RW;
  if (p->fuse && !fused_ff_ok(V, RW, KS, rwf))
    return fail(ERNN_E_INVALID, "fused recompute rows per wavefront: the adjoint's RW, or 8 beside RW = 4 with "
                                "KS in {2,4}");
  BlockTiling na, nf;
  ERNN_TRY(make_blocking(p, true, V, RW, KS, &na));
  if (p->fuse) {
    const int r = make_blocking(p, false, V, rwf, KS, &nf);
    if (r != ERNN_OK) {
      free_blocking(na);
      return r;
    }
  }
  free_blocking(p->blk_adj);
  free_blocking(p->blk_ff);
  p->blk_adj = na;
  p->blk_ff = nf;
  p->ff_rw = p->fuse ? rwf : 0;
  return ERNN_OK;
}

void destroy_plan(Plan* p) {
  if (!p) return;
  drop_graphs(p);
  free_tiling(p->fwd);
  free_tiling(p->adj);
  free_blocking(p->blk);
  free_blocking(p->blk_adj);
  free_blocking(p->blk_ff);
  if (p->d_inst) (void)hipFree(p->d_inst);
  if (p->d_call) (void)hipFree(p->d_call);
  if (p->d_set_first) (void)hipFree(p->d_set_first);
  if (p->d_set_list) (void)hipFree(p->d_set_list);
  if (p->d_set_begin) (void)hipFree(p->d_set_begin);
  if (p->d_set_ext_off) (void)hipFree(p->d_set_ext_off);
  if (p->d_set_grad_off) (void)hipFree(p->d_set_grad_off);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  delete p;
}

int validate(const ernn_desc_t* d) {
  if (!d) return fail(ERNN_E_INVALID, "null descriptor");
  if (d->n_inst < 1 || d->n_sets < 1) return fail(ERNN_E_INVALID, "need n_inst >= 1 and n_sets >= 1");
  if (d->param_mode < ERNN_PARAM_CONST6 || d->param_mode > ERNN_PARAM_FIELD6)
    return fail(ERNN_E_INVALID, "unknown param_mode");
  if (d->rec_stride < 1) return fail(ERNN_E_INVALID, "rec_stride must be >= 1");
  if (d->ckpt_every < 2) return fail(ERNN_E_INVALID, "ckpt_every must be >= 2");
  if (d->fwd_bpc < 0 || d->fwd_bpc > 64 || d->adj_bpc < 0 || d->adj_bpc > 64)
    return fail(ERNN_E_INVALID, "blocks per CU must be in [0, 64]");
  if ((d->fwd_v != 0 || d->fwd_r != 0) && !fwd_tile_ok(d->param_mode, d->fwd_v, d->fwd_r))
    return fail(ERNN_E_INVALID, "forward tile (V, R) must have V in {1,2,4}, R a power of two, V*R <= 16 "
                                "(six fields: <= 8)");
  if ((d->adj_v != 0 || d->adj_r != 0) && !adj_tile_ok(d->param_mode, d->adj_v, d->adj_r))
    return fail(ERNN_E_INVALID, "adjoint tile (V, R) must have V in {1,2,4}, R a power of two, V*R <= 16 "
                                "(damage field: V*R <= 8 and R <= 4, six fields: V*R <= 4 and R <= 2)");
  if (d->blk_ks > 1 && !block_ok(d->param_mode, d->blk_v, d->blk_rw, d->blk_ks))
    return fail(ERNN_E_INVALID, "time-blocked forward shape (V, RW, KS) must have V in {1,2}, RW in {4,8}, "
                                "KS in {2,4,8}, 4 RW > 2 KS (damage field / six fields: not V = 2 with RW = 8), "
                                "or be (1, 8, 5)");
  if (d->ablk_ks > 1 && !d->reverse_u && !ablk_ok(d->param_mode, d->fuse != 0, d->ablk_v, d->ablk_rw, d->ablk_ks))
    return fail(ERNN_E_INVALID, kAblkShapes);
  if (d->fixed_c0 != 0 && d->fixed_c0 != 1) return fail(ERNN_E_INVALID, "fixed_c0 must be 0 or 1");
  if (d->fixed_c0 && d->param_mode != ERNN_PARAM_DFIELD)
    return fail(ERNN_E_INVALID, "fixed_c0 applies to the damage-field mode (ERNN_PARAM_DFIELD) only");
  if (d->fixed_c0 && d->fuse) return fail(ERNN_E_INVALID, "fixed_c0 needs fuse = 0");
  if (d->reverse_u != 0 && d->reverse_u != 1) return fail(ERNN_E_INVALID, "reverse_u must be 0 or 1");
  if (d->reverse_u && d->param_mode != ERNN_PARAM_DFIELD)
    return fail(ERNN_E_INVALID, "reverse_u applies to the damage-field mode (ERNN_PARAM_DFIELD) only");
  if (d->reverse_u && d->fuse) return fail(ERNN_E_INVALID, "reverse_u needs fuse = 0");
  if (d->ortho != 0 && d->ortho != 1) return fail(ERNN_E_INVALID, "ortho must be 0 or 1");
  if (d->ortho && d->fixed_c0 != 1)
    return fail(ERNN_E_INVALID, "ortho applies to a damage field over known constants (fixed_c0 = 1) only");
  if (d->reverse_u && (d->ablk_ks <= 1 || !rev_block_ok(d->ablk_v, d->ablk_rw, d->ablk_ks, d->fixed_c0 != 1)))
    return fail(ERNN_E_INVALID, "reverse_u needs a time-blocked adjoint of shape (1, RW, KS), RW in {4, 6, 8}, "
                                "KS in {2, 4}, or RW in {8, 10}, KS in {5, 6} (joint gradient: not RW = 10)");
  if (d->ablk_ff_rw != 0 && (d->ablk_ks <= 1 || !d->fuse || !fused_ff_ok(d->ablk_v, d->ablk_rw, d->ablk_ks, d->ablk_ff_rw)))
    return fail(ERNN_E_INVALID, "ablk_ff_rw: 0, or on a fuse plan with a time-blocked adjoint the adjoint's RW, or 8 "
                                "beside RW = 4 with KS in {2,4}");
  if (!d->inst) return fail(ERNN_E_INVALID, "null instance array");
  if (d->n_src < 0 || d->n_rec < 0) return fail(ERNN_E_INVALID, "negative source/receiver count");
  if ((d->n_src > 0 && !d->src) || (d->n_rec > 0 && !d->rec))
    return fail(ERNN_E_INVALID, "null source/receiver array");
  std::vector<int> set_nx(d->n_sets, -1), set_ny(d->n_sets, -1);
  for (int b = 0; b < d->n_inst; ++b) {
    const ernn_instance_t& I = d->inst[b];
    char buf[160];
    if (I.nx < 1 || I.ny < 1 || I.nx > 65536 || I.ny > 65536) {
      snprintf(buf, sizeof buf, "instance %d: grid %dx%d outside [1, 65536]", b, I.nx, I.ny);
      return fail(ERNN_E_INVALID, buf);
    }
    // the kernels address a padded field with 32-bit element offsets
    if ((round_up(I.nx, kRowAlign) + 3) * (kLeftPad + round_up(I.ny, kColAlign)) >= ((int64_t)1 << 30)) {
      snprintf(buf, sizeof buf, "instance %d: grid %dx%d too large (a padded field must hold < 2^30 cells)", b,
               I.nx, I.ny);
      return fail(ERNN_E_CAPACITY, buf);
    }
    if (I.nt < 1 || I.nt > 100000000) return fail(ERNN_E_INVALID, "nt outside [1, 1e8]");
    if (I.set < 0 || I.set >= d->n_sets) return fail(ERNN_E_INVALID, "instance set out of range");
    if (!(I.h > 0) || !(I.dt > 0) || !(I.rho > 0) || !std::isfinite(I.h) || !std::isfinite(I.dt) ||
        !std::isfinite(I.rho))
      return fail(ERNN_E_INVALID, "h, dt and rho must be finite and positive");
    if (I.src_begin < 0 || I.src_count < 0 || (int64_t)I.src_begin + I.src_count > d->n_src)
      return fail(ERNN_E_INVALID, "instance source slice out of range");
    if (!d->snapshot &&
        (I.rec_begin < 0 || I.rec_count < 0 || (int64_t)I.rec_begin + I.rec_count > d->n_rec))
      return fail(ERNN_E_INVALID, "instance receiver slice out of range");
    for (int e = 0; e < I.src_count; ++e) {
      const ernn_source_t& s = d->src[I.src_begin + e];
      if (s.i < 0 || s.i >= I.nx || s.j < 0 || s.j >= I.ny || (s.comp != 0 && s.comp != 1) ||
          !std::isfinite(s.w)) {
        snprintf(buf, sizeof buf, "instance %d: source %d at (%d,%d) comp %d invalid", b, e, s.i, s.j, s.comp);
        return fail(ERNN_E_INVALID, buf);
      }
    }
    if (!d->snapshot) {
      for (int e = 0; e < I.rec_count; ++e) {
        const ernn_receiver_t& r = d->rec[I.rec_begin + e];
        if (r.i < 0 || r.i >= I.nx || r.j < 0 || r.j >= I.ny) {
          snprintf(buf, sizeof buf, "instance %d: receiver %d at (%d,%d) outside the grid", b, e, r.i, r.j);
          return fail(ERNN_E_INVALID, buf);
        }
      }
    }
    if (d->param_mode != ERNN_PARAM_CONST6) {
      if (set_nx[I.set] < 0) { set_nx[I.set] = I.nx; set_ny[I.set] = I.ny; }
      else if (set_nx[I.set] != I.nx || set_ny[I.set] != I.ny)
        return fail(ERNN_E_INVALID, "instances of one field set must share the grid size");
    }
  }
  for (int s = 0; s < d->n_sets; ++s) {
    bool used = false;
    for (int b = 0; b < d->n_inst && !used; ++b) used = d->inst[b].set == s;
    if (!used) return fail(ERNN_E_INVALID, "every parameter set needs at least one instance");
  }
  return ERNN_OK;
}

int build_plan(const ernn_desc_t* d, Plan* p) {
  ERNN_HIP(hipGetDevice(&p->device));
  ERNN_HIP(hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, p->device));
  p->mode = d->param_mode;
  p->n_inst = d->n_inst;
  p->n_sets = d->n_sets;
  p->P = d->rec_stride;
  p->snapshot = d->snapshot ? 1 : 0;
  p->K = d->ckpt_every;
  p->use_graph = d->use_graph != 0;
  p->fuse = d->fuse != 0;
  p->lbank = d->fixed_c0 == 1 && p->mode == ERNN_PARAM_DFIELD;
  p->rtr = p->mode == ERNN_PARAM_DFIELD && d->reverse_u == 1;
  p->ortho = p->lbank && d->ortho == 1;
  // kernels over a damage field: kDOnly banks L (fixed_c0) and, with reversed u
  // (nothing banked), runs the forward's sliding-row path with the bits of kDField
  p->kmode = (p->lbank || p->rtr) ? kDOnly : p->mode;
  p->n_coef = p->mode == ERNN_PARAM_DFIELD ? 1 : (p->mode == ERNN_PARAM_FIELD6 ? 6 : 0);
  p->n_acc = p->n_coef;
  p->inst_h.assign(d->inst, d->inst + d->n_inst);
  if (d->n_src) p->src_h.assign(d->src, d->src + d->n_src);
  if (d->n_rec && !p->snapshot) p->rec_h.assign(d->rec, d->rec + d->n_rec);
  p->nt_max = 0;
  for (const auto& I : p->inst_h) p->nt_max = std::max(p->nt_max, I.nt);
  p->S = (p->nt_max + p->K - 1) / p->K;

  // sets: first instance, instance lists, external field / gradient offsets
  p->set_first.assign(p->n_sets, -1);
  p->set_begin.assign(p->n_sets + 1, 0);
  for (int b = 0; b < p->n_inst; ++b) {
    const int s = p->inst_h[b].set;
    if (p->set_first[s] < 0) p->set_first[s] = b;
    p->set_begin[s + 1]++;
  }
  for (int s = 0; s < p->n_sets; ++s) p->set_begin[s + 1] += p->set_begin[s];
  p->set_list.assign(p->n_inst, 0);
  {
    std::vector<int> fill(p->set_begin.begin(), p->set_begin.end() - 1);
    for (int b = 0; b < p->n_inst; ++b) p->set_list[fill[p->inst_h[b].set]++] = b;
  }
  p->set_ext_off.assign(p->n_sets, 0);
  p->set_grad_off.assign(p->n_sets, 0);
  int64_t ext = 0;
  for (int s = 0; s < p->n_sets; ++s) {
    const auto& I = p->inst_h[p->set_first[s]];
    p->set_ext_off[s] = ext;
    p->set_grad_off[s] = ext;
    ext += (int64_t)p->n_coef * I.nx * I.ny;
  }
  p->field_floats = ext;
  p->grad_field_doubles = ext;

  // traces
  p->trace_off.assign(p->n_inst + 1, 0);
  for (int b = 0; b < p->n_inst; ++b) {
    const auto& I = p->inst_h[b];
    const int64_t nf = (I.nt + p->P - 1) / p->P;
    const int64_t X = p->snapshot ? (int64_t)I.nx * I.ny : (int64_t)I.rec_count;
    p->trace_off[b + 1] = p->trace_off[b] + 2 * nf * X;
  }

  // workspace layout (floats, 64-float = 256-byte aligned pieces)
  p->inst.assign(p->n_inst, InstDev{});
  int64_t off = 0;
  for (int b = 0; b < p->n_inst; ++b) {
    const auto& I = p->inst_h[b];
    InstDev& D = p->inst[b];
    D.nx = I.nx;
    D.ny = I.ny;
    D.pitch = (int)(kLeftPad + round_up(I.ny, kColAlign));
    D.nt = I.nt;
    D.set = I.set;
    D.n_rec = p->snapshot ? 0 : I.rec_count;
    D.n_frames = (I.nt + p->P - 1) / p->P;
    D.hinv2 = (float)(1.0 / (I.h * I.h));
    D.dt2 = (float)(I.dt * I.dt);
    D.rho = (float)I.rho;
    D.gscale = (double)D.dt2 / (double)D.rho;
    D.field = (round_up(I.nx, kRowAlign) + 3) * (int64_t)D.pitch;
    D.pool_off = off;
    off += round_up((int64_t)p->n_state() * kStateFields * D.field, 64);
    D.bank_off = off;
    off += round_up((int64_t)p->n_bank() * kBankFields * D.field, 64);
    D.trace_off = p->trace_off[b];
    D.wav_off = 0;
  }
  {
    int64_t w = 0;
    for (int b = 0; b < p->n_inst; ++b) {
      p->inst[b].wav_off = w;
      w += p->inst_h[b].nt;
    }
  }
  if (p->n_coef) {
    std::vector<int64_t> coef_off(p->n_sets);
    for (int s = 0; s < p->n_sets; ++s) {
      coef_off[s] = off;
      off += round_up((int64_t)p->n_coef * p->inst[p->set_first[s]].field, 64);
    }
    for (int b = 0; b < p->n_inst; ++b) p->inst[b].coef_off = coef_off[p->inst[b].set];
    for (int b = 0; b < p->n_inst; ++b) {
      p->inst[b].acc64_off = off / 2;
      off += round_up((int64_t)2 * p->n_acc * p->inst[b].field, 64);
    }
  }
  {
    // blocks of the smallest time-blocked output tile (8 x 48 cells), one slot per wavefront
    int64_t nb = 0;
    for (const auto& I : p->inst_h) nb += (int64_t)((I.nx + 7) / 8) * ((I.ny + 47) / 48);
    p->partial_cap = std::max<int64_t>(tiles_for(*p, 1, 1), kMaxBlockWaves * nb);
  }
  p->partial_off = off / 2;
  off += round_up(2 * kPartials * p->partial_cap, 64);
  p->ws_floats = off;

  ERNN_TRY(dev_upload(&p->d_inst, p->inst));
  ERNN_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_call), sizeof(CallDev)));
  ERNN_HIP(hipMemset(p->d_call, 0, sizeof(CallDev)));
  ERNN_TRY(dev_upload(&p->d_set_first, p->set_first));
  ERNN_TRY(dev_upload(&p->d_set_list, p->set_list));
  ERNN_TRY(dev_upload(&p->d_set_begin, p->set_begin));
  ERNN_TRY(dev_upload(&p->d_set_ext_off, p->set_ext_off));
  ERNN_TRY(dev_upload(&p->d_set_grad_off, p->set_grad_off));
  ERNN_TRY(build_tiling(p, &p->fwd, d->fwd_v ? d->fwd_v : 1, d->fwd_r ? d->fwd_r : 4, d->fwd_bpc));
  const int adj_r0 = (d->param_mode == ERNN_PARAM_FIELD6) ? 2 : 4;
  ERNN_TRY(build_tiling(p, &p->adj, d->adj_v ? d->adj_v : 1, d->adj_r ? d->adj_r : adj_r0, d->adj_bpc));
  ERNN_TRY(set_fwd_blocking(p, d->blk_v, d->blk_rw, d->blk_ks));
  ERNN_TRY(set_adj_blocking(p, d->ablk_v, d->ablk_rw, d->ablk_ks, d->ablk_ff_rw));
  ERNN_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
  return ERNN_OK;
}

CallDev to_dev(const ernn_call_t* c) {
  CallDev v{};
  v.ws = static_cast<float*>(c->workspace);
  v.params = c->params;
  v.field = c->field;
  v.wav = c->wavelets;
  v.traces = c->traces;
  v.gtr = c->grad_traces;
  v.grad_params = c->grad_params;
  v.grad_field = c->grad_field;
  return v;
}

int check_call(const Plan* p, const ernn_call_t* c, bool adjoint, bool grads = true) {
  if (!c) return fail(ERNN_E_INVALID, "null call");
  if (!c->workspace) return fail(ERNN_E_INVALID, "null workspace");
  if ((reinterpret_cast<uintptr_t>(c->workspace) & 255) != 0)
    return fail(ERNN_E_INVALID, "workspace must be 256-byte aligned");
  if (!c->wavelets) return fail(ERNN_E_INVALID, "null wavelets");
  if (p->mode != ERNN_PARAM_FIELD6 && !c->params) return fail(ERNN_E_INVALID, "null params");
  if (p->mode != ERNN_PARAM_CONST6 && !c->field) return fail(ERNN_E_INVALID, "null field");
  if (!adjoint && !c->traces) return fail(ERNN_E_INVALID, "null traces");
  if (adjoint) {
    if (!c->grad_traces) return fail(ERNN_E_INVALID, "null grad_traces");
    if (grads && p->mode != ERNN_PARAM_FIELD6 && !p->lbank && !c->grad_params)
      return fail(ERNN_E_INVALID, "null grad_params");
    if (grads && p->mode != ERNN_PARAM_CONST6 && !c->grad_field)
      return fail(ERNN_E_INVALID, "null grad_field");
  }
  return ERNN_OK;
}

StepArgs base_args(const Plan* p, const Tiling& t) {
  StepArgs a{};
  a.inst = p->d_inst;
  a.tiles = t.d_tiles;
  a.src = t.d_src;
  a.rec = t.d_rec;
  a.call = p->d_call;
  a.partial_off = p->partial_off;
  a.n_tiles = t.n_tiles;
  a.mode = p->kmode;
  a.stride = p->P;
  a.snapshot = p->snapshot;
  a.rows = t.R;
  a.fin_base = p->rtr ? p->ck(0) : 0;   // reversed u starts each segment from a checkpoint
  a.fin_k = p->rtr ? p->K : 0;
  return a;
}

int enqueue_pack(const Plan* p, hipStream_t st) {
  for (int s = 0; s < p->n_sets && p->n_coef; ++s) {
    const auto& I = p->inst_h[p->set_first[s]];
    ERNN_HIP(launch_pack_fields(p->d_call, p->d_inst, p->d_set_first, p->d_set_ext_off, p->n_coef, s,
                                (int64_t)I.nx * I.ny, st));
  }
  return ERNN_OK;
}

BlockArgs block_args(const Plan* p, const BlockTiling& bt) {
  BlockArgs a{};
  a.inst = p->d_inst;
  a.blocks = bt.d_blocks;
  a.src = bt.d_src;
  a.rec = bt.d_rec;
  a.call = p->d_call;
  a.n_blocks = bt.n_blocks;
  a.mode = p->kmode;
  a.stride = p->P;
  a.snapshot = p->snapshot;
  a.rw = bt.RW;
  a.ksmax = bt.KS;
  a.fin_base = p->rtr ? p->ck(0) : 0;
  a.fin_k = p->rtr ? p->K : 0;
  a.ortho = p->ortho && a.mode == kDOnly;
  return a;
}
BlockArgs block_args(const Plan* p) { return block_args(p, p->blk); }

// Time-blocked forward-pass launch from level n0 (state slot *in) up to level
// `end` at most: writes the checkpoint levels inside it, its last level to the
// ring slot it does not read (when that level is no checkpoint), and the
// levels of the last checkpoint segment to its bank (lbank: step n banks L^n
// of the level it leaves).  *in becomes the slot of the last level.
void blk_forward_args(const Plan* p, BlockArgs& a, int n0, int end, int* in) {
  const int B = (p->S - 1) * p->K;
  a.n0 = n0;
  a.ks = std::min(a.ksmax, end - n0);
  a.s_in = *in;
  for (int j = 0; j < kMaxBlockSteps; ++j) a.st_store[j] = a.bank_store[j] = -1;
  for (int j = 0; j < a.ks; ++j) {
    const int m = n0 + 1 + j;
    if (p->is_ck(m)) a.st_store[j] = p->ck(m / p->K);
    else if (j == a.ks - 1) a.st_store[j] = p->other_ring(*in);
    const int lev = p->lbank ? m - 1 : m;   // the level this step banks
    if (lev > B && !p->rtr) a.bank_store[j] = p->bank(p->S - 1, lev);
  }
  *in = a.st_store[a.ks - 1];
}

// Last step + 1 of the recompute of segment s: its top level, or with lbank
// one step further, whose step banks L^top (the last segment's top is the
// last level, whose term the adjoint skips: lam^{nt+1} = 0).
int recompute_end(const Plan* p, int s) {
  const int top = p->seg_top(s);
  return p->lbank ? std::min(top + 1, p->nt_max) : top;
}

// Time-blocked recompute launch of segment s from level n0 (state slot *in):
// every level into the segment's bank (lbank: L^n of every level n > sK
// the launch leaves), the last one also to the ring slot the launch does not
// read unless it ends the recompute.
void blk_recompute_args(const Plan* p, BlockArgs& a, int s, int n0, int* in) {
  const int end = recompute_end(p, s);
  a.n0 = n0;
  a.ks = std::min(a.ksmax, end - n0);
  a.s_in = *in;
  for (int j = 0; j < kMaxBlockSteps; ++j) {
    a.st_store[j] = -1;
    if (j >= a.ks) a.bank_store[j] = -1;
    else if (!p->lbank) a.bank_store[j] = p->bank(s, n0 + 1 + j);
    else a.bank_store[j] = n0 + j > s * p->K ? p->bank(s, n0 + j) : -1;
  }
  if (n0 + a.ks < end) {
    a.st_store[a.ks - 1] = p->other_ring(*in);
    *in = a.st_store[a.ks - 1];
  }
}

BlockAdjArgs block_adj_args(const Plan* p) {
  BlockAdjArgs a{};
  a.inst = p->d_inst;
  a.blocks = p->blk_adj.d_blocks;
  a.src = p->blk_adj.d_src;
  a.rec = p->blk_adj.d_rec;
  a.call = p->d_call;
  a.partial_off = p->partial_off;
  a.n_blocks = p->blk_adj.n_blocks;
  a.mode = (p->rtr && !p->lbank) ? kDField : p->kmode;   // reversed u: joint or dJ/dD only
  a.stride = p->P;
  a.snapshot = p->snapshot;
  a.rw = p->blk_adj.RW;
  a.ksmax = p->blk_adj.KS;
  a.ortho = p->ortho && a.mode == kDOnly;
  return a;
}

// Time-blocked adjoint launch over lam levels m0 + 1 -> m0 - ks + 1 (state
// slot *lin, then the other lam slot), reading u^m of segment s from its bank.
void blk_adjoint_args(const Plan* p, BlockAdjArgs& a, int s, int m0, int ks, int* lin) {
  a.m0 = m0;
  a.ks = ks;
  a.s_in = *lin;
  a.s_out = p->other_lam(*lin);
  for (int j = 0; j < kMaxBlockSteps; ++j) a.bank_u[j] = j < ks ? p->bank(s, m0 - j) : 0;
  *lin = a.s_out;
}

// Reversed-u adjoint launch over lam levels m0 + 1 -> m0 - ks + 1 of segment
// s: u runs backward from checkpoint s + 1 (the level min(top, nt) of each
// instance) through the u ring.
void rev_adjoint_args(const Plan* p, BlockAdjArgs& a, int s, int m0, int ks, int* lin, int* uin) {
  a.m0 = m0;
  a.ks = ks;
  a.s_in = *lin;
  a.s_out = p->other_lam(*lin);
  a.u_in = *uin;
  a.u_out = p->other_ur(*uin);
  a.u_ck = p->ck(s + 1);
  a.seg_top = p->seg_top(s);
  for (int j = 0; j < kMaxBlockSteps; ++j) a.bank_u[j] = 0;
  *lin = a.s_out;
  *uin = a.u_out;
}

// forward step n of the single-step schedule: level n (slot in) -> n + 1
void step_forward_args(const Plan* p, StepArgs& a, int n, int* in, bool bank_last_segment) {
  const int m = n + 1;
  a.step = n;
  a.s_in = *in;
  a.s_out = p->is_ck(m) ? p->ck(m / p->K) : p->other_ring(*in);
  const int lev = p->lbank ? n : m;   // the level this step banks (u^m, or L^n)
  a.bank = (bank_last_segment && !p->rtr && lev > (p->S - 1) * p->K) ? p->bank(p->S - 1, lev) : -1;
  *in = a.s_out;
}
// recompute step n of segment s: level n -> n + 1 into the bank (lbank: L^n,
// for n > sK) and a ring slot
void step_recompute_args(const Plan* p, StepArgs& a, int s, int n, int* in) {
  a.step = n;
  a.s_in = *in;
  a.s_out = p->other_ring(*in);
  a.bank = !p->lbank ? p->bank(s, n + 1) : (n > s * p->K ? p->bank(s, n) : -1);
  *in = a.s_out;
}
// adjoint step m of segment s: lam level m + 1 -> m, u^m from the bank
void step_adjoint_args(const Plan* p, StepArgs& a, int s, int m, int* lin) {
  a.step = m;
  a.s_in = *lin;
  a.s_out = p->other_lam(*lin);
  a.bank = p->bank(s, m);
  *lin = a.s_out;
}

// The forward pass keeps the checkpoints and, for the last checkpoint
// segment, every level's u in that segment's bank, as its recompute would: the
// adjoint that follows on the same workspace starts on them instead of
// recomputing the segment (seg_ws).
int enqueue_forward(const Plan* p, hipStream_t st) {
  ERNN_TRY(enqueue_pack(p, st));
  int in = p->ck(0);
  if (p->blk.KS > 1) {
    BlockArgs b = block_args(p);
    b.record = 1;
    for (int n0 = 0; n0 < p->nt_max; n0 += p->blk.KS) {
      blk_forward_args(p, b, n0, p->nt_max, &in);
      ERNN_HIP(launch_fwd_block(b, p->blk.V, p->blk.grid, st));
    }
    return ERNN_OK;
  }
  StepArgs a = base_args(p, p->fwd);
  a.record = 1;
  for (int n = 0; n < p->nt_max; ++n) {
    step_forward_args(p, a, n, &in, true);
    ERNN_HIP(launch_fwd_step(a, p->fwd.V, p->fwd.blocks, p->fwd.loop, st));
  }
  return ERNN_OK;
}

// skip_last: the last segment's states are in its bank (left by the forward
// pass on this workspace), so its recompute is skipped.
int enqueue_adjoint(const Plan* p, hipStream_t st, bool skip_last) {
  ERNN_TRY(enqueue_pack(p, st));
  // zero: the two lam slots (lam level nt + 1 is the zero state), per-tile
  // partials, fp64 per-cell accumulators
  for (int b = 0; b < p->n_inst; ++b) {
    const InstDev& D = p->inst[b];
    ERNN_HIP(launch_fill_ws(p->d_call, D.pool_off + (int64_t)p->lam(0) * kStateFields * D.field,
                            2 * kStateFields * D.field, st));
    if (p->n_acc)
      ERNN_HIP(launch_fill_ws(p->d_call, 2 * D.acc64_off, 2 * (int64_t)p->n_acc * D.field, st));
  }
  ERNN_HIP(launch_fill_ws(p->d_call, 2 * p->partial_off, 2 * kPartials * p->partial_cap, st));

  StepArgs f = base_args(p, p->fwd);
  f.record = 0;
  StepArgs a = base_args(p, p->adj);
  BlockArgs rb = block_args(p);
  rb.record = 0;
  // recompute segment s whole (time-blocked when enabled)
  auto recompute_segment = [&](int s) -> int {
    int in = p->ck(s);
    const int end = recompute_end(p, s);
    if (p->blk.KS > 1) {
      for (int n0 = s * p->K; n0 < end; n0 += p->blk.KS) {
        blk_recompute_args(p, rb, s, n0, &in);
        ERNN_HIP(launch_fwd_block(rb, p->blk.V, p->blk.grid, st));
      }
      return ERNN_OK;
    }
    for (int n = s * p->K; n < end; ++n) {
      step_recompute_args(p, f, s, n, &in);
      ERNN_HIP(launch_fwd_step(f, p->fwd.V, p->fwd.blocks, p->fwd.loop, st));
    }
    return ERNN_OK;
  };
  int lin = p->lam(0);
  if (p->fuse && p->blk_adj.KS > 1) {
    // The last segment is recomputed alone (unless stored); afterwards every
    // time-blocked adjoint launch of segment s carries the next KS recompute
    // steps of segment s - 1 into the other bank (segment s - 1 has K steps,
    // so a short last segment leaves recompute launches to run alone).
    if (!skip_last) ERNN_TRY(recompute_segment(p->S - 1));
    const int KS = p->blk_adj.KS;
    FusedBlockArgs fb{};
    fb.adj = block_adj_args(p);
    fb.fwd = block_args(p, p->blk_ff);
    fb.fwd.record = 0;
    fb.adj_grid = p->blk_adj.grid;
    const int grid = p->blk_adj.grid + p->blk_ff.grid;
    for (int s = p->S - 1; s >= 0; --s) {
      const int top = p->seg_top(s), bot = s * p->K + 1;
      int n0 = (s - 1) * p->K;   // next recompute step of segment s - 1
      int rin = s >= 1 ? p->ck(s - 1) : 0;
      for (int m0 = top; m0 >= bot; m0 -= KS) {
        const int ks = std::min(KS, m0 - bot + 1);
        blk_adjoint_args(p, fb.adj, s, m0, ks, &lin);
        if (s >= 1 && n0 < s * p->K) {
          blk_recompute_args(p, fb.fwd, s - 1, n0, &rin);
          ERNN_HIP(launch_fused_block(fb, p->blk_adj.V, grid, st));
          n0 += KS;
        } else {
          ERNN_HIP(launch_adj_block(fb.adj, p->blk_adj.V, p->blk_adj.grid, st));
        }
      }
      for (; s >= 1 && n0 < s * p->K; n0 += KS) {
        BlockArgs r = block_args(p, p->blk_ff);
        r.record = 0;
        blk_recompute_args(p, r, s - 1, n0, &rin);
        ERNN_HIP(launch_fwd_block(r, p->blk_ff.V, p->blk_ff.grid, st));
      }
    }
  } else if (p->rtr) {
    // reversed u: each segment's launches run u backward from the checkpoint
    // of its top level beside lam; no recompute, no bank
    BlockAdjArgs ba = block_adj_args(p);
    int uin = p->ur(0);
    for (int s = p->S - 1; s >= 0; --s) {
      const int top = p->seg_top(s), bot = s * p->K + 1;
      for (int m0 = top; m0 >= bot; m0 -= p->blk_adj.KS) {
        rev_adjoint_args(p, ba, s, m0, std::min(p->blk_adj.KS, m0 - bot + 1), &lin, &uin);
        ERNN_HIP(launch_rev_block(ba, p->blk_adj.V, p->blk_adj.grid, st));
      }
    }
  } else if (!p->fuse && p->blk_adj.KS > 1) {
    BlockAdjArgs ba = block_adj_args(p);
    for (int s = p->S - 1; s >= 0; --s) {
      const int top = p->seg_top(s), bot = s * p->K + 1;
      if (s < p->S - 1 || !skip_last) ERNN_TRY(recompute_segment(s));
      for (int m0 = top; m0 >= bot; m0 -= p->blk_adj.KS) {
        blk_adjoint_args(p, ba, s, m0, std::min(p->blk_adj.KS, m0 - bot + 1), &lin);
        ERNN_HIP(launch_adj_block(ba, p->blk_adj.V, p->blk_adj.grid, st));
      }
    }
  } else if (!p->fuse) {
    for (int s = p->S - 1; s >= 0; --s) {
      const int top = p->seg_top(s), bot = s * p->K + 1;
      if (s < p->S - 1 || !skip_last) ERNN_TRY(recompute_segment(s));
      for (int m = top; m >= bot; --m) {
        step_adjoint_args(p, a, s, m, &lin);
        ERNN_HIP(launch_adj_step(a, p->adj.V, p->adj.blocks, p->adj.loop, st));
      }
    }
  } else {
    // The last segment is recomputed alone (unless stored); afterwards every
    // adjoint step of segment s shares its launch with one recompute step of
    // segment s - 1, which fills the other bank (segment s - 1 always has K
    // steps).
    if (!skip_last) ERNN_TRY(recompute_segment(p->S - 1));
    FusedArgs fa{};
    fa.fwd = base_args(p, p->adj);   // the fused kernel runs both on the adjoint's tiles
    fa.fwd.record = 0;
    for (int s = p->S - 1; s >= 0; --s) {
      const int top = p->seg_top(s), bot = s * p->K + 1;
      int rin = s >= 1 ? p->ck(s - 1) : 0;
      int idx = 0;
      for (int m = top; m >= bot; --m, ++idx) {
        step_adjoint_args(p, a, s, m, &lin);
        if (s >= 1 && idx < p->K) {
          fa.adj = a;
          step_recompute_args(p, fa.fwd, s - 1, (s - 1) * p->K + idx, &rin);
          ERNN_HIP(launch_fused_step(fa, p->adj.V, p->adj.fused_blocks, st));
        } else {
          ERNN_HIP(launch_adj_step(a, p->adj.V, p->adj.blocks, p->adj.loop, st));
        }
      }
      // a short last segment leaves recompute steps of segment s - 1 to run alone
      for (; s >= 1 && idx < p->K; ++idx) {
        step_recompute_args(p, f, s - 1, (s - 1) * p->K + idx, &rin);
        ERNN_HIP(launch_fwd_step(f, p->fwd.V, p->fwd.blocks, p->fwd.loop, st));
      }
    }
  }
  if (p->mode != ERNN_PARAM_FIELD6 && !p->lbank) {
    if (p->blk_adj.KS > 1)
      ERNN_HIP(launch_finalize_const(p->d_call, p->partial_off, p->d_set_list, p->d_set_begin,
                                     p->blk_adj.d_inst_first, blk_waves(p->blk_adj.RW), p->n_sets, st));
    else
      ERNN_HIP(launch_finalize_const(p->d_call, p->partial_off, p->d_set_list, p->d_set_begin, p->adj.d_inst_first, 1,
                                     p->n_sets, st));
  }
  if (p->n_acc) {
    for (int s = 0; s < p->n_sets; ++s) {
      const auto& I = p->inst_h[p->set_first[s]];
      ERNN_HIP(launch_finalize_field(p->d_call, p->d_inst, p->d_set_list, p->d_set_begin, p->d_set_grad_off,
                                     p->n_acc, s, (int64_t)I.nx * I.ny, st));
    }
  }
  return ERNN_OK;
}

using Enqueue = int (*)(const Plan*, hipStream_t);
int enqueue_adjoint_full(const Plan* p, hipStream_t st) { return enqueue_adjoint(p, st, false); }
int enqueue_adjoint_skip(const Plan* p, hipStream_t st) { return enqueue_adjoint(p, st, true); }

// Capture an enqueue function into a graph on the plan's private stream.
int capture(Plan* p, Enqueue fn, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  ERNN_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
  const int r = fn(p, p->cap_stream);
  const hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
  if (r != ERNN_OK) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return fail(ERNN_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(ERNN_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  return ERNN_OK;
}

int run_pass(Plan* p, const ernn_call_t* c, hipStream_t st, bool adjoint) {
  ERNN_HIP(hipSetDevice(p->device));
  // An adjoint uses up the stored last segment of its workspace; a forward
  // pass stores it anew.
  const bool skip = adjoint && c->workspace == p->seg_ws;
  p->seg_ws = nullptr;
  ERNN_HIP(launch_set_call(p->d_call, to_dev(c), st));
  Enqueue fn = !adjoint ? enqueue_forward : (skip ? enqueue_adjoint_skip : enqueue_adjoint_full);
  if (!p->use_graph) {
    ERNN_TRY(fn(p, st));
  } else {
    hipGraphExec_t* g = !adjoint ? &p->g_fwd : (skip ? &p->g_adj_skip : &p->g_adj);
    // Every captured kernel reaches the workspace and the per-call buffers
    // only through d_call, so one capture serves every call.
    if (!*g) ERNN_TRY(capture(p, fn, g));
    ERNN_HIP(hipGraphLaunch(*g, st));
  }
  if (!adjoint) p->seg_ws = c->workspace;
  return ERNN_OK;
}

}  // namespace

// The other C-ABI source files (ernn_fem.cpp) report failures through the
// same per-thread message.
namespace ernn {
int set_last_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace ernn

extern "C" {

int ernn_abi_version(void) { return ERNN_ABI_VERSION; }

const char* ernn_last_error(void) { return g_err.c_str(); }

int ernn_plan_create(const ernn_desc_t* desc, void** plan_out) {
  if (!plan_out) return fail(ERNN_E_INVALID, "null plan_out");
  *plan_out = nullptr;
  ERNN_TRY(validate(desc));
  Plan* p = new Plan();
  const int r = build_plan(desc, p);
  if (r != ERNN_OK) {
    destroy_plan(p);
    return r;
  }
  *plan_out = p;
  return ERNN_OK;
}

int ernn_plan_destroy(void* plan) {
  destroy_plan(static_cast<Plan*>(plan));
  return ERNN_OK;
}

int ernn_plan_info(const void* plan, ernn_plan_info_t* info) {
  if (!plan || !info) return fail(ERNN_E_INVALID, "null argument");
  const Plan* p = static_cast<const Plan*>(plan);
  std::memset(info, 0, sizeof *info);
  info->workspace_bytes = p->ws_floats * (int64_t)sizeof(float);
  info->trace_floats = p->trace_off.back();
  info->grad_field_doubles = p->grad_field_doubles;
  info->field_floats = p->field_floats;
  info->n_frames_max = (p->nt_max + p->P - 1) / p->P;
  info->ckpt_every = p->K;
  info->n_ckpt = p->S;
  info->fwd_v = p->fwd.V;
  info->fwd_r = p->fwd.R;
  info->adj_v = p->adj.V;
  info->adj_r = p->adj.R;
  info->fwd_tiles = p->fwd.n_tiles;
  info->adj_tiles = p->adj.n_tiles;
  info->use_graph = p->use_graph ? 1 : 0;
  info->fwd_bpc = p->fwd.bpc;
  info->adj_bpc = p->adj.bpc;
  info->fwd_blocks = p->fwd.blocks;
  info->adj_blocks = p->adj.blocks;
  info->n_cu = p->n_cu;
  info->fuse = p->fuse ? 1 : 0;
  info->blk_v = p->blk.V;
  info->blk_rw = p->blk.RW;
  info->blk_ks = p->blk.KS;
  info->blk_blocks = p->blk.n_blocks;
  info->ablk_v = p->blk_adj.V;
  info->ablk_rw = p->blk_adj.RW;
  info->ablk_ks = p->blk_adj.KS;
  info->ablk_blocks = p->blk_adj.n_blocks;
  info->ablk_ff_rw = p->ff_rw;
  info->fixed_c0 = p->lbank ? 1 : 0;
  info->reverse_u = p->rtr ? 1 : 0;
  info->ortho = p->ortho ? 1 : 0;
  return ERNN_OK;
}

int ernn_trace_offsets(const void* plan, int64_t* trace_off) {
  if (!plan || !trace_off) return fail(ERNN_E_INVALID, "null argument");
  const Plan* p = static_cast<const Plan*>(plan);
  std::copy(p->trace_off.begin(), p->trace_off.end(), trace_off);
  return ERNN_OK;
}

int ernn_workspace_init(void* plan, void* workspace, void* stream) {
  if (!plan || !workspace) return fail(ERNN_E_INVALID, "null argument");
  Plan* p = static_cast<Plan*>(plan);
  if ((reinterpret_cast<uintptr_t>(workspace) & 255) != 0)
    return fail(ERNN_E_INVALID, "workspace must be 256-byte aligned");
  ERNN_HIP(hipSetDevice(p->device));
  if (workspace == p->seg_ws) p->seg_ws = nullptr;
  ERNN_HIP(launch_fill_zero(static_cast<float*>(workspace), p->ws_floats, static_cast<hipStream_t>(stream)));
  return ERNN_OK;
}

int ernn_forward(void* plan, const ernn_call_t* call, void* stream) {
  if (!plan) return fail(ERNN_E_INVALID, "null plan");
  Plan* p = static_cast<Plan*>(plan);
  ERNN_TRY(check_call(p, call, false));
  return run_pass(p, call, static_cast<hipStream_t>(stream), false);
}

int ernn_adjoint(void* plan, const ernn_call_t* call, void* stream) {
  if (!plan) return fail(ERNN_E_INVALID, "null plan");
  Plan* p = static_cast<Plan*>(plan);
  ERNN_TRY(check_call(p, call, true));
  return run_pass(p, call, static_cast<hipStream_t>(stream), true);
}

int ernn_plan_set_tiling(void* plan, int which, int v, int r, int bpc) {
  if (!plan || (which != 0 && which != 1)) return fail(ERNN_E_INVALID, "bad plan or which");
  Plan* p = static_cast<Plan*>(plan);
  ERNN_HIP(hipSetDevice(p->device));
  ERNN_HIP(hipDeviceSynchronize());
  drop_graphs(p);
  return build_tiling(p, which == 0 ? &p->fwd : &p->adj, v, r, bpc);
}

int ernn_plan_set_blocking(void* plan, int which, int v, int rw, int ks) {
  if (!plan || which < 0 || which > 2) return fail(ERNN_E_INVALID, "bad plan or which");
  Plan* p = static_cast<Plan*>(plan);
  if (which == 2) {   // the fused kernel's recompute half
    if (!p->fuse || p->blk_adj.KS <= 1 || v != p->blk_adj.V || ks != p->blk_adj.KS ||
        !fused_ff_ok(v, p->blk_adj.RW, ks, rw))
      return fail(ERNN_E_INVALID, "fused recompute shape: needs a fuse plan with time-blocked adjoint (V, RW, KS); "
                                  "(V, RW, KS) itself or (1, 8, KS) beside RW = 4, KS in {2,4}");
    ERNN_HIP(hipSetDevice(p->device));
    ERNN_HIP(hipDeviceSynchronize());
    drop_graphs(p);
    BlockTiling nf;
    ERNN_TRY(make_blocking(p, false, v, rw, ks, &nf));
    free_blocking(p->blk_ff);
    p->blk_ff = nf;
    p->ff_rw = rw;
    return ERNN_OK;
  }
  ERNN_HIP(hipSetDevice(p->device));
  ERNN_HIP(hipDeviceSynchronize());
  drop_graphs(p);
  // a new adjoint shape resets the recompute half to the adjoint's rows
  return which == 0 ? set_fwd_blocking(p, v, rw, ks) : set_adj_blocking(p, v, rw, ks, 0);
}

int ernn_time_steps(void* plan, int which, int steps, const ernn_call_t* call, void* stream,
                    float* ms_out) {
  if (!plan || !ms_out || steps < 1 || which < 0 || which > 4) return fail(ERNN_E_INVALID, "bad argument");
  Plan* p = static_cast<Plan*>(plan);
  if (which == 2 && !p->fuse) return fail(ERNN_E_INVALID, "fused timing needs a plan with fuse = 1");
  if (which == 3 && p->blk.KS <= 1) return fail(ERNN_E_INVALID, "time-blocked timing needs blk_ks > 1");
  if (which == 4 && p->blk_adj.KS <= 1) return fail(ERNN_E_INVALID, "time-blocked adjoint timing needs ablk_ks > 1");
  if (which == 1 && p->rtr)
    return fail(ERNN_E_INVALID, "a reverse_u plan has no single-step adjoint (which = 4 times its reversed launches)");
  ERNN_TRY(check_call(p, call, which == 1 || which == 2 || which == 4, /*grads=*/false));  // never finalizes
  hipStream_t st = static_cast<hipStream_t>(stream);
  ERNN_HIP(hipSetDevice(p->device));
  if (call->workspace == p->seg_ws) p->seg_ws = nullptr;   // the timed launches overwrite the state slots
  ERNN_HIP(launch_set_call(p->d_call, to_dev(call), st));
  // both events are destroyed on every exit path
  struct Events {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  ERNN_HIP(hipEventCreate(&ev.e0));
  ERNN_HIP(hipEventCreate(&ev.e1));
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  ERNN_HIP(hipEventRecord(e0, st));
  const int nsteps = std::min(steps, p->nt_max);
  // The timed launches follow the schedules' slot use: the adjoint reads u^m
  // from bank 0 (fuse plans: bank 1, while segment 0 is recomputed into bank 0).
  if (which == 4 && p->rtr) {   // `steps` reversed-u adjoint launches within segment 0
    const int KS = p->blk_adj.KS;
    const int launches = std::min(steps, (p->nt_max + KS - 1) / KS);
    BlockAdjArgs ba = block_adj_args(p);
    int lin = p->lam(0), uin = p->ur(0);
    const int top0 = p->seg_top(0);
    for (int L = 0; L < launches; ++L) {
      // m0 walks down segment 0 from its top and starts over below KS
      const int span = std::max(top0 - KS + 1, 1);
      const int m0 = top0 - (L * KS) % span;
      rev_adjoint_args(p, ba, 0, m0, std::min(KS, m0), &lin, &uin);
      ERNN_HIP(launch_rev_block(ba, p->blk_adj.V, p->blk_adj.grid, st));
    }
  } else if (which == 4) {   // `steps` time-blocked adjoint launches of KS steps each
    const int KS = p->blk_adj.KS;
    const int launches = std::min(steps, (p->nt_max + KS - 1) / KS);
    const int useg = p->fuse ? 1 : 0;   // bank_of(useg) is the bank the adjoint reads
    FusedBlockArgs fb{};
    fb.adj = block_adj_args(p);
    fb.fwd = block_args(p, p->blk_ff);
    fb.fwd.record = 0;
    fb.adj_grid = p->blk_adj.grid;
    int lin = p->lam(0), rin = p->ck(0), n0 = 0;
    const int top0 = p->seg_top(0);
    for (int L = 0; L < launches; ++L) {
      // reverse steps within one segment's bank: m0 cycles over useg K + KS .. useg K + K
      const int span = std::max(KS, p->K - p->K % KS);
      const int m0 = useg * p->K + KS + (L * KS) % std::max(span - KS + 1, 1);
      blk_adjoint_args(p, fb.adj, useg, std::min(m0, useg * p->K + p->K), std::min(KS, p->K), &lin);
      if (!p->fuse) {
        ERNN_HIP(launch_adj_block(fb.adj, p->blk_adj.V, p->blk_adj.grid, st));
        continue;
      }
      if (n0 >= top0) { n0 = 0; rin = p->ck(0); }
      blk_recompute_args(p, fb.fwd, 0, n0, &rin);
      n0 += KS;
      ERNN_HIP(launch_fused_block(fb, p->blk_adj.V, p->blk_adj.grid + p->blk_ff.grid, st));
    }
  } else if (which == 3) {   // `steps` time-blocked forward launches of KS steps each
    BlockArgs b = block_args(p);
    b.record = 1;
    const int launches = std::min(steps, (p->nt_max + p->blk.KS - 1) / p->blk.KS);
    int in = p->ck(0);
    for (int L = 0; L < launches; ++L) {
      blk_forward_args(p, b, L * p->blk.KS, p->nt_max, &in);
      ERNN_HIP(launch_fwd_block(b, p->blk.V, p->blk.grid, st));
    }
  } else if (which == 0) {
    StepArgs a = base_args(p, p->fwd);
    a.record = 1;
    int in = p->ck(0);
    for (int n = 0; n < nsteps; ++n) {
      step_forward_args(p, a, n, &in, false);
      ERNN_HIP(launch_fwd_step(a, p->fwd.V, p->fwd.blocks, p->fwd.loop, st));
    }
  } else {
    StepArgs a = base_args(p, p->adj);
    FusedArgs fa{};
    fa.fwd = base_args(p, p->adj);
    fa.fwd.record = 0;
    const int useg = which == 2 ? 1 : 0;
    int lin = p->lam(0), rin = p->ck(0);
    for (int q = 0; q < nsteps; ++q) {
      const int m = useg * p->K + p->K - q % p->K;   // u^m from segment useg's bank
      step_adjoint_args(p, a, useg, m, &lin);
      if (which == 1) {
        ERNN_HIP(launch_adj_step(a, p->adj.V, p->adj.blocks, p->adj.loop, st));
      } else {   // recompute steps of segment 0 into bank 0 while the adjoint reads bank 1
        const int n = q % p->K;
        if (n == 0) rin = p->ck(0);
        fa.adj = a;
        step_recompute_args(p, fa.fwd, 0, n, &rin);
        ERNN_HIP(launch_fused_step(fa, p->adj.V, p->adj.fused_blocks, st));
      }
    }
  }
  ERNN_HIP(hipEventRecord(e1, st));
  ERNN_HIP(hipEventSynchronize(e1));
  ERNN_HIP(hipEventElapsedTime(ms_out, e0, e1));
  return ERNN_OK;
}

}  // extern "C"
