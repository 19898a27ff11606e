// Host side of the FEM observed-data generator: descriptor validation, the
// tile and nodal-load tables, the run schedule (zero state and frames, pack
// the damage factor, form the edge-flux matrix, nt - 2 step launches) with
// optional hipGraph capture, and the ernn_fem_* C ABI of include/ernn.h.
//
// Reference: Parameter_inversion/FEM_generate_data.py (PF) :196-590 and
// Damage_inversion/FEM_generate_data.py (DF) :770-1224.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ernn.h"
#include "ernn_internal.h"

namespace ernn {
int set_last_error(int code, const std::string& msg);
bool fem_shape_ok(int V, int R);
hipError_t launch_fem_step(const FemStepArgs& a, int V, int R, int blocks, hipStream_t st);
hipError_t launch_fem_setup(const FemCallDev* call, int64_t field, double dx, double dy, hipStream_t st);
hipError_t launch_fem_pack(const FemCallDev* call, int64_t field, int nx, int ny, int pitch, hipStream_t st);
hipError_t launch_fem_zero(const FemCallDev* call, int which, int64_t n, hipStream_t st);
hipError_t launch_fem_loads(const FemCallDev* call, const SrcEv* loads, int n, int64_t field, int pitch,
                            hipStream_t st);
bool fem_block_shape_ok(int RW, int KS);
hipError_t launch_fem_block(const FemBlockArgs& a, int grid, hipStream_t st);
hipError_t launch_fem_set_call(FemCallDev* dst, const FemCallDev& v, hipStream_t st);
}  // namespace ernn

using namespace ernn;

namespace {

int fail(int code, const std::string& msg) { return set_last_error(code, msg); }

#define FEM_HIP(x)                                                                   \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      return fail(ERNN_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define FEM_TRY(x)              \
  do {                          \
    int r_ = (x);               \
    if (r_ != ERNN_OK) return r_; \
  } while (0)

struct FemPlan {
  int device = 0;
  int nx = 0, ny = 0, nt = 0, mode = 0, stride = 1;
  double dx = 0, dy = 0, dt = 0, rho = 0;
  float scale = 0.7f;
  int V = 1, R = 1;
  int pitch = 0;
  int64_t field = 0;       // floats per padded field
  int64_t ws_floats = 0;
  int64_t frame = 0;       // nx * ny
  int n_frames = 0;
  int n_tiles = 0, blocks = 0;
  float dt2 = 0, rinv = 0, m_even = 0, m_odd = 0;
  int use_graph = 0;
  int RW = 0, KS = 0;                 // time-blocked shape (KS = 0: one step per launch)
  int n_blocks = 0, blk_grid = 0;
  std::vector<ernn_source_t> loads;   // interior loads
  FemBlock* d_blocks = nullptr;
  FemTile* d_tiles = nullptr;
  SrcEv* d_loads = nullptr;
  FemCallDev* d_call = nullptr;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t graph = nullptr;
};

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

void free_tables(FemPlan* p) {
  if (p->d_tiles) (void)hipFree(p->d_tiles);
  if (p->d_loads) (void)hipFree(p->d_loads);
  if (p->d_blocks) (void)hipFree(p->d_blocks);
  p->d_tiles = nullptr;
  p->d_loads = nullptr;
  p->d_blocks = nullptr;
}

// Time-blocked output tiles of (WB RW - 2 KS) rows x (64 - 2 KS) columns over
// the interior, in row-band order; the grid maps block b to logical block
// (b % 8) x blocks_per_xcd + b / 8 so that each XCD sweeps a band of rows.
// A block is flagged when a load lies in its region (output tile + KS halo).
int build_blocks(FemPlan* p) {
  const int wb = blk_waves(p->RW);   // wavefronts per block
  const int orows = wb * p->RW - 2 * p->KS, ocols = kWave - 2 * p->KS;
  const int br = (p->nx - 2 + orows - 1) / orows, bc = (p->ny - 2 + ocols - 1) / ocols;
  std::vector<FemBlock> blocks;
  for (int a = 0; a < br; ++a)
    for (int b = 0; b < bc; ++b) {
      FemBlock bl{1 + a * orows, 1 + b * ocols, 0, 0};
      for (const auto& l : p->loads)
        if (l.i >= bl.oi0 - p->KS && l.i < bl.oi0 - p->KS + wb * p->RW && l.j >= bl.oj0 - p->KS &&
            l.j < bl.oj0 - p->KS + kWave)
          bl.has_loads = 1;
      blocks.push_back(bl);
    }
  p->n_blocks = (int)blocks.size();
  p->blk_grid = 8 * ((p->n_blocks + 7) / 8);
  FEM_HIP(hipMalloc(&p->d_blocks, sizeof(FemBlock) * blocks.size()));
  FEM_HIP(hipMemcpy(p->d_blocks, blocks.data(), sizeof(FemBlock) * blocks.size(), hipMemcpyHostToDevice));
  return ERNN_OK;
}

void destroy(FemPlan* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->graph) (void)hipGraphExecDestroy(p->graph);
  free_tables(p);
  if (p->d_call) (void)hipFree(p->d_call);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  delete p;
}

// Tiles of R node rows x 63 node columns over the interior nodes (rows 1 ..
// nx-2 from row 1, columns 1 .. ny-2 from column 1: a wavefront's lane 0
// holds the column before the tile), ordered row band by row band so that
// each XCD's contiguous share of tiles is a band of rows; tiles holding a
// nodal load are flagged (only they read the load fields).
constexpr int kFemTileCols = kWave - 1;

int build_tables(FemPlan* p) {
  free_tables(p);
  const int cols = kFemTileCols;
  const int tr = (p->nx - 2 + p->R - 1) / p->R;
  const int tc = (p->ny - 2 + cols - 1) / cols;
  p->n_tiles = tr * tc;
  const int per_xcd = (p->n_tiles + 7) / 8;
  p->blocks = 8 * std::max(1, (per_xcd + kWavesPerBlock - 1) / kWavesPerBlock);
  std::vector<int> has(p->n_tiles, 0);
  for (const auto& l : p->loads) has[((l.i - 1) / p->R) * tc + (l.j - 1) / cols] = 1;
  std::vector<FemTile> tiles(p->n_tiles);
  for (int a = 0; a < tr; ++a)
    for (int b = 0; b < tc; ++b) tiles[a * tc + b] = FemTile{1 + a * p->R, 1 + b * cols, has[a * tc + b], 0};
  FEM_HIP(hipMalloc(&p->d_tiles, sizeof(FemTile) * tiles.size()));
  FEM_HIP(hipMemcpy(p->d_tiles, tiles.data(), sizeof(FemTile) * tiles.size(), hipMemcpyHostToDevice));
  if (p->KS > 1) FEM_TRY(build_blocks(p));
  if (!p->loads.empty()) {
    std::vector<SrcEv> flat;
    for (const auto& l : p->loads) flat.push_back(SrcEv{l.i, l.j, l.w, l.comp});
    FEM_HIP(hipMalloc(&p->d_loads, sizeof(SrcEv) * flat.size()));
    FEM_HIP(hipMemcpy(p->d_loads, flat.data(), sizeof(SrcEv) * flat.size(), hipMemcpyHostToDevice));
  }
  return ERNN_OK;
}

int check_call(const FemPlan* p, const ernn_fem_call_t* c) {
  if (!c) return fail(ERNN_E_INVALID, "null call");
  if (!c->workspace || !c->params || !c->wavelet || !c->frames_x || !c->frames_y)
    return fail(ERNN_E_INVALID, "workspace, params, wavelet and both frame arrays are required");
  if ((reinterpret_cast<uintptr_t>(c->workspace) & 255) != 0)
    return fail(ERNN_E_INVALID, "workspace must be 256-byte aligned");
  if (p->mode == ERNN_FEM_DAMAGE && !c->field) return fail(ERNN_E_INVALID, "damage mode needs the factor field");
  return ERNN_OK;
}

FemCallDev to_dev(const ernn_fem_call_t* c) {
  FemCallDev d;
  d.ws = static_cast<float*>(c->workspace);
  d.params = c->params;
  d.field = c->field;
  d.wav = c->wavelet;
  d.frames_x = c->frames_x;
  d.frames_y = c->frames_y;
  return d;
}

FemStepArgs step_args(const FemPlan* p, int n) {
  FemStepArgs a{};
  a.call = p->d_call;
  a.tiles = p->d_tiles;
  a.field = p->field;
  a.rec_off = ((n + 1) % p->stride == 0) ? (int64_t)((n + 1) / p->stride) * p->frame : -1;
  a.n_tiles = p->n_tiles;
  a.nx = p->nx;
  a.ny = p->ny;
  a.pitch = p->pitch;
  a.step = n;
  a.mode = p->mode;
  a.dt2 = p->dt2;
  a.rinv = p->rinv;
  a.m_even = p->m_even;
  a.m_odd = p->m_odd;
  a.scale = p->scale;
  return a;
}

// State, damage factor and flux matrix; frames zeroed (frame 0 = u^0 = 0 and
// the boundary ring stays zero in every frame).
int enqueue_setup(const FemPlan* p, hipStream_t st) {
  FEM_HIP(launch_fem_zero(p->d_call, 0, (int64_t)kFemSlots * p->field, st));
  if (p->mode == ERNN_FEM_DAMAGE) FEM_HIP(launch_fem_pack(p->d_call, p->field, p->nx, p->ny, p->pitch, st));
  FEM_HIP(launch_fem_loads(p->d_call, p->d_loads, (int)p->loads.size(), p->field, p->pitch, st));
  FEM_HIP(launch_fem_setup(p->d_call, p->field, p->dx, p->dy, st));
  FEM_HIP(launch_fem_zero(p->d_call, 1, (int64_t)p->n_frames * p->frame, st));
  FEM_HIP(launch_fem_zero(p->d_call, 2, (int64_t)p->n_frames * p->frame, st));
  return ERNN_OK;
}

FemBlockArgs block_args(const FemPlan* p) {
  FemBlockArgs a{};
  a.call = p->d_call;
  a.blocks = p->d_blocks;
  a.field = p->field;
  a.frame = p->frame;
  a.n_blocks = p->n_blocks;
  a.blocks_per_xcd = p->blk_grid / 8;
  a.nx = p->nx;
  a.ny = p->ny;
  a.pitch = p->pitch;
  a.mode = p->mode;
  a.stride = p->stride;
  a.dt2 = p->dt2;
  a.rinv = p->rinv;
  a.m_even = p->m_even;
  a.m_odd = p->m_odd;
  a.scale = p->scale;
  a.rw = p->RW;
  a.ksmax = p->KS;
  return a;
}

// Steps n = 1 .. n_end - 1.  Time-blocked: launch L runs steps n0 .. n0 + ks - 1
// from pair L & 1 into pair (L + 1) & 1 (u^0 = u^1 = 0 in pair 0).
int enqueue_steps(const FemPlan* p, int n_end, hipStream_t st) {
  if (p->KS > 1) {
    FemBlockArgs a = block_args(p);
    for (int n0 = 1, L = 0; n0 < n_end; ++L) {
      a.n0 = n0;
      a.ks = std::min(p->KS, n_end - n0);
      a.src_pair = L & 1;
      a.dst_pair = (L + 1) & 1;
      FEM_HIP(launch_fem_block(a, p->blk_grid, st));
      n0 += a.ks;
    }
    return ERNN_OK;
  }
  for (int n = 1; n < n_end; ++n) FEM_HIP(launch_fem_step(step_args(p, n), p->V, p->R, p->blocks, st));
  return ERNN_OK;
}

int enqueue_run(const FemPlan* p, hipStream_t st) {
  FEM_TRY(enqueue_setup(p, st));
  return enqueue_steps(p, p->nt - 1, st);
}

int capture(FemPlan* p) {
  hipGraph_t g = nullptr;
  FEM_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
  const int r = enqueue_run(p, p->cap_stream);
  const hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
  if (r != ERNN_OK) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return fail(ERNN_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(ERNN_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  return ERNN_OK;
}

}  // namespace

extern "C" {

int ernn_fem_plan_create(const ernn_fem_desc_t* d, void** plan_out) {
  if (!d || !plan_out) return fail(ERNN_E_INVALID, "null argument");
  *plan_out = nullptr;
  if (d->nx < 3 || d->ny < 3) return fail(ERNN_E_INVALID, "nx and ny must be >= 3 (an interior node)");
  if (d->nt < 2) return fail(ERNN_E_INVALID, "nt must be >= 2");
  if (d->mode != ERNN_FEM_CONST && d->mode != ERNN_FEM_DAMAGE) return fail(ERNN_E_INVALID, "unknown mode");
  if (!(d->dx > 0) || !(d->dy > 0) || !(d->dt > 0) || !(d->rho > 0) || !std::isfinite(d->dx) ||
      !std::isfinite(d->dy) || !std::isfinite(d->dt) || !std::isfinite(d->rho))
    return fail(ERNN_E_INVALID, "dx, dy, dt and rho must be positive and finite");
  if (d->rec_stride < 1) return fail(ERNN_E_INVALID, "rec_stride must be >= 1");
  if (!std::isfinite(d->rec_scale)) return fail(ERNN_E_INVALID, "rec_scale must be finite");
  if (d->n_loads < 0 || (d->n_loads > 0 && !d->loads)) return fail(ERNN_E_INVALID, "bad load array");
  int V = d->tile_v, R = d->tile_r;
  if (V == 0 && R == 0) {
    V = 1;
    R = 1;
  }
  if (!fem_shape_ok(V, R)) return fail(ERNN_E_INVALID, "tile (V, R) must be (1, 1), (1, 2), (1, 4) or (1, 8)");
  const int KS = d->blk_ks > 1 ? d->blk_ks : 0, RW = KS ? d->blk_rw : 0;
  if (KS && !fem_block_shape_ok(RW, KS))
    return fail(ERNN_E_INVALID, "time-blocked shape (RW, KS) must be (2, 2), (2, 4), (4, 2), (4, 4), (8, 2), (8, 4) or (8, 8)");
  const int64_t pitch = kLeftPad + round_up(d->ny, kColAlign);
  const int64_t field = (int64_t)(d->nx + kFemExtraRows) * pitch;
  if (field >= (int64_t(1) << 30)) return fail(ERNN_E_CAPACITY, "grid too large (a padded field must hold < 2^30 floats)");
  FemPlan* p = new FemPlan();
  p->nx = d->nx;
  p->ny = d->ny;
  p->nt = d->nt;
  p->mode = d->mode;
  p->stride = d->rec_stride;
  p->dx = d->dx;
  p->dy = d->dy;
  p->dt = d->dt;
  p->rho = d->rho;
  p->scale = d->rec_scale;
  p->V = V;
  p->R = R;
  p->RW = RW;
  p->KS = KS;
  p->pitch = (int)pitch;
  p->field = field;
  p->ws_floats = (int64_t)kFemSlots * field + 64;
  p->frame = (int64_t)d->nx * d->ny;
  p->n_frames = (d->nt + d->rec_stride - 1) / d->rec_stride;
  p->use_graph = d->use_graph ? 1 : 0;
  // the reference's scalars in its own rounding: dt*dt and 1/rho enter FP32
  // arithmetic as FP32 (PF:482-493); the lumped mass is rho * area / 3 in
  // FP32 (PF:335), added once per element touching the node (PF:333-337)
  p->dt2 = (float)(d->dt * d->dt);
  p->rinv = (float)(1.0 / d->rho);
  const float m = (float)(d->rho * (d->dx * d->dy / 2.0) / 3.0);
  float s = 0.f;
  for (int k = 1; k <= 8; ++k) {
    s += m;
    if (k == 4) p->m_odd = s;
  }
  p->m_even = s;
  for (int64_t k = 0; k < d->n_loads; ++k) {
    const ernn_source_t& l = d->loads[k];
    if (l.i < 0 || l.i >= d->nx || l.j < 0 || l.j >= d->ny || (l.comp != 0 && l.comp != 1) || !std::isfinite(l.w)) {
      destroy(p);
      return fail(ERNN_E_INVALID, "load " + std::to_string(k) + ": node outside the grid, bad component or weight");
    }
    if (l.i >= 1 && l.i <= d->nx - 2 && l.j >= 1 && l.j <= d->ny - 2) p->loads.push_back(l);
  }
  // one weight per (node, component): duplicates are summed in FP32 in array order
  std::stable_sort(p->loads.begin(), p->loads.end(), [](const ernn_source_t& x, const ernn_source_t& y) {
    return x.comp != y.comp ? x.comp < y.comp : (x.i != y.i ? x.i < y.i : x.j < y.j);
  });
  std::vector<ernn_source_t> merged;
  for (const auto& l : p->loads) {
    if (!merged.empty() && merged.back().i == l.i && merged.back().j == l.j && merged.back().comp == l.comp)
      merged.back().w += l.w;
    else
      merged.push_back(l);
  }
  p->loads.swap(merged);
  int r = ERNN_OK;
  hipError_t e = hipGetDevice(&p->device);
  if (e == hipSuccess) e = hipMalloc(&p->d_call, sizeof(FemCallDev));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) r = fail(ERNN_E_HIP, std::string("plan set-up: ") + hipGetErrorString(e));
  if (r == ERNN_OK) r = build_tables(p);
  if (r != ERNN_OK) {
    destroy(p);
    return r;
  }
  *plan_out = p;
  return ERNN_OK;
}

int ernn_fem_plan_destroy(void* plan) {
  if (!plan) return fail(ERNN_E_INVALID, "null plan");
  FemPlan* p = static_cast<FemPlan*>(plan);
  (void)hipSetDevice(p->device);
  (void)hipDeviceSynchronize();
  destroy(p);
  return ERNN_OK;
}

int ernn_fem_plan_info(const void* plan, ernn_fem_info_t* info) {
  if (!plan || !info) return fail(ERNN_E_INVALID, "null argument");
  const FemPlan* p = static_cast<const FemPlan*>(plan);
  *info = ernn_fem_info_t{};
  info->workspace_bytes = p->ws_floats * (int64_t)sizeof(float);
  info->frame_floats = (int64_t)p->n_frames * p->frame;
  info->n_frames = p->n_frames;
  info->tile_v = p->V;
  info->tile_r = p->R;
  info->n_tiles = p->n_tiles;
  info->blocks = p->blocks;
  info->n_loads = (int32_t)p->loads.size();
  info->use_graph = p->use_graph;
  info->blk_rw = p->RW;
  info->blk_ks = p->KS;
  info->blk_blocks = p->n_blocks;
  return ERNN_OK;
}

int ernn_fem_run(void* plan, const ernn_fem_call_t* call, void* stream) {
  if (!plan) return fail(ERNN_E_INVALID, "null plan");
  FemPlan* p = static_cast<FemPlan*>(plan);
  FEM_TRY(check_call(p, call));
  hipStream_t st = static_cast<hipStream_t>(stream);
  FEM_HIP(hipSetDevice(p->device));
  FEM_HIP(launch_fem_set_call(p->d_call, to_dev(call), st));
  if (!p->use_graph) return enqueue_run(p, st);
  // the captured kernels reach every buffer through d_call: one capture serves every call
  if (!p->graph) FEM_TRY(capture(p));
  FEM_HIP(hipGraphLaunch(p->graph, st));
  return ERNN_OK;
}

int ernn_fem_set_tiling(void* plan, int v, int r) {
  if (!plan) return fail(ERNN_E_INVALID, "null plan");
  FemPlan* p = static_cast<FemPlan*>(plan);
  if (!fem_shape_ok(v, r)) return fail(ERNN_E_INVALID, "tile (V, R) must be (1, 1), (1, 2), (1, 4) or (1, 8)");
  FEM_HIP(hipSetDevice(p->device));
  FEM_HIP(hipDeviceSynchronize());
  if (p->graph) {
    (void)hipGraphExecDestroy(p->graph);
    p->graph = nullptr;
  }
  p->V = v;
  p->R = r;
  return build_tables(p);
}

int ernn_fem_set_blocking(void* plan, int rw, int ks) {
  if (!plan) return fail(ERNN_E_INVALID, "null plan");
  FemPlan* p = static_cast<FemPlan*>(plan);
  if (ks > 1 && !fem_block_shape_ok(rw, ks))
    return fail(ERNN_E_INVALID, "time-blocked shape (RW, KS) must be (2, 2), (2, 4), (4, 2), (4, 4), (8, 2), (8, 4) or (8, 8)");
  FEM_HIP(hipSetDevice(p->device));
  FEM_HIP(hipDeviceSynchronize());
  if (p->graph) {
    (void)hipGraphExecDestroy(p->graph);
    p->graph = nullptr;
  }
  p->KS = ks > 1 ? ks : 0;
  p->RW = p->KS ? rw : 0;
  return build_tables(p);
}

int ernn_fem_time_steps(void* plan, int steps, const ernn_fem_call_t* call, void* stream, float* ms_out) {
  if (!plan || !ms_out || steps < 1) return fail(ERNN_E_INVALID, "bad argument");
  FemPlan* p = static_cast<FemPlan*>(plan);
  FEM_TRY(check_call(p, call));
  hipStream_t st = static_cast<hipStream_t>(stream);
  FEM_HIP(hipSetDevice(p->device));
  FEM_HIP(launch_fem_set_call(p->d_call, to_dev(call), st));
  FEM_TRY(enqueue_setup(p, st));
  hipEvent_t e0, e1;
  FEM_HIP(hipEventCreate(&e0));
  FEM_HIP(hipEventCreate(&e1));
  FEM_HIP(hipEventRecord(e0, st));
  const int r = enqueue_steps(p, std::min(steps + 1, p->nt - 1), st);
  FEM_HIP(hipEventRecord(e1, st));
  FEM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FEM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (r != ERNN_OK) return r;
  *ms_out = ms;
  return ERNN_OK;
}

}  // extern "C"
