// Internal layout shared by the HIP kernels (ernn_kernels.hip) and the host
// plan builder (ernn_plan.cpp).  Plain structs, identical on host and device.
//
// State layout in the workspace ("padded field"): a field of an nx x ny grid
// occupies round_up(nx, 16) + 3 rows of `pitch` floats.  Row r of the grid
// (-1 <= r <= round_up(nx, 16) + 1) starts at (r + 1) * pitch, and column j
// (-32 <= j < pitch - 32) sits at +32 + j, so interior cell (i, j) is at
// (i + 1) * pitch + 32 + j.  Everything outside 0 <= i < nx, 0 <= j < ny is a
// ghost cell that stays zero: the reference's zero padding
// (Parameter_inversion/Inversion_adam.py:50).  pitch = 32 + round_up(ny, 256),
// so every interior row starts 128-byte aligned and a wavefront's last halo
// read (column round_up(ny, 64 V)) lands at most on the zero left pad of the
// next row; the extra ghost rows let a tile of up to 16 rows load all its
// rows without bounds checks.
#pragma once
#include <stdint.h>

namespace ernn {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kLeftPad = 32;
constexpr int kColAlign = 256;
constexpr int kRowAlign = 16;

// Kernel modes: the ABI's parameter modes, plus kDOnly for a damage field
// over known constants (fixed_c0): the bank then holds (dt^2/rho) A(C0) u^m
// instead of u^m and the adjoint forms dJ/dD only.
enum ParamMode : int { kConst6 = 0, kDField = 1, kField6 = 2, kDOnly = 3 };

// Recurrence state.  Both recurrences, u^{n+1} = 2 u^n - u^{n-1} + d^n
// (PI:286-291) and the reverse-time adjoint, are carried as (u, c, v) per
// cell and component: v^n = u^n - u^{n-1} is the increment, u the value and c
// the rounding error of the last sum u + v (Kahan compensation).  One step is
//   v' = v + d,  y = v' + c,  u' = u + y,  c' = y - (u' - u),
// so the FP32 rounding of the state does not accumulate over thousands of
// steps.  A state slot holds the six padded fields u_x, u_y, c_x, c_y, v_x,
// v_y of one time level; a bank slot holds u_x, u_y only (the recomputed
// forward states the adjoint reads for its gradient terms), or with kDOnly
// the two components of (dt^2/rho) A(C0) u.
constexpr int kStateFields = 6;
constexpr int kBankFields = 2;

// fp64 partial sums per tile for the constant gradients (kConst6, kDField).
// Slots 0-5 hold S11, S22, S12, S16, S26, S66 with the stencil weight h^-2 (or
// h^-2 / 4 for slot 2) applied; slots 6-7 hold the mixed-derivative parts of
// C16 and C26 (weight h^-2 / 4).  dJ/dC = (S11, S22, S12, S16 + 2 s6,
// S26 + 2 s7, S66 + S12).
constexpr int kPartials = 8;

// One instance = one shot on one grid with one parameter set.
struct InstDev {
  int nx, ny, pitch, nt;
  int set;          // parameter set index
  int n_rec;        // receivers (receiver mode)
  int n_frames;     // recorded frames, ceil(nt / P)
  int pad0;
  float hinv2;      // fl32(h^-2)
  float dt2;        // fl32(dt^2)
  float rho;        // fl32(rho)
  float pad1;
  double gscale;    // (double)dt2 / (double)rho: d(u^{n+1})/dC = gscale * dA/dC u^n
  int64_t field;    // floats per padded field, (nx + 3) * pitch
  int64_t pool_off; // float offset of this instance's state slots [n_state][kStateFields][field]
  int64_t bank_off; // float offset of its bank slots [n_bank][kBankFields][field]
  int64_t coef_off; // float offset of the packed coefficient fields of its set
  int64_t acc64_off;// double offset of its fp64 per-cell accumulators
  int64_t trace_off;// float offset of its traces in the trace array
  int64_t wav_off;  // float offset of its source time function
};

// One tile = R rows x (64 * V) columns of one instance, swept by one wavefront.
struct TileDev {
  int inst, i0, j0, set;
  int src_begin, src_end;   // source events of this tile (sorted by row)
  int rec_begin, rec_end;   // receiver events of this tile (sorted by row)
};

struct SrcEv { int i, j; float w; int comp; };
struct RecEv { int i, j, rec, pad; };

// Per-call device pointers, written by a one-thread kernel before each
// forward/adjoint pass so that captured graphs stay valid across calls.
struct CallDev {
  float* ws;               // workspace
  const float* params;     // [n_sets][6] constants (kConst6, kDField)
  const float* field;      // [n_sets][n_coef][nx*ny] stiffness fields (kDField: D, kField6: 6 C)
  const float* wav;        // source time functions
  float* traces;           // forward output
  const float* gtr;        // adjoint input: dJ/d traces (same layout)
  double* grad_params;     // [n_sets][6]
  double* grad_field;      // [n_sets][nx*ny] (kDField) or [n_sets][6][nx*ny] (kField6)
};

// Per-step scalars (kernel argument by value).
struct StepArgs {
  const InstDev* inst;
  const TileDev* tiles;
  const SrcEv* src;
  const RecEv* rec;
  const CallDev* call;
  int64_t partial_off;     // double offset in the workspace of [n_tiles][kPartials] fp64 partials
  int n_tiles;
  int mode;
  int step;                // forward: n (level n -> n+1); adjoint: m (lam level m+1 -> m)
  int s_in, s_out;         // state slots of the input and output levels
  int bank;                // forward: bank slot receiving u^{n+1} (or -1); adjoint: bank slot holding u^m
  int stride;              // recording stride P
  int record;              // forward: write traces on recorded steps
  int snapshot;            // traces are whole fields
  int rows;                // R
  // reversed-u plans (forward): the step that reaches an instance's last level
  // I.nt also stores it into state slot fin_base + ceil(I.nt / fin_k) (fin_k = 0: off)
  int fin_base, fin_k;
};

// A launch that runs one adjoint step and, on the same tiles, one recompute
// step of the previous checkpoint segment (they touch disjoint slots).
struct FusedArgs {
  StepArgs adj;
  StepArgs fwd;
};

// Time-blocked forward: one launch advances up to kMaxBlockSteps steps over
// block regions whose halo (one cell per step on every side) covers the extra
// steps.  A block is blk_waves(RW) wavefronts stacked vertically, each
// holding RW rows x 64 V columns of the region in registers; the output tile
// is the region shrunk by KS cells on every side.
constexpr int kMaxBlockSteps = 8;

// Wavefronts per time-blocked block: two-row bands (RW = 2) stack 8, which
// keeps the region as tall as RW = 4 with half the serial work per
// wavefront; every other shape stacks kWavesPerBlock.  Only the FEM kernel
// builds RW = 2: on the ERNN stencil the two-row shapes lost to RW = 4
// (one barrier per step across 8 wavefronts costs more than the shorter
// chain saves).
constexpr int kMaxBlockWaves = 8;
constexpr int blk_waves(int rw) { return rw == 2 ? kMaxBlockWaves : kWavesPerBlock; }

struct BlockDev {
  InstDev I;                // the instance's record, carried inline: a block's first field loads then wait
                            // on one dependent load (this record) instead of two (record, then instance)
  int inst, oi0, oj0, set;  // output tile origin (grid coordinates)
  int src_begin, src_end;   // sources inside the block's region, sorted by row
  int rec_begin, rec_end;   // receivers inside the output tile, sorted by row
  int rec_shared;           // 1: two receivers share a cell (the kernels walk the list instead of a cell map)
  int pad;
};

struct BlockArgs {
  const InstDev* inst;
  const BlockDev* blocks;
  const SrcEv* src;
  const RecEv* rec;
  const CallDev* call;
  int n_blocks;
  int mode;
  int n0;                        // first step of the launch: starts from level n0
  int ks;                        // steps in this launch (<= KS of the kernel)
  int s_in;                      // state slot of level n0
  int st_store[kMaxBlockSteps];  // state slot receiving level n0+1+s, or -1
  int bank_store[kMaxBlockSteps];// bank slot receiving u^{n0+1+s}, or -1
  int stride;                    // recording stride P
  int record;
  int snapshot;
  int rw, ksmax;                 // compiled shape: rows per wavefront, steps per launch
  int fin_base, fin_k;           // as StepArgs: an instance's last level also into slot fin_base + ceil(I.nt / fin_k)
  int ortho;                     // kDOnly: C16 = C26 = 0 in every set (the operator sums drop those terms)
};

// Time-blocked adjoint (constant and damage modes): one launch runs the
// reverse steps m = m0, m0 - 1, ..., m0 - ks + 1 on the block regions.
struct BlockAdjArgs {
  const InstDev* inst;
  const BlockDev* blocks;        // receiver lists cover the whole region (injection)
  const SrcEv* src;              // the region's point forces (the reversed-u adjoint)
  const RecEv* rec;
  const CallDev* call;
  int64_t partial_off;           // [blocks][blk_waves(RW)][kPartials] fp64 partials
  int n_blocks;
  int mode;
  int m0, ks;                    // lam levels m0+1 -> m0-ks+1
  int s_in, s_out;               // state slots of lam level m0+1 (input) and m0-ks+1 (output)
  int bank_u[kMaxBlockSteps];    // bank slot holding u^{m0-s}
  int stride;
  int snapshot;
  int rw, ksmax;
  // reversed-u launches (rev_block_kernel): u runs backward beside lam.  The
  // launch that holds an instance's first level of the segment, min(seg_top,
  // I.nt), loads (u, c, v) of that level from checkpoint slot u_ck; later
  // launches read (u^{m0}, c, v^{m0+1}) from u ring slot u_in.  The launch
  // writes (u, c, v) of its level m0 - ks, v of the level above, to u_out.
  int u_in, u_out, u_ck, seg_top;
  int ortho;                     // as BlockArgs (reversed u over known constants)
};

// Fused time-blocked backward: a time-blocked adjoint launch of segment s and
// a time-blocked recompute launch of segment s - 1 (same V, RW, KS) in one
// grid; blocks below adj_grid run the adjoint.
struct FusedBlockArgs {
  BlockAdjArgs adj;
  BlockArgs fwd;
  int adj_grid;
  int pad;
};

// --------------------------------------------------------------------------
// FEM generator (P1 triangles on the checkerboard mesh, lumped mass, central
// differences).  Node (i, j), i along x, is at (i + 1) * pitch + kLeftPad + j
// of a padded field of nx + kFemExtraRows rows (the layout above, with room
// for a tile of up to kFemMaxRows rows past the last interior row).  The
// boundary ring of nodes is held at zero, so it and the ghost cells are the
// zero padding of the interior unknowns.
// Workspace: ux slots 0-3, uy slots 4-7 as two pairs each (pair p = slots
// 2p, 2p + 1 of the component).  One step per launch keeps u^m in slot m & 1
// of pair 0; the time-blocked launches read (u^{n0-1}, u^{n0}) from pair
// L & 1 (older first) and write their last two states to the other pair.
// Slot 8: the packed damage factor; slots 9, 10: the nodal load weights of
// the x and y components; the 4 x 4 edge-flux matrix at 11 fields.
constexpr int kFemMaxRows = 8;
constexpr int kFemExtraRows = kFemMaxRows + 3;
constexpr int kFemUySlot = 4;
constexpr int kFemDSlot = 8;
constexpr int kFemLoadSlot = 9;
constexpr int kFemSlots = 11;

enum FemMode : int { kFemConst = 0, kFemDamage = 1 };

// One tile = R node rows x 63 node columns, swept by one wavefront.
struct FemTile {
  int i0, j0;
  int has_loads;          // 1: a nodal load lies in the tile (the kernel reads the load fields)
  int pad;
};

struct FemCallDev {
  float* ws;
  const float* params;    // C11, C22, C12, C16, C26, C66 (damage mode: the intact C0)
  const float* field;     // damage mode: per-node factor, nx * ny, row-major (x, y)
  const float* wav;       // time factors, nt
  float* frames_x;        // [n_frames][nx][ny]
  float* frames_y;
};

// Time-blocked FEM launch: a block of blk_waves(RW) wavefronts x RW node
// rows x 64 node columns advances up to KS steps from registers; the output
// tile is the block region shrunk by KS nodes on every side.
constexpr int kFemMaxBlockSteps = 8;

struct FemBlock {
  int oi0, oj0;           // output tile origin (grid node)
  int has_loads;          // 1: a nodal load lies in the block region
  int pad;
};

struct FemBlockArgs {
  const FemCallDev* call;
  const FemBlock* blocks;
  int64_t field;
  int64_t frame;          // nx * ny
  int n_blocks, blocks_per_xcd;
  int nx, ny, pitch;
  int mode;
  int n0;                 // first step: reads u^{n0-1}, u^{n0}, produces u^{n0+1} .. u^{n0+ks}
  int ks;                 // steps in this launch (<= KS)
  int src_pair, dst_pair;
  int stride;             // frame k = u^{k stride}
  float dt2, rinv, m_even, m_odd, scale;
  int rw, ksmax;          // compiled shape
};

struct FemStepArgs {
  const FemCallDev* call;
  const FemTile* tiles;
  int64_t field;          // floats per padded field
  int64_t rec_off;        // float offset of the frame this step writes, or -1
  int n_tiles;
  int nx, ny, pitch;
  int step;               // n: u^{n+1} from u^n (slot n & 1) and u^{n-1}, written over u^{n-1}
  int mode;
  float dt2;              // fl32(dt^2)
  float rinv;             // fl32(1 / rho)
  float m_even, m_odd;    // lumped mass of interior nodes with i + j even (8 elements) / odd (4)
  float scale;            // frame scale
};

}  // namespace ernn
