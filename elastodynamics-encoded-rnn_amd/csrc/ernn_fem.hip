// HIP kernels of the FEM observed-data generator on gfx950 (CDNA4, wave64).
//
// Reference semantics (paths relative to the reference repository root):
//   PF = Parameter_inversion/FEM_generate_data.py, DF = Damage_inversion/FEM_generate_data.py
//   mesh            PF:117-189: cell (i, j) with corners n0 = (i, j), n1 = (i+1, j),
//                   n2 = (i+1, j+1), n3 = (i, j+1) is split along n0-n2 when i + j
//                   is even ([n0,n1,n2], [n0,n2,n3]) and along n1-n3 when odd
//                   ([n0,n1,n3], [n1,n2,n3])
//   lumped mass     PF:327-339 (interior nodes: 8 elements when i + j is even, 4 when odd)
//   internal force  PF:417-457, per-element C at the node nearest the centroid DF:1003-1026
//   time step       PF:481-497: u^{n+1} = 2u^n - u^{n-1} + dt^2 (f_ext - K u^n) / M, boundary held at 0
//   frames          x0.7 (PF:525-526), every stride-th state (PF:926-930)
//
// The reference scatters each element's force into its three nodes with
// index_add_ (atomics, one pass per Gauss point).  Here each node gathers:
// the gradients of a P1 triangle on this mesh are edge differences, so a
// triangle with x-edge (a -> b, one row apart) and y-edge (c -> d, one column
// apart) has constant strain from the two edge differences
//   e = (dX ux, dX uy, dY ux, dY uy),  dX = u(b) - u(a),  dY = u(d) - u(c),
// and its nodal forces are +-PX on a, b and +-PY on c, d with
//   (PX.x, PX.y, PY.x, PY.y) = Q e,
//   Q = [[rx C11, rx C16, C16/2, C12/2], [rx C16, rx C66, C66/2, C26/2],
//        [C16/2, C66/2, ry C66, ry C26], [C12/2, C26/2, ry C26, ry C22]],
//   rx = dy / (2 dx), ry = dx / (2 dy)
// (|detJ| x the three Gauss weights = the triangle area dx dy / 2; Q is the
// exact element matrix B^T C B area written on edge differences, formed in
// FP64 and rounded once).  Every cell's two triangles give four corner
// forces; a node's force is the sum over its four cells.
//
// Work decomposition: one wavefront owns R node rows x 63 node columns and
// sweeps its cell rows top to bottom from registers.  Lane l holds node
// column j0 - 1 + l and computes the one cell to its right, (ci, j0 - 1 + l),
// per cell row; the right neighbour's u comes from lane l + 1 by a DPP
// wavefront shift (lane 63 loads it), and the left cell's corner forces
// F2, F3 come from lane l - 1 the same way, so every cell is computed once
// per cell row.  Lane 0 only feeds lane 1 (its own left cell belongs to the
// previous tile), which is why a tile stores 63 columns.  A cell row's lower
// corner forces are carried in registers to the next node row: (R + 1) / R
// cells per node, each u value loaded once per tile, no atomics, no LDS.
// (A first version computed the V + 1 cells around V columns per lane:
// four cells per node at R = 1, VALU-bound at 80 % of peak on 2049^2.)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ernn_internal.h"
#include "ernn_device.h"

namespace ernn {

struct Flux {
  float px, py, qx, qy;   // PX = (px, py) on the x-edge nodes, PY = (qx, qy) on the y-edge nodes
};

__device__ __forceinline__ Flux tri_flux(const float (&Q)[16], float ex, float ey, float fx, float fy) {
  Flux r;
  r.px = fmaf(Q[3], fy, fmaf(Q[2], fx, fmaf(Q[1], ey, Q[0] * ex)));
  r.py = fmaf(Q[7], fy, fmaf(Q[6], fx, fmaf(Q[5], ey, Q[4] * ex)));
  r.qx = fmaf(Q[11], fy, fmaf(Q[10], fx, fmaf(Q[9], ey, Q[8] * ex)));
  r.qy = fmaf(Q[15], fy, fmaf(Q[14], fx, fmaf(Q[13], ey, Q[12] * ex)));
  return r;
}

__device__ __forceinline__ Flux scale_flux(Flux f, float s) {
  f.px *= s; f.py *= s; f.qx *= s; f.qy *= s;
  return f;
}

// One node update in the reference's operation order (PF:482-493):
// u^{n+1} = (2u^n - u^{n-1}) + (dt^2 (f_ext - f_int)) / M.  With f_ext = 0
// this equals (2u^n - u^{n-1}) - (dt^2 f_int) / M bit for bit.
__device__ __forceinline__ float fem_update(float u, float up, float fext, float fint, float dt2, float M) {
  const float base = 2.0f * u - up;
  return base + (dt2 * (fext - fint)) / M;
}

// DPP wave_shr:1 with bound_ctrl: lane l receives lane l - 1's value, lane 0 gets 0.
__device__ __forceinline__ float from_left(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}

// Corner forces of the cell (ci, c) with corners n0 = (ci, c), n1 = (ci + 1, c),
// n2 = (ci + 1, c + 1), n3 = (ci, c + 1), from u at its four corners (a: row
// ci, b: row ci + 1; 0: column c, 1: column c + 1) and the y-edge difference
// of row ci (l*, carried from the cell row above); also returns the y-edge
// difference of row ci + 1.  d*: the damage factor at the corners.
struct Corners {
  float f0x, f0y, f1x, f1y, f2x, f2y, f3x, f3y;
  float rxd, ryd;
};

template <int MODE>
__device__ __forceinline__ Corners cell_corners(const float (&Q)[16], bool even, float xa0, float xa1, float xb0,
                                                float xb1, float ya0, float ya1, float yb0, float yb1, float lx,
                                                float ly, float da0, float da1, float db0, float db1) {
  Corners k;
  // x-edges at columns c (bottom) and c + 1 (top), y-edges of rows ci (left) and ci + 1 (right)
  const float h0x = xb0 - xa0, h0y = yb0 - ya0;
  const float h1x = xb1 - xa1, h1y = yb1 - ya1;
  k.rxd = xb1 - xb0;
  k.ryd = yb1 - yb0;
  // triangle 1 holds the bottom x-edge n0 -> n1 and, for even cells, the
  // right y-edge n1 -> n2 ([n0,n1,n2]), for odd cells the left n0 -> n3
  // ([n0,n1,n3]); triangle 2 holds the top x-edge n3 -> n2 and the other y-edge
  Flux t1 = tri_flux(Q, h0x, h0y, even ? k.rxd : lx, even ? k.ryd : ly);
  Flux t2 = tri_flux(Q, h1x, h1y, even ? lx : k.rxd, even ? ly : k.ryd);
  if constexpr (MODE == kFemDamage) {
    // C of a triangle = factor at the node nearest its centroid:
    // [n0,n1,n2] -> n1, [n0,n2,n3] -> n3, [n0,n1,n3] -> n0, [n1,n2,n3] -> n2
    t1 = scale_flux(t1, even ? db0 : da0);
    t2 = scale_flux(t2, even ? da1 : db1);
  }
  const float lqx = even ? t2.qx : t1.qx, lqy = even ? t2.qy : t1.qy;   // PY of the left-edge triangle
  const float rqx = even ? t1.qx : t2.qx, rqy = even ? t1.qy : t2.qy;   // PY of the right-edge triangle
  k.f0x = -t1.px - lqx; k.f0y = -t1.py - lqy;
  k.f1x = t1.px - rqx;  k.f1y = t1.py - rqy;
  k.f2x = t2.px + rqx;  k.f2y = t2.py + rqy;
  k.f3x = lqx - t2.px;  k.f3y = lqy - t2.py;
  return k;
}

// DPP wave_shl:1 with bound_ctrl: lane l receives lane l + 1's value, lane 63 gets 0.
__device__ __forceinline__ float from_right(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

template <int MODE, int R>
__device__ __forceinline__ void fem_tile(const FemStepArgs& a, const FemTile& tl, gfloat* ws, const gfloat* wav,
                                         gfloat* frx, gfloat* fry, const float (&Q)[16]) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t pitch = (uint32_t)a.pitch;
  const int i0 = tl.i0, j0 = tl.j0;
  const int c = j0 - 1 + lane;                   // the lane's node column
  const uint32_t lo = (uint32_t)lane;            // from column j0 - 1
  const uint32_t ho = (uint32_t)(lane + (lane == kWave - 1));   // lane 63 also reads column j0 + 63
  const int cur = a.step & 1;
  const gfloat* uxc = ws + (int64_t)cur * a.field;
  const gfloat* uyc = ws + (int64_t)(kFemUySlot + cur) * a.field;
  gfloat* uxp = ws + (int64_t)(1 - cur) * a.field;
  gfloat* uyp = ws + (int64_t)(kFemUySlot + 1 - cur) * a.field;
  const gfloat* dfl = ws + (int64_t)kFemDSlot * a.field;
  const gfloat* wlx = ws + (int64_t)kFemLoadSlot * a.field;
  const gfloat* wly = ws + (int64_t)(kFemLoadSlot + 1) * a.field;
  const bool loads = tl.has_loads != 0;   // wavefront-uniform

  // every load of the tile first: R + 2 rows of u^n (and of the damage
  // factor), R rows of u^{n-1} (and of the load weights, in tiles that have any)
  RawRow<1> rx[R + 2], ry[R + 2], rd[MODE == kFemDamage ? R + 2 : 1];
#pragma unroll
  for (int r = 0; r < R + 2; ++r) {
    const uint32_t ro = row_off(pitch, i0 - 1 + r, j0 - 1);
    rx[r] = load_raw<1>(uxc, ro, lo, ho);
    ry[r] = load_raw<1>(uyc, ro, lo, ho);
    if constexpr (MODE == kFemDamage) rd[r] = load_raw<1>(dfl, ro, lo, ho);
  }
  float px[R], py[R], lwx[R], lwy[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t ro = row_off(pitch, i0 + r, j0 - 1) + lo;
    px[r] = *at(uxp, ro);
    py[r] = *at(uyp, ro);
    lwx[r] = loads ? *at(wlx, ro) : 0.f;
    lwy[r] = loads ? *at(wly, ro) : 0.f;
  }
  const float wv = wav[a.step];

  // Row<1>: v[1] = the lane's column, v[2] = the column to its right
  Row<1> xa = make_row(rx[0], lane), ya = make_row(ry[0], lane);
  Row<1> da;
  if constexpr (MODE == kFemDamage) da = make_row(rd[0], lane);
  float lx = xa.v[2] - xa.v[1], ly = ya.v[2] - ya.v[1];   // y-edge of the cell row's upper node row
  float gx = 0.f, gy = 0.f;                                // forces of node row ci from cell row ci - 1

#pragma unroll
  for (int r = 0; r <= R; ++r) {
    const int ci = i0 - 1 + r;      // cell row: nodes ci (row a) and ci + 1 (row b)
    const Row<1> xb = make_row(rx[r + 1], lane), yb = make_row(ry[r + 1], lane);
    Row<1> db;
    if constexpr (MODE == kFemDamage) db = make_row(rd[r + 1], lane);
    const bool even = ((ci + c) & 1) == 0;
    const Corners k = cell_corners<MODE>(Q, even, xa.v[1], xa.v[2], xb.v[1], xb.v[2], ya.v[1], ya.v[2], yb.v[1],
                                         yb.v[2], lx, ly, MODE == kFemDamage ? da.v[1] : 0.f,
                                         MODE == kFemDamage ? da.v[2] : 0.f, MODE == kFemDamage ? db.v[1] : 0.f,
                                         MODE == kFemDamage ? db.v[2] : 0.f);
    // the left cell (ci, c - 1) gives node (ci, c) its n3 and node (ci + 1, c) its n2
    const float l2x = from_left(k.f2x), l2y = from_left(k.f2y);
    const float l3x = from_left(k.f3x), l3y = from_left(k.f3y);
    if (r >= 1) {                   // node row i = ci: cells (ci - 1, .) carried in g, cells (ci, .) here
      const int i = ci;
      const float M = ((i + c) & 1) ? a.m_odd : a.m_even;
      // nodal load: f_ext = ((w(t_n) g) / rho) M  (PF:482-484); g = 0 gives f_ext = 0 exactly
      const float fex = ((wv * lwx[r - 1]) * a.rinv) * M;
      const float fey = ((wv * lwy[r - 1]) * a.rinv) * M;
      const float fx = (gx + k.f0x) + l3x;
      const float fy = (gy + k.f0y) + l3y;
      const float nx_ = fem_update(xa.v[1], px[r - 1], fex, fx, a.dt2, M);
      const float ny_ = fem_update(ya.v[1], py[r - 1], fey, fy, a.dt2, M);
      if (lane >= 1 && i <= a.nx - 2 && c >= 1 && c <= a.ny - 2) {
        const uint32_t ro = row_off(pitch, i, j0 - 1) + lo;
        *at(uxp, ro) = nx_;
        *at(uyp, ro) = ny_;
        if (a.rec_off >= 0) {
          const int64_t fo = a.rec_off + (int64_t)i * a.ny + c;
          frx[fo] = nx_ * a.scale;
          fry[fo] = ny_ * a.scale;
        }
      }
    }
    gx = k.f1x + l2x;
    gy = k.f1y + l2y;
    xa = xb;
    ya = yb;
    if constexpr (MODE == kFemDamage) da = db;
    lx = k.rxd;
    ly = k.ryd;
  }
}

template <int MODE, int R>
__global__ void __launch_bounds__(kBlock) fem_step_kernel(FemStepArgs a) {
  const FemCallDev& c = *a.call;
  gfloat* ws = (gfloat*)c.ws;
  const gfloat* qp = ws + (int64_t)kFemSlots * a.field;
  float Q[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) Q[k] = qp[k];
  const TileRange tr = tile_range(a.n_tiles);
  for (int t = tr.t; t < tr.end; t += tr.step)
    fem_tile<MODE, R>(a, a.tiles[t], ws, (const gfloat*)c.wav, (gfloat*)c.frames_x, (gfloat*)c.frames_y, Q);
}

// --------------------------------------------------------------------------
// Time-blocked step: KS steps per launch.  A block of WB wavefronts (4, or 8
// for RW = 2: a shorter serial chain per wavefront and twice the wavefronts
// in flight) holds a region of WB RW node rows x 64 node columns (wavefront
// w: rows w RW .. w RW + RW - 1, lane l: column l) with u^n and u^{n-1} in
// registers.  Each
// step a wavefront takes the row above and below its band from its
// neighbours through LDS (double-buffered by step parity: one barrier per
// step), the column neighbour and the left cell's corners by DPP with zero
// fill, so values within s nodes of the region edge are wrong after s steps;
// the output tile is the region shrunk by KS nodes on every side.  Nodes
// outside the grid interior are set to zero after every step (the boundary
// ring held at zero).  Every node's arithmetic is fem_tile's, so the frames
// and states equal the one-step kernel's bit for bit.
// --------------------------------------------------------------------------
template <int MODE, int RW, int KS>
__global__ void __launch_bounds__(kWave * blk_waves(RW)) fem_block_kernel(FemBlockArgs a) {
  constexpr int WB = blk_waves(RW);
  const int lb = (int)(blockIdx.x & 7) * a.blocks_per_xcd + (int)(blockIdx.x >> 3);
  // the call and block records load side by side ahead of the exit (a
  // clamped, always valid index), so the field loads wait on one scalar
  // round trip after the kernel arguments: the exit tests a copy of n_blocks
  // passed through an empty asm that takes the loaded values as inputs (not
  // volatile, so later uniform loads stay scalar)
  const FemCallDev& C = *a.call;
  gfloat* ws = (gfloat*)C.ws;
  const FemBlock bl = a.blocks[max(min(lb, a.n_blocks - 1), 0)];
  int n_blocks = a.n_blocks;
  asm("" : "+s"(n_blocks) : "s"((uint64_t)ws), "s"(bl.oi0), "s"(bl.oj0), "s"(bl.has_loads));
  if (lb >= n_blocks) return;   // whole block: no barrier is reached by anyone
  const gfloat* qp = ws + (int64_t)kFemSlots * a.field;
  float Q[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) Q[k] = qp[k];
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & (kWave - 1));
  const int rj0 = bl.oj0 - KS;
  const int c = rj0 + lane;                       // the lane's grid column
  const int r0 = bl.oi0 - KS + w * RW;            // the wavefront's first grid row
  const uint32_t pitch = (uint32_t)a.pitch;
  const bool col_in = c >= 0 && c < a.ny;
  const bool col_int = c >= 1 && c <= a.ny - 2;
  const gfloat* sx_old = ws + (int64_t)(2 * a.src_pair) * a.field;
  const gfloat* sx_new = ws + (int64_t)(2 * a.src_pair + 1) * a.field;
  const gfloat* sy_old = ws + (int64_t)(kFemUySlot + 2 * a.src_pair) * a.field;
  const gfloat* sy_new = ws + (int64_t)(kFemUySlot + 2 * a.src_pair + 1) * a.field;
  const gfloat* dfl = ws + (int64_t)kFemDSlot * a.field;
  const gfloat* wlx = ws + (int64_t)kFemLoadSlot * a.field;
  const gfloat* wly = ws + (int64_t)(kFemLoadSlot + 1) * a.field;
  const bool loads = bl.has_loads != 0;

  // the band: u^{n0} (cx, cy) and u^{n0-1} (ox, oy); outside the grid 0
  float cx[RW], cy[RW], ox[RW], oy[RW], lwx[RW], lwy[RW];
  float dd[MODE == kFemDamage ? RW + 2 : 1];
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int gi = r0 + k;
    const bool in = col_in && gi >= 0 && gi < a.nx;
    const uint32_t o = in ? row_off(pitch, gi, rj0) + (uint32_t)lane : 0u;
    cx[k] = in ? *at(sx_new, o) : 0.f;
    cy[k] = in ? *at(sy_new, o) : 0.f;
    ox[k] = in ? *at(sx_old, o) : 0.f;
    oy[k] = in ? *at(sy_old, o) : 0.f;
    lwx[k] = (in && loads) ? *at(wlx, o) : 0.f;
    lwy[k] = (in && loads) ? *at(wly, o) : 0.f;
  }
  if constexpr (MODE == kFemDamage) {
#pragma unroll
    for (int k = 0; k < RW + 2; ++k) {
      const int gi = r0 - 1 + k;
      const bool in = col_in && gi >= 0 && gi < a.nx;
      dd[k] = in ? *at(dfl, row_off(pitch, gi, rj0) + (uint32_t)lane) : 0.f;
    }
  }
  float ddr[MODE == kFemDamage ? RW + 2 : 1];
  if constexpr (MODE == kFemDamage) {
#pragma unroll
    for (int k = 0; k < RW + 2; ++k) ddr[k] = from_right(dd[k]);
  }

  // edge rows: [buffer][wavefront][top, bottom][x, y][lane]
  __shared__ float edge[2][WB][2][2][kWave];
  edge[0][w][0][0][lane] = cx[0];
  edge[0][w][0][1][lane] = cy[0];
  edge[0][w][1][0][lane] = cx[RW - 1];
  edge[0][w][1][1][lane] = cy[RW - 1];
  __syncthreads();

  gfloat* frx = (gfloat*)C.frames_x;
  gfloat* fry = (gfloat*)C.frames_y;
  const gfloat* wav = (const gfloat*)C.wav;
  // the output tile: region rows KS .. WB RW - KS - 1, lanes KS .. 63 - KS
  const bool lane_out = lane >= KS && lane < kWave - KS && col_int;
  for (int s = 0; s < a.ks; ++s) {
    const int buf = s & 1;
    const int n = a.n0 + s;
    const float wv = wav[n];
    const float ax = w > 0 ? edge[buf][w - 1][1][0][lane] : 0.f;
    const float ay = w > 0 ? edge[buf][w - 1][1][1][lane] : 0.f;
    const float bx = w < WB - 1 ? edge[buf][w + 1][0][0][lane] : 0.f;
    const float by = w < WB - 1 ? edge[buf][w + 1][0][1][lane] : 0.f;
    // cell rows r = 0 .. RW: nodes r0 - 1 + r (row a) and r0 + r (row b)
    float xa0 = ax, ya0 = ay, xa1 = from_right(ax), ya1 = from_right(ay);
    float lx = xa1 - xa0, ly = ya1 - ya0;
    float gx = 0.f, gy = 0.f;
    float nxv[RW], nyv[RW];
#pragma unroll
    for (int r = 0; r <= RW; ++r) {
      const int ci = r0 - 1 + r;
      const float xb0 = r < RW ? cx[r] : bx, yb0 = r < RW ? cy[r] : by;
      const float xb1 = from_right(xb0), yb1 = from_right(yb0);
      const bool even = ((ci + c) & 1) == 0;
      const Corners k = cell_corners<MODE>(Q, even, xa0, xa1, xb0, xb1, ya0, ya1, yb0, yb1, lx, ly,
                                           MODE == kFemDamage ? dd[r] : 0.f, MODE == kFemDamage ? ddr[r] : 0.f,
                                           MODE == kFemDamage ? dd[r + 1] : 0.f,
                                           MODE == kFemDamage ? ddr[r + 1] : 0.f);
      const float l2x = from_left(k.f2x), l2y = from_left(k.f2y);
      const float l3x = from_left(k.f3x), l3y = from_left(k.f3y);
      if (r >= 1) {                 // node row r0 + r - 1 = ci
        const int i = ci;
        const float M = ((i + c) & 1) ? a.m_odd : a.m_even;
        const float fex = ((wv * lwx[r - 1]) * a.rinv) * M;
        const float fey = ((wv * lwy[r - 1]) * a.rinv) * M;
        const float fx = (gx + k.f0x) + l3x;
        const float fy = (gy + k.f0y) + l3y;
        const bool inter = col_int && i >= 1 && i <= a.nx - 2;
        nxv[r - 1] = inter ? fem_update(xa0, ox[r - 1], fex, fx, a.dt2, M) : 0.f;
        nyv[r - 1] = inter ? fem_update(ya0, oy[r - 1], fey, fy, a.dt2, M) : 0.f;
      }
      gx = k.f1x + l2x;
      gy = k.f1y + l2y;
      xa0 = xb0; xa1 = xb1; ya0 = yb0; ya1 = yb1;
      lx = k.rxd;
      ly = k.ryd;
    }
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      ox[k] = cx[k];
      oy[k] = cy[k];
      cx[k] = nxv[k];
      cy[k] = nyv[k];
    }
    if ((n + 1) % a.stride == 0) {  // frame (n + 1) / stride from the output tile
      const int64_t fo0 = (int64_t)((n + 1) / a.stride) * a.frame;
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        const int rr = w * RW + k;
        const int gi = r0 + k;
        if (lane_out && rr >= KS && rr < WB * RW - KS && gi >= 1 && gi <= a.nx - 2) {
          const int64_t fo = fo0 + (int64_t)gi * a.ny + c;
          frx[fo] = cx[k] * a.scale;
          fry[fo] = cy[k] * a.scale;
        }
      }
    }
    edge[buf ^ 1][w][0][0][lane] = cx[0];
    edge[buf ^ 1][w][0][1][lane] = cy[0];
    edge[buf ^ 1][w][1][0][lane] = cx[RW - 1];
    edge[buf ^ 1][w][1][1][lane] = cy[RW - 1];
    __syncthreads();
  }
  // the last two states of the output tile into the other pair (older first)
  gfloat* dx_old = ws + (int64_t)(2 * a.dst_pair) * a.field;
  gfloat* dx_new = ws + (int64_t)(2 * a.dst_pair + 1) * a.field;
  gfloat* dy_old = ws + (int64_t)(kFemUySlot + 2 * a.dst_pair) * a.field;
  gfloat* dy_new = ws + (int64_t)(kFemUySlot + 2 * a.dst_pair + 1) * a.field;
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int rr = w * RW + k;
    const int gi = r0 + k;
    if (lane_out && rr >= KS && rr < WB * RW - KS && gi >= 1 && gi <= a.nx - 2) {
      const uint32_t o = row_off(pitch, gi, rj0) + (uint32_t)lane;
      *at(dx_old, o) = ox[k];
      *at(dx_new, o) = cx[k];
      *at(dy_old, o) = oy[k];
      *at(dy_new, o) = cy[k];
    }
  }
}

// Q from the six constants (FP64, rounded once); one thread.
__global__ void fem_setup_kernel(const FemCallDev* call, int64_t field, double dx, double dy) {
  const FemCallDev& c = *call;
  const double C11 = c.params[0], C22 = c.params[1], C12 = c.params[2];
  const double C16 = c.params[3], C26 = c.params[4], C66 = c.params[5];
  const double rx = dy / (2.0 * dx), ry = dx / (2.0 * dy);
  const double Q[16] = {rx * C11,  rx * C16,  0.5 * C16, 0.5 * C12,
                        rx * C16,  rx * C66,  0.5 * C66, 0.5 * C26,
                        0.5 * C16, 0.5 * C66, ry * C66,  ry * C26,
                        0.5 * C12, 0.5 * C26, ry * C26,  ry * C22};
  float* q = c.ws + (int64_t)kFemSlots * field;
  for (int k = 0; k < 16; ++k) q[k] = (float)Q[k];
}

// Damage factor (nx x ny, row-major) into workspace slot 4 (padded layout).
__global__ void fem_pack_kernel(const FemCallDev* call, int64_t field, int nx, int ny, int pitch) {
  const FemCallDev& c = *call;
  const int64_t n = (int64_t)nx * ny;
  float* dst = c.ws + (int64_t)kFemDSlot * field;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / ny), j = (int)(k % ny);
    dst[(int64_t)(i + 1) * pitch + kLeftPad + j] = c.field[k];
  }
}

// Nodal load weights into workspace slots 5 (x) and 6 (y); one thread per
// load, every (node, component) once (the plan merges duplicates).
__global__ void fem_loads_kernel(const FemCallDev* call, const SrcEv* loads, int n, int64_t field, int pitch) {
  const FemCallDev& c = *call;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const SrcEv s = loads[k];
  c.ws[(int64_t)(kFemLoadSlot + (s.comp ? 1 : 0)) * field + (int64_t)(s.i + 1) * pitch + kLeftPad + s.j] = s.w;
}

// Zero n floats of the workspace (state slots) or of both frame arrays.
__global__ void fem_zero_kernel(const FemCallDev* call, int which, int64_t n) {
  const FemCallDev& c = *call;
  float* p = which == 0 ? c.ws : (which == 1 ? c.frames_x : c.frames_y);
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    p[k] = 0.f;
}

__global__ void fem_set_call_kernel(FemCallDev* dst, FemCallDev v) { *dst = v; }

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
#define ERNN_FEM_SHAPES(X) X(1) X(2) X(4) X(8)

template <int MODE>
hipError_t launch_fem_m(const FemStepArgs& a, int R, int blocks, hipStream_t st) {
#define ERNN_FEM_CASE(r)                                                                         \
  if (R == r) {                                                                                  \
    hipLaunchKernelGGL((fem_step_kernel<MODE, r>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a); \
    return hipGetLastError();                                                                    \
  }
  ERNN_FEM_SHAPES(ERNN_FEM_CASE)
#undef ERNN_FEM_CASE
  return hipErrorInvalidValue;
}

bool fem_shape_ok(int V, int R) {
  if (V != 1) return false;
#define ERNN_FEM_OK(r) if (R == r) return true;
  ERNN_FEM_SHAPES(ERNN_FEM_OK)
#undef ERNN_FEM_OK
  return false;
}

hipError_t launch_fem_step(const FemStepArgs& a, int V, int R, int blocks, hipStream_t st) {
  if (V != 1) return hipErrorInvalidValue;
  return a.mode == kFemDamage ? launch_fem_m<kFemDamage>(a, R, blocks, st)
                              : launch_fem_m<kFemConst>(a, R, blocks, st);
}

#define ERNN_FEM_BLOCK_SHAPES(X) X(2, 2) X(2, 4) X(4, 2) X(4, 4) X(8, 2) X(8, 4) X(8, 8)

template <int MODE>
hipError_t launch_fem_block_m(const FemBlockArgs& a, int grid, hipStream_t st) {
#define ERNN_FEM_BCASE(RW_, KS_)                                                                      \
  if (a.rw == RW_ && a.ksmax == KS_) {                                                                \
    hipLaunchKernelGGL((fem_block_kernel<MODE, RW_, KS_>), dim3((unsigned)grid),                      \
                       dim3(kWave * blk_waves(RW_)), 0, st, a);                               \
    return hipGetLastError();                                                                         \
  }
  ERNN_FEM_BLOCK_SHAPES(ERNN_FEM_BCASE)
#undef ERNN_FEM_BCASE
  return hipErrorInvalidValue;
}

bool fem_block_shape_ok(int RW, int KS) {
#define ERNN_FEM_BOK(RW_, KS_) if (RW == RW_ && KS == KS_) return true;
  ERNN_FEM_BLOCK_SHAPES(ERNN_FEM_BOK)
#undef ERNN_FEM_BOK
  return false;
}

hipError_t launch_fem_block(const FemBlockArgs& a, int grid, hipStream_t st) {
  if (a.ks < 1 || a.ks > a.ksmax) return hipErrorInvalidValue;
  return a.mode == kFemDamage ? launch_fem_block_m<kFemDamage>(a, grid, st)
                              : launch_fem_block_m<kFemConst>(a, grid, st);
}

hipError_t launch_fem_setup(const FemCallDev* call, int64_t field, double dx, double dy, hipStream_t st) {
  hipLaunchKernelGGL(fem_setup_kernel, dim3(1), dim3(1), 0, st, call, field, dx, dy);
  return hipGetLastError();
}

hipError_t launch_fem_pack(const FemCallDev* call, int64_t field, int nx, int ny, int pitch, hipStream_t st) {
  const int64_t n = (int64_t)nx * ny;
  const int64_t blocks = std::min<int64_t>((n + kBlock - 1) / kBlock, 65536);
  hipLaunchKernelGGL(fem_pack_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, call, field, nx, ny, pitch);
  return hipGetLastError();
}

hipError_t launch_fem_loads(const FemCallDev* call, const SrcEv* loads, int n, int64_t field, int pitch,
                            hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fem_loads_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, call, loads, n,
                     field, pitch);
  return hipGetLastError();
}

hipError_t launch_fem_zero(const FemCallDev* call, int which, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + kBlock - 1) / kBlock, 65536);
  hipLaunchKernelGGL(fem_zero_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, call, which, n);
  return hipGetLastError();
}

hipError_t launch_fem_set_call(FemCallDev* dst, const FemCallDev& v, hipStream_t st) {
  hipLaunchKernelGGL(fem_set_call_kernel, dim3(1), dim3(1), 0, st, dst, v);
  return hipGetLastError();
}

}  // namespace ernn
