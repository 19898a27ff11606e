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
// Work decomposition: one wavefront owns R node rows x 64 V node columns and
// sweeps its cell rows top to bottom from registers; lane l holds V columns
// and gets the column neighbours by DPP wavefront shifts (ernn_device.h), the
// edge lanes load the halo columns.  A cell row's lower corner forces are
// carried in registers to the next node row, so each u value is loaded once
// per tile and no atomics or LDS are used.
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

template <int MODE, int V, int R>
__device__ __forceinline__ void fem_tile(const FemStepArgs& a, const FemTile& tl, gfloat* ws, const gfloat* wav,
                                         gfloat* frx, gfloat* fry, const float (&Q)[16]) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t pitch = (uint32_t)a.pitch;
  const uint32_t lo = lane_off<V>(lane), ho = halo_off<V>(lane);
  const int i0 = tl.i0, j0 = tl.j0;
  const int cur = a.step & 1;
  const gfloat* uxc = ws + (int64_t)cur * a.field;
  const gfloat* uyc = ws + (int64_t)(2 + cur) * a.field;
  gfloat* uxp = ws + (int64_t)(1 - cur) * a.field;
  gfloat* uyp = ws + (int64_t)(3 - cur) * a.field;
  const gfloat* dfl = ws + 4 * a.field;

  // every load of the tile first: R + 2 rows of u^n (and of the damage
  // factor), R rows of u^{n-1}
  RawRow<V> rx[R + 2], ry[R + 2], rd[MODE == kFemDamage ? R + 2 : 1];
#pragma unroll
  for (int r = 0; r < R + 2; ++r) {
    const uint32_t ro = row_off(pitch, i0 - 1 + r, j0);
    rx[r] = load_raw<V>(uxc, ro, lo, ho);
    ry[r] = load_raw<V>(uyc, ro, lo, ho);
    if constexpr (MODE == kFemDamage) rd[r] = load_raw<V>(dfl, ro, lo, ho);
  }
  float px[R][V], py[R][V];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t ro = row_off(pitch, i0 + r, j0) + lo;
    load_vec<V>(at(uxp, ro), px[r]);
    load_vec<V>(at(uyp, ro), py[r]);
  }
  const float wv = wav[a.step];

  const int c0 = j0 + lane * V;     // the lane's first node column
  Row<V> xa = make_row(rx[0], lane), ya = make_row(ry[0], lane);
  Row<V> da;
  if constexpr (MODE == kFemDamage) da = make_row(rd[0], lane);
  float lx[V + 1], ly[V + 1];       // y-edge differences of the upper row of the cell row
#pragma unroll
  for (int k = 0; k <= V; ++k) {
    lx[k] = xa.v[k + 1] - xa.v[k];
    ly[k] = ya.v[k + 1] - ya.v[k];
  }
  float gx[V], gy[V];               // forces of node row ci from cell row ci - 1
#pragma unroll
  for (int e = 0; e < V; ++e) gx[e] = gy[e] = 0.f;

#pragma unroll
  for (int r = 0; r <= R; ++r) {
    const int ci = i0 - 1 + r;      // cell row: nodes ci (row a) and ci + 1 (row b)
    const Row<V> xb = make_row(rx[r + 1], lane), yb = make_row(ry[r + 1], lane);
    Row<V> db;
    if constexpr (MODE == kFemDamage) db = make_row(rd[r + 1], lane);
    float hx[V + 2], hy[V + 2];     // x-edge differences at columns c0 - 1 + k
#pragma unroll
    for (int k = 0; k < V + 2; ++k) {
      hx[k] = xb.v[k] - xa.v[k];
      hy[k] = yb.v[k] - ya.v[k];
    }
    float rxd[V + 1], ryd[V + 1];   // y-edge differences of row b
#pragma unroll
    for (int k = 0; k <= V; ++k) {
      rxd[k] = xb.v[k + 1] - xb.v[k];
      ryd[k] = yb.v[k + 1] - yb.v[k];
    }
    // cells k = 0 .. V (column c0 - 1 + k): corner forces F0 (n0), F1 (n1), F2 (n2), F3 (n3)
    float f0x[V + 1], f0y[V + 1], f1x[V + 1], f1y[V + 1], f2x[V + 1], f2y[V + 1], f3x[V + 1], f3y[V + 1];
#pragma unroll
    for (int k = 0; k <= V; ++k) {
      const bool even = ((ci + c0 - 1 + k) & 1) == 0;
      // triangle 1 holds the bottom x-edge n0 -> n1 and, for even cells, the
      // right y-edge n1 -> n2 ([n0,n1,n2]), for odd cells the left n0 -> n3
      // ([n0,n1,n3]); triangle 2 holds the top x-edge n3 -> n2 and the other y-edge
      const float y1x = even ? rxd[k] : lx[k], y1y = even ? ryd[k] : ly[k];
      const float y2x = even ? lx[k] : rxd[k], y2y = even ? ly[k] : ryd[k];
      Flux t1 = tri_flux(Q, hx[k], hy[k], y1x, y1y);
      Flux t2 = tri_flux(Q, hx[k + 1], hy[k + 1], y2x, y2y);
      if constexpr (MODE == kFemDamage) {
        // C of a triangle = factor at the node nearest its centroid:
        // [n0,n1,n2] -> n1, [n0,n2,n3] -> n3, [n0,n1,n3] -> n0, [n1,n2,n3] -> n2
        t1 = scale_flux(t1, even ? db.v[k] : da.v[k]);
        t2 = scale_flux(t2, even ? da.v[k + 1] : db.v[k + 1]);
      }
      const float lqx = even ? t2.qx : t1.qx, lqy = even ? t2.qy : t1.qy;   // PY of the left-edge triangle
      const float rqx = even ? t1.qx : t2.qx, rqy = even ? t1.qy : t2.qy;   // PY of the right-edge triangle
      f0x[k] = -t1.px - lqx;  f0y[k] = -t1.py - lqy;
      f1x[k] = t1.px - rqx;   f1y[k] = t1.py - rqy;
      f2x[k] = t2.px + rqx;   f2y[k] = t2.py + rqy;
      f3x[k] = lqx - t2.px;   f3y[k] = lqy - t2.py;
    }
    if (r >= 1) {                   // node row i = ci: cells (ci - 1, .) carried in g, cells (ci, .) here
      const int i = ci;
      float fex[V], fey[V];
#pragma unroll
      for (int e = 0; e < V; ++e) fex[e] = fey[e] = 0.f;
      float M[V];
#pragma unroll
      for (int e = 0; e < V; ++e) M[e] = ((i + c0 + e) & 1) ? a.m_odd : a.m_even;
      // nodal loads of this row: f_ext = ((w(t_n) g) / rho) M  (PF:482-484)
      for (int q = tl.ev_begin; q < tl.ev_end; ++q) {
        const SrcEv s = a.loads[q];
        if (s.i != i) continue;
        const int d = s.j - c0;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          if (d == e) {
            const float f = ((wv * s.w) * a.rinv) * M[e];
            if (s.comp) fey[e] += f; else fex[e] += f;
          }
        }
      }
      const uint32_t ro = row_off(pitch, i, j0) + lo;
      const bool row_in = i <= a.nx - 2;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const float fx = (gx[e] + f0x[e + 1]) + f3x[e];
        const float fy = (gy[e] + f0y[e + 1]) + f3y[e];
        const float nx_ = fem_update(xa.v[e + 1], px[r - 1][e], fex[e], fx, a.dt2, M[e]);
        const float ny_ = fem_update(ya.v[e + 1], py[r - 1][e], fey[e], fy, a.dt2, M[e]);
        const int col = c0 + e;
        if (row_in && col >= 1 && col <= a.ny - 2) {
          *at(uxp, ro + e) = nx_;
          *at(uyp, ro + e) = ny_;
          if (a.rec_off >= 0) {
            const int64_t fo = a.rec_off + (int64_t)i * a.ny + col;
            frx[fo] = nx_ * a.scale;
            fry[fo] = ny_ * a.scale;
          }
        }
      }
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
      gx[e] = f1x[e + 1] + f2x[e];
      gy[e] = f1y[e + 1] + f2y[e];
    }
    xa = xb;
    ya = yb;
    if constexpr (MODE == kFemDamage) da = db;
#pragma unroll
    for (int k = 0; k <= V; ++k) {
      lx[k] = rxd[k];
      ly[k] = ryd[k];
    }
  }
}

template <int MODE, int V, int R>
__global__ void __launch_bounds__(kBlock) fem_step_kernel(FemStepArgs a) {
  const FemCallDev& c = *a.call;
  gfloat* ws = (gfloat*)c.ws;
  const gfloat* qp = ws + (int64_t)kFemSlots * a.field;
  float Q[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) Q[k] = qp[k];
  const TileRange tr = tile_range(a.n_tiles);
  for (int t = tr.t; t < tr.end; t += tr.step)
    fem_tile<MODE, V, R>(a, a.tiles[t], ws, (const gfloat*)c.wav, (gfloat*)c.frames_x, (gfloat*)c.frames_y, Q);
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
  float* dst = c.ws + 4 * field;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / ny), j = (int)(k % ny);
    dst[(int64_t)(i + 1) * pitch + kLeftPad + j] = c.field[k];
  }
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
#define ERNN_FEM_SHAPES(X) X(1, 1) X(1, 2) X(1, 4) X(1, 8) X(2, 2) X(2, 4) X(4, 2) X(4, 4)

template <int MODE>
hipError_t launch_fem_m(const FemStepArgs& a, int V, int R, int blocks, hipStream_t st) {
#define ERNN_FEM_CASE(v, r)                                                                         \
  if (V == v && R == r) {                                                                           \
    hipLaunchKernelGGL((fem_step_kernel<MODE, v, r>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a); \
    return hipGetLastError();                                                                       \
  }
  ERNN_FEM_SHAPES(ERNN_FEM_CASE)
#undef ERNN_FEM_CASE
  return hipErrorInvalidValue;
}

bool fem_shape_ok(int V, int R) {
#define ERNN_FEM_OK(v, r) if (V == v && R == r) return true;
  ERNN_FEM_SHAPES(ERNN_FEM_OK)
#undef ERNN_FEM_OK
  return false;
}

hipError_t launch_fem_step(const FemStepArgs& a, int V, int R, int blocks, hipStream_t st) {
  return a.mode == kFemDamage ? launch_fem_m<kFemDamage>(a, V, R, blocks, st)
                              : launch_fem_m<kFemConst>(a, V, R, blocks, st);
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
