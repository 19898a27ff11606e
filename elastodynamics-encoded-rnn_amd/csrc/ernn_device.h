// Device helpers shared by the HIP kernel files (ernn_kernels.hip,
// ernn_fem.hip): global-address-space views, 32-bit offset addressing,
// register rows with DPP column neighbours, and the XCD-aware tile order
// over the padded field layout of ernn_internal.h.
#pragma once
#include <hip/hip_runtime.h>
#include "ernn_internal.h"

namespace ernn {


// Global-memory views of the call's buffers.  The pointers reach the kernels
// through the device-side argument block, so the compiler would treat them as
// generic (flat) pointers; a flat access counts against both the vector-memory
// and the LDS counters, which makes the first cross-lane shuffle of a tile wait
// for every load in flight.  In the global address space, loads retire in
// order and a row can be consumed while the next rows are still arriving.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) f32x2 gfloat2;
typedef __attribute__((address_space(1))) f32x4 gfloat4;

// Reads of tables that no kernel writes while it runs (wavelets, source and
// receiver events) through the constant address space: uniform ones become
// scalar loads, counted by lgkmcnt.  As vector loads after a store they would
// wait on vmcnt, which on CDNA also counts the stores still in flight.
__device__ __forceinline__ float ld_const(const gfloat* p) {
  return *(const float __attribute__((address_space(4)))*)p;
}
// Four dwords of a 16-byte record, loaded one by one so that they merge into
// one s_load_dwordx4 (an aggregate copy can narrow to three dwords, which
// only a vector load serves).
__device__ __forceinline__ void ld_const4(const void* p, int (&q)[4]) {
  const int __attribute__((address_space(4)))* c = (const int __attribute__((address_space(4)))*)p;
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = c[k];
}
__device__ __forceinline__ SrcEv ld_src(const SrcEv* p) {
  int q[4];
  ld_const4(p, q);
  return SrcEv{q[0], q[1], __int_as_float(q[2]), q[3]};
}
__device__ __forceinline__ RecEv ld_rec(const RecEv* p) {
  int q[4];
  ld_const4(p, q);
  return RecEv{q[0], q[1], q[2], q[3]};
}
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) char gbyte;

// base + off elements, with the byte offset formed in 32 bits (padded fields
// hold fewer than 2^30 elements): the load or store then takes the 64-bit
// base in scalar registers and a 32-bit offset register that every field of
// the tile shares.
template <typename T>
__device__ __forceinline__ T* at(T* base, uint32_t off) {
  return reinterpret_cast<T*>(reinterpret_cast<gbyte*>(base) + (off * (uint32_t)sizeof(float)));
}
template <typename T>
__device__ __forceinline__ const T* at(const T* base, uint32_t off) {
  return reinterpret_cast<const T*>(reinterpret_cast<const gbyte*>(base) + (off * (uint32_t)sizeof(float)));
}
// --------------------------------------------------------------------------
// Register rows
// --------------------------------------------------------------------------
template <int V>
struct Row {
  float v[V + 2];  // v[0]: column j0+lV-1, v[1..V]: own columns, v[V+1]: j0+lV+V
};

template <int V>
__device__ __forceinline__ void load_vec(const gfloat* __restrict__ p, float (&c)[V]) {
  if constexpr (V == 4) {
    const f32x4 t = *reinterpret_cast<const gfloat4*>(p);
    c[0] = t.x; c[1] = t.y; c[2] = t.z; c[3] = t.w;
  } else if constexpr (V == 2) {
    const f32x2 t = *reinterpret_cast<const gfloat2*>(p);
    c[0] = t.x; c[1] = t.y;
  } else {
    c[0] = *p;
  }
}

template <int V>
__device__ __forceinline__ void store_vec(gfloat* __restrict__ p, const float (&c)[V]) {
  if constexpr (V == 4) {
    *reinterpret_cast<gfloat4*>(p) = f32x4{c[0], c[1], c[2], c[3]};
  } else if constexpr (V == 2) {
    *reinterpret_cast<gfloat2*>(p) = f32x2{c[0], c[1]};
  } else {
    *p = c[0];
  }
}

// A row is read in two phases so that a tile can issue every load before
// the first cross-lane shuffle waits for data.  load_raw: ro is the offset of
// column j0 of one padded row; the lane loads its V columns and one more
// value that becomes the halo on the two edge lanes (column j0 - 1 on lane 0,
// j0 + 64 V on lane 63).  The inner lanes re-read their own first column, so
// the load needs no branch; a branch there would keep the compiler from
// hoisting the next rows' loads above this row's shuffles.
// make_row: the column neighbours come from the adjacent lanes by DPP
// whole-wavefront shifts (one VALU move each, no LDS round trip; measured
// 76 vs 78 us per 1024^2 damage-adjoint step against ds_bpermute shuffles).
template <int V>
struct RawRow {
  float c[V];
  float h;
};

template <int V>
__device__ __forceinline__ RawRow<V> load_raw(const gfloat* __restrict__ base, uint32_t ro, uint32_t lo,
                                              uint32_t ho) {
  RawRow<V> r;
  load_vec<V>(at(base, ro + lo), r.c);
  r.h = *at(base, ro + ho);
  return r;
}

template <int V>
__device__ __forceinline__ Row<V> make_row(const RawRow<V>& raw, int lane) {
  Row<V> r;
  // wave_shr:1 / wave_shl:1 (DPP controls 0x138 / 0x130): lane l receives
  // lane l -+ 1's value; the lane shifted in from outside the wavefront keeps
  // `old`, here the halo that lane 0 (left) and lane 63 (right) loaded.
  const int h = __float_as_int(raw.h);
  r.v[0] = __int_as_float(__builtin_amdgcn_update_dpp(h, __float_as_int(raw.c[V - 1]), 0x138, 0xf, 0xf, false));
#pragma unroll
  for (int e = 0; e < V; ++e) r.v[e + 1] = raw.c[e];
  r.v[V + 1] = __int_as_float(__builtin_amdgcn_update_dpp(h, __float_as_int(raw.c[0]), 0x130, 0xf, 0xf, false));
  return r;
}

// make_row for a row without halo values (the time-blocked kernels, whose
// region edge cells leave the valid part after the first step): bound_ctrl
// writes 0 into the lane shifted in from outside the wavefront, so no
// register has to be preset with `old` before each shift.
template <int V>
__device__ __forceinline__ Row<V> make_row_z(const float (&c)[V]) {
  Row<V> r;
  r.v[0] = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(c[V - 1]), 0x138, 0xf, 0xf, true));
#pragma unroll
  for (int e = 0; e < V; ++e) r.v[e + 1] = c[e];
  r.v[V + 1] = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(c[0]), 0x130, 0xf, 0xf, true));
  return r;
}


// Tiles of this wavefront.  Block b runs on XCD b % 8 (dispatch is round
// robin over the 8 XCDs); XCD x owns the contiguous band of tiles
// [x * C, (x + 1) * C), C = ceil(n_tiles / 8), so halo rows and the next
// step's reads of the same band hit that XCD's L2.  Its waves stride through
// the band, so a grid capped below the tile count makes them persistent.
struct TileRange {
  int t, end, step;
};

__device__ __forceinline__ TileRange tile_range(int n_tiles) {
  const int b = blockIdx.x;
  const int per_xcd = gridDim.x >> 3;
  const int x = b & 7;
  const int chunk = (n_tiles + 7) >> 3;
  TileRange r;
  r.t = __builtin_amdgcn_readfirstlane(x * chunk + (b >> 3) * kWavesPerBlock + (int)(threadIdx.x >> 6));
  r.end = min((x + 1) * chunk, n_tiles);
  r.step = per_xcd * kWavesPerBlock;
  return r;
}

// Element offset of column j0 of grid row r in a padded field.  Offsets are
// 32-bit (the plan keeps every padded field below 2^30 elements), so a tile
// addresses each field by one 64-bit base in scalar registers plus one 32-bit
// offset per row that all fields share.
__device__ __forceinline__ uint32_t row_off(uint32_t pitch, int r, int j0) {
  return (uint32_t)(r + 1) * pitch + (uint32_t)(kLeftPad + j0);
}

// Per-lane column offsets: lo = the lane's first column, ho = the value
// load_raw fetches besides (column -1 on lane 0, 64 V on lane 63, else the
// lane's own first column; -1 wraps, and row offsets are >= kLeftPad).
template <int V>
__device__ __forceinline__ uint32_t lane_off(int lane) { return (uint32_t)(lane * V); }
template <int V>
__device__ __forceinline__ uint32_t halo_off(int lane) {
  return (uint32_t)(lane * V - (int)(lane == 0) + V * (int)(lane == kWave - 1));
}

}  // namespace ernn
