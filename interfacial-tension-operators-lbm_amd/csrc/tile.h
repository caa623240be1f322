// Tile geometry shared by the step kernels: one 64-lane wavefront per tile, tiles are vertical
// strips of 64 - 2H owned columns (H halo lanes each side) times `ly` rows of one case.
#pragma once

#include <hip/hip_runtime.h>

#include "step_kernels.h"

namespace itolbm {

__device__ __forceinline__ int wrapi(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

// Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous run of tiles so that
// tiles sharing halo rows/columns also share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return xcd * q + (xcd < r ? xcd : r) + idx;
}

struct Lane {
    int col, xm, xp;      // own column and its periodic x-neighbours
    int lane;
    bool interior;        // owns (stores) its column
};

// Aligned strip (one-row stencils): the 64 lanes own the 64 columns strip*64 .. strip*64+63, so
// every wavefront-wide load of a column row and every store starts on a 64-element boundary.  The
// columns just outside the strip (its halo) are handled by Halo1.
__device__ __forceinline__ Lane make_lane_aligned(int nx, int strip)
{
    Lane L;
    L.lane = threadIdx.x;
    const int xr = strip * 64 + L.lane;
    L.col = wrapi(xr, nx);
    L.xm = (L.col == 0) ? nx - 1 : L.col - 1;
    L.xp = (L.col == nx - 1) ? 0 : L.col + 1;
    L.interior = xr < nx;
    return L;
}

// Bit masks of the velocities with c_x = +1, c_x = -1, c_y = +1, c_y = -1 (D3Q19 ordering).
constexpr unsigned CXP_MASK = (1u << 1) | (1u << 7) | (1u << 9) | (1u << 11) | (1u << 13);
constexpr unsigned CXN_MASK = (1u << 2) | (1u << 8) | (1u << 10) | (1u << 12) | (1u << 14);
constexpr unsigned CYP_MASK = (1u << 3) | (1u << 7) | (1u << 10) | (1u << 16) | (1u << 18);
constexpr unsigned CYN_MASK = (1u << 4) | (1u << 8) | (1u << 9) | (1u << 15) | (1u << 17);

// The two halo columns of an aligned strip, strip*64 - 1 (site 0) and strip*64 + 64 (site 1): lane
// j < 38 pulls population i = j mod 19 of halo site j / 19, i.e. the value at
// (column - c_ix, row - c_iy).  One wavefront load per row fetches all 38.
struct Halo1 {
    int i;        // velocity of this lane's value
    int dy;       // row offset of the pulled value (-c_iy)
    unsigned x;   // column of the pulled value
    bool on;      // lane < 38
};

__device__ __forceinline__ Halo1 make_halo1(int nx, int strip)
{
    Halo1 h;
    const int lane = threadIdx.x;
    h.on = lane < 2 * Q;
    const int site = lane >= Q ? 1 : 0;
    h.i = h.on ? lane - site * Q : 0;   // lanes >= 38: velocity 0 of site 1 (a valid address)
    const int cx = (int)((CXP_MASK >> h.i) & 1u) - (int)((CXN_MASK >> h.i) & 1u);
    const int cy = (int)((CYP_MASK >> h.i) & 1u) - (int)((CYN_MASK >> h.i) & 1u);
    const int hc = strip * 64 + (site ? 64 : -1);
    h.x = (unsigned)wrapi(hc - cx, nx);
    h.dy = -cy;
    return h;
}

// The four halo columns of an aligned strip for two-row stencils: strip*64 - 2, - 1 (sites 0, 1)
// and strip*64 + 64, + 65 (sites 2, 3).  Their 76 pulled values take two wavefront loads per row:
// value g = 19 s + i sits in lane g of load A (g < 64) or lane g - 64 of load B.  Lanes of load B
// past 11 re-read a valid address that no one uses.
struct Halo2 {
    int i[2];        // velocity of this lane's value in loads A, B
    int dy[2];       // -c_iy
    unsigned x[2];   // pulled column
};

__device__ __forceinline__ Halo2 make_halo2(int nx, int strip)
{
    Halo2 h;
    const int lane = threadIdx.x;
#pragma unroll
    for (int slot = 0; slot < 2; ++slot) {
        int g = slot * 64 + lane;
        if (g >= 4 * Q) g = 4 * Q - 1;
        const int site = g / Q, i = g - site * Q;
        const int cx = (int)((CXP_MASK >> i) & 1u) - (int)((CXN_MASK >> i) & 1u);
        const int cy = (int)((CYP_MASK >> i) & 1u) - (int)((CYN_MASK >> i) & 1u);
        const int hc = strip * 64 + (site < 2 ? site - 2 : 64 + site - 2);
        h.i[slot] = i;
        h.dy[slot] = -cy;
        h.x[slot] = (unsigned)wrapi(hc - cx, nx);
    }
    return h;
}

// Row y + dy (dy in {-1, 0, 1}, per lane) wrapped periodically, for any ny >= 1 and any y: the three
// wrapped rows are uniform values (scalar arithmetic) and each lane selects its own.
__device__ __forceinline__ int wrap_row3(int y, int dy, int ny)
{
    const int ym = wrapi(y - 1, ny), yc = wrapi(y, ny), yp = wrapi(y + 1, ny);
    return dy < 0 ? ym : (dy > 0 ? yp : yc);
}

// ------------------------------------------------------------------ buffer-resource access
// A buffer resource holds a uniform base in scalar registers; an access gives the lane's byte offset
// (VGPR) and a uniform byte offset (SGPR), so no per-lane 64-bit address arithmetic is needed.  The
// resource's extent bounds the accesses: one outside it returns zero / is dropped and cannot reach
// other memory.  Offsets are 32-bit: one resource spans at most 4 GiB.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int RSRC_WORD3 = 0x00020000;   // raw 32-bit data format (CDNA)
constexpr int AUX_NT = 2;                // nontemporal store: not read again by the launch

template <typename T>
__device__ __forceinline__ rsrc_t make_rsrc(__attribute__((address_space(1))) const T* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)(unsigned long long)base, (short)0, (int)bytes, RSRC_WORD3);
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
template <typename R> __device__ R buf_load(rsrc_t r, unsigned voff, unsigned soff);
template <typename R> __device__ void buf_store(R v, rsrc_t r, unsigned voff, unsigned soff);
template <typename R>
__device__ __forceinline__ R buf_load64(rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <typename R>
__device__ __forceinline__ void buf_store64(R v, rsrc_t r, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, AUX_NT);
}
template <> __device__ __forceinline__ double buf_load<double>(rsrc_t r, unsigned voff, unsigned soff)
{
    return buf_load64<double>(r, voff, soff);
}
template <> __device__ __forceinline__ void buf_store<double>(double v, rsrc_t r, unsigned voff, unsigned soff)
{
    buf_store64(v, r, voff, soff);
}
template <> __device__ __forceinline__ unsigned buf_load<unsigned>(rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
template <> __device__ __forceinline__ void buf_store<unsigned>(unsigned v, rsrc_t r, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, AUX_NT);
}

// Value of lane `src` (a compile-time constant after unrolling) in every lane.
__device__ __forceinline__ float lane_value(float v, int src)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
}
__device__ __forceinline__ unsigned lane_value(unsigned v, int src)
{
    return (unsigned)__builtin_amdgcn_readlane((int)v, src);
}
__device__ __forceinline__ double lane_value(double v, int src)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The value of lane - 1 (from_left) or lane + 1 (from_right) of the wavefront, and `edge` in lane 0
// (from_left) or lane 63 (from_right), whose neighbour lies outside: one whole-wavefront DPP move
// (wave_shr:1 / wave_shl:1; the lane without a source keeps the `old` operand, bound_ctrl off).
__device__ __forceinline__ float from_left(float v, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float v, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ double dpp_pair(double v, double edge, int ctrl_left)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long e = __builtin_bit_cast(unsigned long long, edge);
    unsigned lo, hi;
    if (ctrl_left) {
        lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)e, (int)(unsigned)b, 0x138, 0xF, 0xF, false);
        hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(e >> 32), (int)(unsigned)(b >> 32), 0x138, 0xF, 0xF, false);
    } else {
        lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)e, (int)(unsigned)b, 0x130, 0xF, 0xF, false);
        hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(e >> 32), (int)(unsigned)(b >> 32), 0x130, 0xF, 0xF, false);
    }
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double from_left(double v, double edge) { return dpp_pair(v, edge, 1); }
__device__ __forceinline__ double from_right(double v, double edge) { return dpp_pair(v, edge, 0); }

}  // namespace itolbm
