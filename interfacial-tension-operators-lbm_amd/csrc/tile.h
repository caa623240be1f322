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

__device__ __forceinline__ Lane make_lane(int nx, int strip, int H)
{
    Lane L;
    L.lane = threadIdx.x;
    const int WI = 64 - 2 * H;
    const int xr = strip * WI + L.lane - H;
    L.col = wrapi(xr, nx);
    L.xm = (L.col == 0) ? nx - 1 : L.col - 1;
    L.xp = (L.col == nx - 1) ? 0 : L.col + 1;
    L.interior = (L.lane >= H) && (L.lane < 64 - H) && (xr < nx);
    return L;
}

}  // namespace itolbm
