// Host-visible declarations of the step kernels' launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include "lattice.h"

namespace itolbm {

// kernel families: one launch per family advances every case of that family by one step
// KIND_GEO: cases with solid sites (any operator), stepped by the site-parallel kernels
enum { KIND_P = 0, KIND_F3 = 1, KIND_F1 = 2, KIND_F2 = 3, KIND_GEO = 4, N_KINDS = 5 };

// Device-resident entry of one case in a launch's case table.
struct CaseDev {
    CaseConst cc;
    void* pop[2];        // post-collision state buffers: pop[parity] is read, pop[parity ^ 1] written
    double* aux;         // [6][S]: mediator momentum gm xyz, force F xyz (diagnostic steps), or nullptr
    long long S;         // population stride (elements)
    int ly;              // rows per wavefront segment
    int nstrip, nseg;    // tiles of the case: strips of 64 - 2H columns x segments of ly rows
    int tile0;           // first tile of the case in the launch
    int kind;            // kernel family of the case (KIND_*)
};

struct StepLaunch {
    const CaseDev* tab;      // case table
    const int* tile_case;    // tile -> case index
    int parity;              // which of pop[0] / pop[1] holds the input state
    int record;              // write aux for cases that have it
    int depth;               // FP32/FP16 kernels: rows in flight ahead (0 = default)
    int ntiles;              // tiles of the launch (read by the one-launch FP64 kernel)
    int compact;             // FP64: the P1/P3 and F3 cases hold the compact form (below)
};

// Compact FP64 form (P1/P3 and F3 cases): row blocks [y][CSLOTS][px] of doubles (px >= nx: the slot
// pitch, CaseConst::px).  A collided site
// stores its 19 post-collision totals f_i in slots 0..18 and the scalars its recolouring was formed
// from in slots CS_CR (colour fraction cR), CS_RHO (density), CS_G .. CS_G + 2 (mediator gradient);
// slots 24..38 are unused.  A raw site (one the collision left untouched, or an uploaded state) holds
// f_R,i in slot i, f_B,i in slot cs_raw_b(i) and CS_RAW in slot CS_RHO (a collided site's density is
// > 1e-30 or NaN, never <= 1e-30).
constexpr int CSLOTS = 39;
constexpr int CS_CR = 19, CS_RHO = 20, CS_G = 21;
constexpr double CS_RAW = -1.0;
__host__ __device__ constexpr int cs_raw_b(int i) { return i == 0 ? 19 : 20 + i; }
// doubles per site of an FP64 state buffer (either form)
constexpr int F64_SLOTS = CSLOTS;

int step_halo(int op);
// kernel family of an operator (OpId)
inline int step_kind(int op)
{
    switch (op) {
    case OP_P1:
    case OP_P3: return KIND_P;
    case OP_F3: return KIND_F3;
    case OP_F1: return KIND_F1;
    default: return KIND_F2;
    }
}

// Static scratch of the allocation-free step (static_step.hip): a table entry slot, then the
// tile -> case map of at most SCRATCH_MAX_TILES tiles.
constexpr int SCRATCH_MAX_TILES = 1 << 20;
constexpr size_t SCRATCH_TABLE_BYTES = 256;
constexpr size_t SCRATCH_BYTES = SCRATCH_TABLE_BYTES + sizeof(int) * (size_t)SCRATCH_MAX_TILES;
inline CaseDev* scratch_table(void* base) { return (CaseDev*)base; }
inline int* scratch_tiles(void* base) { return (int*)((char*)base + SCRATCH_TABLE_BYTES); }
// device address of the scratch array on `device`, which must be current: looked up once per
// device with hipGetSymbolAddress, then kept
hipError_t scratch_setup(int device, void** out);
// writes the table entry d and a tile map of ntiles tiles of case 0 into the scratch at `base`
hipError_t launch_fill_static_table(void* base, const CaseDev& d, int ntiles, hipStream_t stream);

// plain streaming copy of `bytes` bytes (16-byte aligned pointers), stream_copy.hip
hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

hipError_t launch_step_f64(int kind, const StepLaunch& a, int ntiles, hipStream_t stream);
// every family of an FP64 table in one launch of `grid` blocks (a.ntiles tiles; grid = a.ntiles: a
// wavefront per tile; grid = step_f64_fused_capacity(): persistent)
hipError_t launch_step_f64_fused(const StepLaunch& a, int grid, hipStream_t stream);
// blocks of step_fused resident on the current device at once, a multiple of 8 (0 if unknown), for
// P1/P3 and F3 cases in the compact or the full form
int step_f64_fused_capacity(bool compact);

// One FP64 step of one case with solid sites (step_f64.hip): `in` holds the post-collision state at
// its own site, `out` receives the next one; phi [S], g [3][S] and nt [6][S] are the case's field
// arrays; aux [6][S] (gm, F) is written when not null.
struct GeoLaunch {
    CaseConst cc;
    const double* in;
    double* out;
    const unsigned char* solid;   // nx * ny sites, 1 = solid
    double wett;                  // the wall's phase for the mediators (wett_R)
    double* phi;
    double* g;
    double* nt;
    double* aux;
    long long S;
};
hipError_t launch_step_geo(int kind, const GeoLaunch& a, hipStream_t stream);

// One case's FP64 state from the full form to the compact one (every site raw) or back, from `in` to
// `out` (the case's other state buffer); the 38 populations of every site are unchanged
hipError_t launch_f64_convert(const double* in, double* out, bool to_compact, double beta, int nx, int px, long long n,
                              hipStream_t stream);

// FP64 state in row blocks [y][colour][19][nx], or in the compact form [y][CSLOTS][px] (recolouring
// factor beta, slot pitch px), -> [colour][19][S]: every site's post-collision populations at the site
hipError_t launch_f64_rows_to_soa(const double* pop, bool compact, double beta, int nx, int px, long long n,
                                  long long S, double* soa, hipStream_t stream);

// FP32-compute forms; storage FP32 (elem = 4) or FP16 (elem = 2): per population the pair
// (s_i / w_i, e_i), s_i = fR_i + fB_i - w_i, e_i = (fR_i - fB_i) / (fR_i + fB_i), at ((y * NS + slot) px + x)
// with NS = 19 slots, or NS = 14 for the z-symmetric storage (sym; slot_of in lattice.h)
hipError_t launch_step_lp(int elem, bool sym, int kind, const StepLaunch& a, int ntiles, hipStream_t stream);

// low-precision state -> FP64 colour populations fR, fB [19][S]
hipError_t launch_lp_to_f64(int elem, bool sym, const void* pop, int nx, int px, long long S, long long n, double rmid,
                            double rhalf, double* fR, double* fB, hipStream_t stream);
// drop initial condition written on the device in a reduced-precision form
hipError_t launch_lp_init_drop(int elem, bool sym, void* pop, int nx, int px, int ny, double R, double kp, int x0, int y0,
                               double v0, double rmid, double rhalf, hipStream_t stream);
// in-place change of the reference density deviation r(e) = mid + half e of a reduced-precision state
hipError_t launch_lp_recentre(int elem, bool sym, void* pop, long long npairs, int px, double mid0, double half0,
                              double mid1, double half1, hipStream_t stream);

}  // namespace itolbm
