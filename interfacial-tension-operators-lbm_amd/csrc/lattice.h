// D3Q19 lattice constants and shared parameter blocks for the device kernels.
//
// Ordering and weights follow the reference lattice (LBM_functions.cpp:366-497):
//   i = 0 rest; 1..6 = +x -x +y -y +z -z; 7..18 = (1,1,0) (-1,-1,0) (1,-1,0) (-1,1,0) (1,0,1)
//   (-1,0,-1) (1,0,-1) (-1,0,1) (0,-1,-1) (0,1,1) (0,-1,1) (0,1,-1);  w = 1/3, 1/18, 1/36.
// With nz = 1 the z component of every velocity wraps onto the same plane, so a population
// streams by (cx, cy) only.
#pragma once

#include <cstdint>

namespace itolbm {

constexpr int Q = 19;

// HIP host and device both see these tables (constexpr, folded at compile time).
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1};

constexpr double W0 = 1. / 3.;
constexpr double W1 = 1. / 18.;
constexpr double W2 = 1. / 36.;

// Lattice weights of the FP32 / FP16 forms.  The FP32 collision hands the density deviation
// sum_i s_i on multiplied by the exact sum of its weights; with the nearest floats to 1/3, 1/18,
// 1/36 that sum is 1 + 1.5e-8, which compounds into a drift of rho - 1 of 3e-4 over 20000 steps.
// These weights sum to exactly 1: WL1 is the float nearest 1/18 for which 1 - 12 WL1 is a float,
// WL2 = WL1 / 2, WL0 = 1 - 12 WL1 (relative offsets from 1/3, 1/18, 1/36: 1.2e-7, -6e-8, -6e-8).
constexpr float WL0 = 0x1.555558p-2f;
constexpr float WL1 = 0x1.c71c7p-5f;
constexpr float WL2 = 0x1.c71c7p-6f;

// Stored form of a population in the FP32 / FP16 forms.  The deviation s_i = f_i - w_i is kept
// multiplied by a power of two near 1 / w_i (4, 16, 32 for the rest, axis and diagonal velocities), so
// that encoding and decoding scale exactly, less the case's reference density in the same units:
//   h_i = 2^k s_i - c_k r(e_i),   c_k = 2^k w_k,   r(e) = ref_mid + ref_half e.
// The weights in those units are exact floats (c_0 = 4 WL0, c_1 = 16 WL1 = c_2 = 32 WL2).
constexpr float SCL[3] = {4.f, 16.f, 32.f};
constexpr float ISCL[3] = {0.25f, 0.0625f, 0.03125f};
constexpr float CWL[3] = {WL0 * 4.f, WL1 * 16.f, WL2 * 32.f};
static_assert(CWL[1] == CWL[2], "axis and diagonal classes share the weight unit");

// velocity class of population i: 0 rest, 1 axis, 2 diagonal
__host__ __device__ constexpr int vclass(int i) { return i == 0 ? 0 : (i < 7 ? 1 : 2); }

// z-mirror-symmetric storage of the FP32 / FP16 forms.  With one node in z, velocity i and its z
// mirror (c_z -> -c_z) pull from the same site, and a state that is symmetric under z -> -z has
// f_6 = f_5, f_13 = f_11, f_14 = f_12, f_17 = f_15, f_18 = f_16.  Every state the programs create is
// (gradients and forces have no z component), and the FP32 / FP16 collision keeps the symmetry
// exactly (its z moment is summed over mirror pairs), so 14 of the 19 populations are stored.
constexpr int NSYM = 14;
constexpr int MIRROR_OF[Q] = {0, 1, 2, 3, 4, 5, 5, 7, 8, 9, 10, 11, 12, 11, 12, 15, 16, 15, 16};   // stored partner
constexpr int SLOT_SYM[Q] = {0, 1, 2, 3, 4, 5, 5, 6, 7, 8, 9, 10, 11, 10, 11, 12, 13, 12, 13};     // slot of i
constexpr int VEL_SYM[NSYM] = {0, 1, 2, 3, 4, 5, 7, 8, 9, 10, 11, 12, 15, 16};                     // velocity of a slot

constexpr bool check_mirrors()
{
    for (int i = 0; i < Q; ++i) {
        const int m = MIRROR_OF[i];
        if (VEL_SYM[SLOT_SYM[i]] != m || MIRROR_OF[m] != m) return false;
        if (CX[i] != CX[m] || CY[i] != CY[m] || vclass(i) != vclass(m)) return false;
        if (m != i && CZ[i] != -CZ[m]) return false;
    }
    for (int s = 0; s < NSYM; ++s)
        if (SLOT_SYM[VEL_SYM[s]] != s) return false;
    return true;
}
static_assert(check_mirrors(), "z-mirror table of the D3Q19 ordering");

// stored slots per site and the slot of velocity i: all 19 (full storage) or 14 (z-symmetric)
__host__ __device__ constexpr int n_slots(bool sym) { return sym ? NSYM : Q; }
__host__ __device__ constexpr int slot_of(bool sym, int i) { return sym ? SLOT_SYM[i] : i; }
__host__ __device__ constexpr bool is_stored(bool sym, int i) { return !sym || MIRROR_OF[i] == i; }

// c_k r(e) = mid(k) + half(k) e, formed identically on the host and the device (c_1 = c_2, so the
// axis and diagonal classes share one pair; scalar fields keep it in registers)
struct LpRef {
    float mid0, half0, mid1, half1;
    __host__ __device__ float mid(int k) const { return k == 0 ? mid0 : mid1; }
    __host__ __device__ float half(int k) const { return k == 0 ? half0 : half1; }
};
__host__ __device__ inline LpRef make_lp_ref(double ref_mid, double ref_half)
{
    const float m = (float)ref_mid, h = (float)ref_half;
    return LpRef{CWL[0] * m, CWL[0] * h, CWL[1] * m, CWL[1] * h};
}

// operator ids of Operadores_tensao.cpp:45-49
enum OpId { OP_P1 = 1, OP_P3 = 3, OP_F1 = 11, OP_F2 = 12, OP_F3 = 13 };

// Everything one case's step needs on the device, computed once on the host so that every
// constant is the same double the reference forms at run time.
struct CaseConst {
    int op;            // OpId
    int nx, ny;
    int px;            // FP64 compact form: elements between the slots of a row block (>= nx)
    double one_over_tau;   // 1.0 / tau            (Collision_operators.cpp:24)
    double A;              // fat_R_B              (P amplitude; sigma for F)
    double B0, B1, B2;     // coef_B by |c|^2 = 0, 1, 2   (Operadores_tensao.cpp:94-109)
    double beta;           // recoll
    double hs;             // 0.5 * sigma          (F1 / F2)
    double kf, bf;         // 0.75 sigma xi, 3 sigma / (8 xi)   (F3)
    double sfac0, sfac1, sfac2;   // (1 - 1/(2 tau)) * w * 3   (Guo source, LBM_functions.cpp:1641-1669)
    // reduced-precision forms only: stored deviations are taken from the reference density
    // 1 + ref_mid + ref_half e (e = +1 pure red, -1 pure blue); zero for FP64
    double ref_mid, ref_half;
};

}  // namespace itolbm
