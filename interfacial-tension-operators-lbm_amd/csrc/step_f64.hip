// FP64 fused time step: pull-stream + colour field + mediator gradient + tension (P or F class)
// + BGK (Guo forcing) + Latva-Kokko-Rothman recolouring, one kernel launch per step.
//
// Reference path: Imm_OPS.cpp:246-372 (mediator pass :252-267, class-F force :271, collide +
// push-stream :275-287), with the per-site functions colisao_comum (Operadores_tensao.cpp:340-373),
// coll_BGK (Collision_operators.cpp:11-49), calcula / dist_eq / source / recolloring
// (LBM_functions.cpp:1099-1171, 889-945, 1602-1673, 1712-1746), perturbacao / coef_B
// (Operadores_tensao.cpp:94-133) and calcula_forca (Operadores_tensao.cpp:202-332).
//
// Layout: row blocks [y][colour][19][nx] (below), holding the POST-COLLISION values of the previous
// step at their own site.  The reference's push
// (f_new[x + c_i] = f*(x)) is done here as a pull: site x reads f*_i(x - c_i).
//
// Work decomposition: one 64-lane wavefront owns a vertical strip of 64 consecutive columns, of
// which the H edge lanes on each side are halo (H = 1 for P1/P3/F3, 2 for F1/F2), and marches down
// `ly` rows.  Each lane pulls the 38 populations of its own column once per row; the colour field of
// the rows in flight stays in registers and x-neighbours come from wavefront shuffles, so each
// population is read from HBM once (halo columns/rows are re-read through L2) and written once.
//
// Arithmetic: every double expression keeps the reference's order of evaluation, so with
// -ffp-contract=off the results equal the reference's bit for bit.  Terms whose lattice
// coefficient is 0 are left out of the running sums: the reference adds them to sums that start at
// +0.0, where adding a signed zero never changes the value.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lattice.h"
#include "step_kernels.h"
#include "tile.h"

namespace itolbm {

namespace {

constexpr double SQRT2 = 1.4142135623730951;  // sqrt(2.0), correctly rounded

// Pointers into device global memory.  Table entries hold generic pointers; casting them to the
// global address space lets the compiler emit global (not flat) loads and stores with a uniform
// 64-bit base and a 32-bit per-lane offset.
using gcdouble = __attribute__((address_space(1))) const double;
using gdouble = __attribute__((address_space(1))) double;

// State layout: row blocks [y][colour][19][nx] (colour 0 red, 1 blue), so the 38 velocity rows of one
// y are contiguous and one buffer resource (base in scalar registers) covers them; an access gives the
// lane's byte offset x * 8 and the velocity row's byte offset as a scalar.
__device__ __forceinline__ long long block_doubles(int nx) { return 2LL * Q * nx; }
__device__ __forceinline__ unsigned block_bytes(int nx) { return (unsigned)(block_doubles(nx) * (long long)sizeof(double)); }
__device__ __forceinline__ unsigned vrow_off(int colour, int i, int nx)
{
    return (unsigned)((colour * Q + i) * nx) * (unsigned)sizeof(double);
}

// Pull the 38 populations of site (col, y): f_i(x) = f*_i(x - c_i).
__device__ __forceinline__ void load_row(gcdouble* __restrict__ in, int nx, int ny, int y, const Lane& L,
                                         double fR[Q], double fB[Q], bool live = true)
{
    const long long rb = block_doubles(nx);
    const unsigned bytes = live ? block_bytes(nx) : 0u;
    const rsrc_t rm = make_rsrc(in + wrapi(y - 1, ny) * rb, bytes);
    const rsrc_t r0 = make_rsrc(in + wrapi(y, ny) * rb, bytes);
    const rsrc_t rp = make_rsrc(in + wrapi(y + 1, ny) * rb, bytes);
    const unsigned om = (unsigned)L.xm * 8u, o0 = (unsigned)L.col * 8u, op = (unsigned)L.xp * 8u;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const unsigned xs = CX[i] > 0 ? om : (CX[i] < 0 ? op : o0);
        const rsrc_t r = CY[i] > 0 ? rm : (CY[i] < 0 ? rp : r0);
        fR[i] = buf_load<double>(r, xs, vrow_off(0, i, nx));
        fB[i] = buf_load<double>(r, xs, vrow_off(1, i, nx));
    }
}

// element index of population (colour, i) of site (x, y) in the row-block layout
__device__ __forceinline__ unsigned long long elem_index(int nx, int y, int colour, int i, unsigned x)
{
    return (unsigned long long)y * (unsigned long long)block_doubles(nx) + (unsigned)((colour * Q + i) * nx) + x;
}

// density (LBM_functions.cpp:1581-1590): sequential sum
__device__ __forceinline__ double dens(const double f[Q])
{
    double r = f[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) r = r + f[i];
    return r;
}

// c_i . a with the reference's left-to-right order, zero-coefficient terms left out
template <int i>
__device__ __forceinline__ double cdot(double a0, double a1, double a2)
{
    constexpr int c0 = CX[i], c1 = CY[i], c2 = CZ[i];
    double s = 0.0;
    bool first = true;
    if constexpr (c0 != 0) { s = (c0 > 0) ? a0 : -a0; first = false; }
    if constexpr (c1 != 0) { const double t = (c1 > 0) ? a1 : -a1; s = first ? t : s + t; first = false; }
    if constexpr (c2 != 0) { const double t = (c2 > 0) ? a2 : -a2; s = first ? t : s + t; first = false; }
    return s;
}

// c_i . a for a direction i known only at run time (uniform across the wavefront): the same operations
// as cdot<i>, the zero components left out by uniform branches
constexpr unsigned CZP_MASK = (1u << 5) | (1u << 11) | (1u << 14) | (1u << 16) | (1u << 17);
constexpr unsigned CZN_MASK = (1u << 6) | (1u << 12) | (1u << 13) | (1u << 15) | (1u << 18);
__device__ __forceinline__ int cx_of(int i) { return (int)((CXP_MASK >> i) & 1u) - (int)((CXN_MASK >> i) & 1u); }
__device__ __forceinline__ int cy_of(int i) { return (int)((CYP_MASK >> i) & 1u) - (int)((CYN_MASK >> i) & 1u); }
__device__ __forceinline__ int cz_of(int i) { return (int)((CZP_MASK >> i) & 1u) - (int)((CZN_MASK >> i) & 1u); }
__device__ __forceinline__ double cdot_rt(int i, double a0, double a1, double a2)
{
    const int c0 = cx_of(i), c1 = cy_of(i), c2 = cz_of(i);
    double s = 0.0;
    bool first = true;
    if (c0 != 0) { s = (c0 > 0) ? a0 : -a0; first = false; }
    if (c1 != 0) { const double t = (c1 > 0) ? a1 : -a1; s = first ? t : s + t; first = false; }
    if (c2 != 0) { const double t = (c2 > 0) ? a2 : -a2; s = first ? t : s + t; first = false; }
    return s;
}

// Nine values of a field around a lane: v[dy + 1][dx + 1] = field(x + dx, y + dy).
struct Nine {
    double v[3][3];
    __device__ __forceinline__ double at(int dx, int dy) const { return v[dy + 1][dx + 1]; }
};

// 3 x 3 neighbourhood of each lane's site from the lanes of rows y-1, y, y+1 (a row per lane), with the
// columns left of lane 0 and right of lane 63 taken from the strip's halo values (hl, hr per row), for
// aligned 64-column strips.
__device__ __forceinline__ Nine gather9h(const double r[3], const double hl[3], const double hr[3])
{
    Nine n;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        n.v[k][0] = from_left(r[k], hl[k]);
        n.v[k][1] = r[k];
        n.v[k][2] = from_right(r[k], hr[k]);
    }
    return n;
}

// Mediator momentum gm = sum_i c_i M_i, M_i(x) = w_i (3 phi(x - c_i)):
// emite_mediadores + propag_site_med + momentum (LBM_functions.cpp:1286-1337, 1056-1087).
__device__ __forceinline__ void mediator_gradient(const Nine& phi, double& gx, double& gy, double& gz)
{
    auto t = [&](int dx, int dy) { return phi.at(dx, dy) * 3.0; };
    const double M1 = W1 * t(-1, 0), M2 = W1 * t(1, 0), M3 = W1 * t(0, -1), M4 = W1 * t(0, 1);
    const double M5 = W1 * t(0, 0);
    const double M7 = W2 * t(-1, -1), M8 = W2 * t(1, 1), M9 = W2 * t(-1, 1), M10 = W2 * t(1, -1);
    const double M11 = W2 * t(-1, 0), M12 = W2 * t(1, 0), M15 = W2 * t(0, 1), M16 = W2 * t(0, -1);
    // M13 = M11, M14 = M12, M17 = M15, M18 = M16, M6 = M5
    gx = M1 - M2;
    gx = gx + M7; gx = gx - M8; gx = gx + M9; gx = gx - M10;
    gx = gx + M11; gx = gx - M12; gx = gx + M11; gx = gx - M12;
    gy = M3 - M4;
    gy = gy + M7; gy = gy - M8; gy = gy - M9; gy = gy + M10;
    gy = gy - M15; gy = gy + M16; gy = gy - M15; gy = gy + M16;
    gz = M5 - M5;
    gz = gz + M11; gz = gz - M12; gz = gz - M11; gz = gz + M12;
    gz = gz - M15; gz = gz + M16; gz = gz + M15; gz = gz - M16;
}

// Isotropic gradient of the phase field, g = 3 sum_{i>=1} w_i c_i phi(x + c_i)
// (calcula_forca, Operadores_tensao.cpp:219-235).
__device__ __forceinline__ void phase_gradient(const Nine& p, double& gx, double& gy, double& gz)
{
    const double ph1 = W1 * p.at(1, 0), ph2 = W1 * p.at(-1, 0), ph3 = W1 * p.at(0, 1), ph4 = W1 * p.at(0, -1);
    const double ph5 = W1 * p.at(0, 0);
    const double ph7 = W2 * p.at(1, 1), ph8 = W2 * p.at(-1, -1), ph9 = W2 * p.at(1, -1), ph10 = W2 * p.at(-1, 1);
    const double ph11 = W2 * p.at(1, 0), ph12 = W2 * p.at(-1, 0), ph15 = W2 * p.at(0, -1), ph16 = W2 * p.at(0, 1);
    double x = 0.0, y = 0.0, z = 0.0;
    x = x + ph1; x = x - ph2; x = x + ph7; x = x - ph8; x = x + ph9; x = x - ph10;
    x = x + ph11; x = x - ph12; x = x + ph11; x = x - ph12;
    y = y + ph3; y = y - ph4; y = y + ph7; y = y - ph8; y = y - ph9; y = y + ph10;
    y = y - ph15; y = y + ph16; y = y - ph15; y = y + ph16;
    z = z + ph5; z = z - ph5; z = z + ph11; z = z - ph12; z = z - ph11; z = z + ph12;
    z = z - ph15; z = z + ph16; z = z + ph15; z = z - ph16;
    gx = 3.0 * x; gy = 3.0 * y; gz = 3.0 * z;
}

// Per-site collision, colisao_comum (Operadores_tensao.cpp:340-373), in two stages so that the
// part that does not depend on the colour gradient (sums, BGK) can run while the next row's loads
// are in flight.
struct Site {
    double f[Q];
    double rho, cR, cB;    // colour sums of the pulled populations
    double rho2;           // calcula's density of f = fR + fB
    double mx, my, mz;     // calcula's momentum of f
    bool skip;             // rho <= 1e-30: the reference leaves the site untouched
};

// f = fR + fB, colour fractions and calcula's sums (LBM_functions.cpp:1136-1169)
__device__ __forceinline__ void site_sums(Site& s, const double fR[Q], const double fB[Q], double rR, double rB)
{
    s.rho = rR + rB;
    s.skip = (s.rho <= 1.0e-30);
    s.cR = rR / s.rho;
    s.cB = 1.0 - s.cR;
#pragma unroll
    for (int i = 0; i < Q; ++i) s.f[i] = fR[i] + fB[i];
    const double* f = s.f;
    double rho2 = f[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) rho2 = rho2 + f[i];
    double mx = f[1] - f[2];
    mx = mx + f[7]; mx = mx - f[8]; mx = mx + f[9]; mx = mx - f[10];
    mx = mx + f[11]; mx = mx - f[12]; mx = mx + f[13]; mx = mx - f[14];
    double my = f[3] - f[4];
    my = my + f[7]; my = my - f[8]; my = my - f[9]; my = my + f[10];
    my = my - f[15]; my = my + f[16]; my = my - f[17]; my = my + f[18];
    double mz = f[5] - f[6];
    mz = mz + f[11]; mz = mz - f[12]; mz = mz - f[13]; mz = mz + f[14];
    mz = mz - f[15]; mz = mz + f[16]; mz = mz + f[17]; mz = mz - f[18];
    s.rho2 = rho2; s.mx = mx; s.my = my; s.mz = mz;
}

// coll_BGK (Collision_operators.cpp:11-49) on s.f; FORCED adds Guo's source with force F
template <bool FORCED>
__device__ __forceinline__ void bgk(const CaseConst& cc, Site& s, double Fx, double Fy, double Fz)
{
    double* f = s.f;
    const double rho2 = s.rho2;
    double vx = 0., vy = 0., vz = 0.;
    if (rho2 != 0.0) {
        const double inv = 1.0 / rho2;
        vx = s.mx * inv; vy = s.my * inv; vz = s.mz * inv;
    }
    double Sx = 0.0, Sy = 0.0, Sz = 0.0, vF = 0.0;
    if constexpr (FORCED) {
        const double ax = Fx / s.rho, ay = Fy / s.rho, az = Fz / s.rho;
        vx = vx + 0.5 * ax; vy = vy + 0.5 * ay; vz = vz + 0.5 * az;
        Sx = ax * rho2; Sy = ay * rho2; Sz = az * rho2;   // F' = a rho'
        vF = 0.0 + Sx * vx;
        vF = vF + Sy * vy;
        vF = vF + Sz * vz;
    }
    const double vquad = vx * vx + vy * vy + vz * vz;
    const double K = 0.5 * vquad * 3.0;
    const double wr0 = W0 * rho2, wr1 = W1 * rho2, wr2 = W2 * rho2;
    const double oot = cc.one_over_tau;
    // dist_eq (LBM_functions.cpp:924-944) and the relaxation, pair by pair (c_o = -c_i)
    f[0] = f[0] + (wr0 * ((1.0 + 0.0) - K) - f[0]) * oot;
    auto pair = [&](int i, int o, double wr, double cv) {
        const double lin = cv * 3.0;
        const double qd = 0.5 * cv * cv * 3.0 * 3.0;
        const double ei = wr * (1.0 + lin + qd - K);
        const double eo = wr * (1.0 - lin + qd - K);
        f[i] = f[i] + (ei - f[i]) * oot;
        f[o] = f[o] + (eo - f[o]) * oot;
    };
    pair(1, 2, wr1, vx);
    pair(3, 4, wr1, vy);
    pair(5, 6, wr1, vz);
    pair(7, 8, wr2, vx + vy);
    pair(9, 10, wr2, vx - vy);
    pair(11, 12, wr2, vx + vz);
    pair(13, 14, wr2, vx - vz);
    pair(16, 15, wr2, vy + vz);
    pair(18, 17, wr2, vy - vz);
    if constexpr (FORCED) {
        // source (LBM_functions.cpp:1641-1669): S_i = sfac_i ((cF + 3 cF cv) - vF), added after the relaxation
        f[0] = f[0] + cc.sfac0 * (0.0 - vF);
#define ITOLBM_SRC(i, sf)                                                   \
    {                                                                       \
        const double cF = cdot<i>(Sx, Sy, Sz);                              \
        const double cv = cdot<i>(vx, vy, vz);                              \
        f[i] = f[i] + sf * (cF + 3.0 * cF * cv - vF);                       \
    }
        ITOLBM_SRC(1, cc.sfac1) ITOLBM_SRC(2, cc.sfac1) ITOLBM_SRC(3, cc.sfac1) ITOLBM_SRC(4, cc.sfac1)
        ITOLBM_SRC(5, cc.sfac1) ITOLBM_SRC(6, cc.sfac1) ITOLBM_SRC(7, cc.sfac2) ITOLBM_SRC(8, cc.sfac2)
        ITOLBM_SRC(9, cc.sfac2) ITOLBM_SRC(10, cc.sfac2) ITOLBM_SRC(11, cc.sfac2) ITOLBM_SRC(12, cc.sfac2)
        ITOLBM_SRC(13, cc.sfac2) ITOLBM_SRC(14, cc.sfac2) ITOLBM_SRC(15, cc.sfac2) ITOLBM_SRC(16, cc.sfac2)
        ITOLBM_SRC(17, cc.sfac2) ITOLBM_SRC(18, cc.sfac2)
#undef ITOLBM_SRC
    }
}

// perturbacao (Operadores_tensao.cpp:116-133): f_i += A|g| (w_i (c.g)^2/|g|^2 - B_i), g = gm
__device__ __forceinline__ void perturb(const CaseConst& cc, double f[Q], double gx, double gy, double gz,
                                        double mod2, double mod)
{
    if (mod2 <= 1.0e-60 || cc.A == 0.0) return;
    const double Am = cc.A * mod;
    f[0] += Am * (0.0 - cc.B0);
    auto pert = [&](int i, int o, double w, double B, double cg) {
        const double t = Am * (w * cg * cg / mod2 - B);
        f[i] += t;
        f[o] += t;
    };
    pert(1, 2, W1, cc.B1, gx);
    pert(3, 4, W1, cc.B1, gy);
    pert(5, 6, W1, cc.B1, gz);
    pert(7, 8, W2, cc.B2, gx + gy);
    pert(9, 10, W2, cc.B2, gx - gy);
    pert(11, 12, W2, cc.B2, gx + gz);
    pert(13, 14, W2, cc.B2, gx - gz);
    pert(16, 15, W2, cc.B2, gy + gz);
    pert(18, 17, W2, cc.B2, gy - gz);
}

// recolloring (LBM_functions.cpp:1712-1746) of s.f with the mediator gradient gm; each pair of
// colour populations goes to out(i, fR_i, fB_i) as soon as it is formed
template <class Out>
__device__ __forceinline__ void recolour(const CaseConst& cc, const Site& s, double gx, double gy, double gz,
                                         double mod, Out out)
{
    const double* f = s.f;
    const double cR = s.cR, cB = s.cB;
    const double fator = cc.beta * cR * cB;
    out(0, cR * f[0], cB * f[0]);
    const double fw1 = fator * W1 * s.rho, fw2 = fator * W2 * s.rho;
    const double mod_d = mod * SQRT2;
    auto recol = [&](int i, int o, double fw, double prod, double den) {
        const double cosp = (mod != 0.0) ? prod / den : 0.0;
        const double t = fw * cosp;
        out(i, cR * f[i] - t, cB * f[i] + t);
        out(o, cR * f[o] - (-t), cB * f[o] + (-t));
    };
    recol(1, 2, fw1, gx, mod);
    recol(3, 4, fw1, gy, mod);
    recol(5, 6, fw1, gz, mod);
    recol(7, 8, fw2, gx + gy, mod_d);
    recol(9, 10, fw2, gx - gy, mod_d);
    recol(11, 12, fw2, gx + gz, mod_d);
    recol(13, 14, fw2, gx - gz, mod_d);
    recol(16, 15, fw2, gy + gz, mod_d);
    recol(18, 17, fw2, gy - gz, mod_d);
}

// nontemporal stores of a site's 38 post-collision populations into row block `ro`
__device__ __forceinline__ void store_pop(rsrc_t ro, int nx, unsigned xoff, int i, double r, double b)
{
    buf_store<double>(r, ro, xoff, vrow_off(0, i, nx));
    buf_store<double>(b, ro, xoff, vrow_off(1, i, nx));
}
__device__ __forceinline__ void store_row(rsrc_t ro, int nx, unsigned xoff, const double fR[Q], const double fB[Q])
{
#pragma unroll
    for (int i = 0; i < Q; ++i) store_pop(ro, nx, xoff, i, fR[i], fB[i]);
}

__device__ __forceinline__ void store_aux(gdouble* __restrict__ aux, long long S, unsigned site, double gx, double gy,
                                          double gz, double Fx, double Fy, double Fz)
{
    (aux + 0 * S)[site] = gx; (aux + 1 * S)[site] = gy; (aux + 2 * S)[site] = gz;
    (aux + 3 * S)[site] = Fx; (aux + 4 * S)[site] = Fy; (aux + 5 * S)[site] = Fz;
}

__device__ __forceinline__ double phase(double rR, double rB) { return (rR - rB) / (rR + rB); }

// One wavefront = one tile of one case: the tile -> case table gives the case, its entry the rest.
struct TileCtx {
    CaseConst cc;
    gcdouble* in;    // row blocks of the current state
    gdouble* out;    // row blocks of the next state
    gdouble* aux;
    long long S;     // aux stride
    int strip, y0, y1;
    int seg;         // segment index of the tile within its case
    int kind;        // kernel family of the case
};

__device__ __forceinline__ bool tile_ctx_at(const StepLaunch& L, int tile, TileCtx& t)
{
    const int c = L.tile_case[tile];
    const CaseDev& d = L.tab[c];
    t.cc = d.cc;
    t.kind = d.kind;
    t.in = (gcdouble*)d.pop[L.parity];
    t.out = (gdouble*)d.pop[L.parity ^ 1];
    t.S = d.S;
    t.aux = L.record ? (gdouble*)d.aux : nullptr;
    const int local = tile - d.tile0;
    t.strip = local % d.nstrip;
    const int seg = local / d.nstrip;
    t.seg = seg;
    t.y0 = seg * d.ly;
    t.y1 = min(t.y0 + d.ly, d.cc.ny);
    return t.y0 < d.cc.ny;
}

__device__ __forceinline__ bool tile_ctx(const StepLaunch& L, TileCtx& t)
{
    return tile_ctx_at(L, xcd_tile(blockIdx.x, gridDim.x), t);
}

// ------------------------------------------------------------------ P1 / P3 and F3 (1-ring)
// Aligned 64-column strips; the colour field of the two columns beside the strip comes from one
// extra 38-lane load per colour and row (Halo1), summed in the same order as every other site's.
template <bool F3>
__device__ __forceinline__ void ring1_tile(const TileCtx& t, double (&hbuf)[2][2 * Q])
{
    const CaseConst& cc = t.cc;
    const int nx = cc.nx, ny = cc.ny;
    const Lane L = make_lane_aligned(nx, t.strip);
    const Halo1 h = make_halo1(nx, t.strip);
    // every lane loads (lanes >= 38 re-read an in-range value that no one uses): no branch
    auto halo_load = [&](int y, double& vr, double& vb) {
        const unsigned long long e = elem_index(nx, wrap_row3(y, h.dy, ny), 0, h.i, h.x);
        vr = t.in[e];
        vb = t.in[e + (unsigned)(Q * nx)];
    };
    // colour field of the two halo columns: their 38 pulled values go through LDS (hbuf) to lane 0
    // (left site) and lane 63 (right site), the only lanes that use them, each summing in index order
    // as dens() (summing on the wavefront's uniform values took 152 lane reads per row)
    auto halo_phase = [&](double vr, double vb, double& l, double& r) {
        const int lane = L.lane;
        if (lane < 2 * Q) {
            hbuf[0][lane] = vr;
            hbuf[1][lane] = vb;
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        double v = 0.0;
        if (lane == 0 || lane == 63) {
            const double* pr = hbuf[0] + (lane ? Q : 0);
            const double* pb = hbuf[1] + (lane ? Q : 0);
            double rR = pr[0], rB = pb[0];
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                rR = rR + pr[i];
                rB = rB + pb[i];
            }
            v = phase(rR, rB);
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // hbuf is written again by the next row
        l = v;
        r = v;
    };

    double cR_[Q], cB_[Q], nR[Q], nB[Q];
    double vr, vb;
    double ph[3], hl[3], hr[3];   // colour field of rows y-1, y, y+1 and of the two halo columns
    load_row(t.in, nx, ny, t.y0 - 1, L, nR, nB);
    halo_load(t.y0 - 1, vr, vb);
    ph[0] = phase(dens(nR), dens(nB));
    halo_phase(vr, vb, hl[0], hr[0]);
    load_row(t.in, nx, ny, t.y0, L, cR_, cB_);
    halo_load(t.y0, vr, vb);
    double rRc = dens(cR_), rBc = dens(cB_);
    ph[1] = phase(rRc, rBc);
    halo_phase(vr, vb, hl[1], hr[1]);

    for (int y = t.y0; y < t.y1; ++y) {
        // next row's loads go out first; the gradient-independent part of this row's collision
        // (sums and, for P1/P3, the whole BGK relaxation) runs while they are in flight
        load_row(t.in, nx, ny, y + 1, L, nR, nB);
        halo_load(y + 1, vr, vb);
        Site s;
        site_sums(s, cR_, cB_, rRc, rBc);
        if constexpr (!F3) bgk<false>(cc, s, 0.0, 0.0, 0.0);

        const double rRn = dens(nR), rBn = dens(nB);
        ph[2] = phase(rRn, rBn);
        halo_phase(vr, vb, hl[2], hr[2]);
        const Nine p9 = gather9h(ph, hl, hr);
        double gx, gy, gz;
        mediator_gradient(p9, gx, gy, gz);
        const double mod2 = gx * gx + gy * gy + gz * gz;
        const double mod = sqrt(mod2);

        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
        if constexpr (F3) {
            // mu grad phi (Operadores_tensao.cpp:273-287, 325-330)
            double g0, g1, g2;
            phase_gradient(p9, g0, g1, g2);
            const double phc = ph[1];
            double lap = 0.0;
#define ITOLBM_LAP(i) lap += (i < 7 ? W1 : W2) * (p9.at(CX[i], CY[i]) - phc);
            ITOLBM_LAP(1) ITOLBM_LAP(2) ITOLBM_LAP(3) ITOLBM_LAP(4) ITOLBM_LAP(5) ITOLBM_LAP(6)
            ITOLBM_LAP(7) ITOLBM_LAP(8) ITOLBM_LAP(9) ITOLBM_LAP(10) ITOLBM_LAP(11) ITOLBM_LAP(12)
            ITOLBM_LAP(13) ITOLBM_LAP(14) ITOLBM_LAP(15) ITOLBM_LAP(16) ITOLBM_LAP(17) ITOLBM_LAP(18)
#undef ITOLBM_LAP
            lap *= 6.0;
            const double mu = 4.0 * cc.bf * phc * (phc * phc - 1.0) - cc.kf * lap;
            Fx = mu * g0; Fy = mu * g1; Fz = mu * g2;
            bgk<true>(cc, s, Fx, Fy, Fz);
        } else {
            perturb(cc, s.f, gx, gy, gz, mod2, mod);
        }

        if (L.interior) {
            const unsigned site = (unsigned)(y * nx + L.col);
            const rsrc_t ro = make_rsrc(t.out + (long long)y * block_doubles(nx), block_bytes(nx));
            const unsigned xoff = (unsigned)L.col * 8u;
            auto out = [=](int i, double r, double b) { store_pop(ro, nx, xoff, i, r, b); };
            if (!s.skip) {
                recolour(cc, s, gx, gy, gz, mod, out);
            } else {   // left untouched (rare)
                double uR[Q], uB[Q];
                load_row(t.in, nx, ny, y, L, uR, uB);
                store_row(ro, nx, xoff, uR, uB);
            }
            if (t.aux) store_aux(t.aux, t.S, site, gx, gy, gz, Fx, Fy, Fz);
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) { cR_[i] = nR[i]; cB_[i] = nB[i]; }
        rRc = rRn; rBc = rBn;
        ph[0] = ph[1]; ph[1] = ph[2];
        hl[0] = hl[1]; hl[1] = hl[2];
        hr[0] = hr[1]; hr[1] = hr[2];
    }
}

template <bool F3>
__global__ void __launch_bounds__(64) step_ring1(StepLaunch launch)
{
    __shared__ double hbuf[2][2 * Q];
    TileCtx t;
    if (!tile_ctx(launch, t)) return;
    ring1_tile<F3>(t, hbuf);
}

// ------------------------------------------------------------------ compact form (P1/P3, F3)
// Recolouring (LBM_functions.cpp:1712-1746) forms a collided site's 38 outgoing populations as
//   f_R,i = cR f_i - t_k,   f_B,i = cB f_i + t_k      (-t_k for the second velocity of pair k)
// with cB = 1 - cR and t_k = (beta cR cB w_k rho) (c_k . gm) / (|gm| |c_k|): functions of the site's 19
// totals f_i and of five scalars, cR, rho and gm.  The compact form stores those 24 doubles per site
// (layout in step_kernels.h), and the site that pulls population i forms f_R,i and f_B,i from its
// sender's stored values with the sender's own expressions, so the values are the ones the sender
// would have stored, bit for bit, and a step moves 384 instead of 608 bytes per lattice update.
//
// A lane forms the terms t_k of its own column's sites once per row; the x - 1 and x + 1 neighbours'
// terms come from the neighbouring lanes by DPP (whole-wavefront shifts by one lane), and those of
// the two columns beside the strip from two halo lanes (x0 - 1 in lane 0, x0 + 64 in lane 63).  A
// row with a raw site among its senders is rebuilt population by population from loads (rare: the
// first step after an upload, or a site with rho <= 1e-30).
// px: elements from one slot of a row block to the next (the slot pitch, nx or more)
__device__ __forceinline__ long long cblock_doubles(int px) { return (long long)CSLOTS * px; }
__device__ __forceinline__ unsigned cblock_bytes(int px) { return (unsigned)(cblock_doubles(px) * (long long)sizeof(double)); }
__device__ __forceinline__ unsigned cslot_off(int slot, int px) { return (unsigned)(slot * px) * (unsigned)sizeof(double); }

// t_k of recolour() for pair k = 0 .. 8, the pairs (1,2) (3,4) (5,6) (7,8) (9,10) (11,12) (13,14) (16,15)
// (18,17): the axis pairs with fw1 and |gm|, the diagonal ones with fw2 and |gm| sqrt 2
template <int K>
__device__ __forceinline__ double pair_t(double fw1, double fw2, double gx, double gy, double gz, double mod,
                                         double mod_d)
{
    double prod;
    if constexpr (K == 0) prod = gx;
    else if constexpr (K == 1) prod = gy;
    else if constexpr (K == 2) prod = gz;
    else if constexpr (K == 3) prod = gx + gy;
    else if constexpr (K == 4) prod = gx - gy;
    else if constexpr (K == 5) prod = gx + gz;
    else if constexpr (K == 6) prod = gx - gz;
    else if constexpr (K == 7) prod = gy + gz;
    else prod = gy - gz;
    const double cosp = (mod != 0.0) ? prod / (K < 3 ? mod : mod_d) : 0.0;
    return (K < 3 ? fw1 : fw2) * cosp;
}

// recolouring terms of a stored site, formed as its recolour() formed them
struct Rc {
    double cR, cB;
    double t[9];
};

__device__ __forceinline__ Rc rc_of(double beta, double cR, double rho, double gx, double gy, double gz)
{
    Rc r;
    r.cR = cR;
    r.cB = 1.0 - cR;
    const double fator = beta * cR * r.cB;
    const double fw1 = fator * W1 * rho, fw2 = fator * W2 * rho;
    const double mod = sqrt(gx * gx + gy * gy + gz * gz);
    const double mod_d = mod * SQRT2;
    r.t[0] = pair_t<0>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[1] = pair_t<1>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[2] = pair_t<2>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[3] = pair_t<3>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[4] = pair_t<4>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[5] = pair_t<5>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[6] = pair_t<6>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[7] = pair_t<7>(fw1, fw2, gx, gy, gz, mod, mod_d);
    r.t[8] = pair_t<8>(fw1, fw2, gx, gy, gz, mod, mod_d);
    return r;
}

// first velocity of its recolouring pair (t_k), or the second (-t_k)
__host__ __device__ constexpr bool pair_first(int i) { return i <= 14 ? (i & 1) != 0 : (i & 1) == 0; }

// f_R,i and f_B,i of velocity I from the sender's total f and terms, as recolour() forms them
template <int I>
__device__ __forceinline__ void csplit(double f, double cR, double cB, double t, double& r, double& b)
{
    if constexpr (I == 0) {
        r = cR * f;
        b = cB * f;
    } else if constexpr (pair_first(I)) {
        r = cR * f - t;
        b = cB * f + t;
    } else {
        r = cR * f - (-t);
        b = cB * f + (-t);
    }
}

// One stored site read for population i (run-time, per lane): its total and scalars.
struct CPop {
    double f, cR, rho, gx, gy, gz;
};

__device__ __forceinline__ CPop cpop_load(gcdouble* __restrict__ in, long long e0, int px, int i)
{
    CPop p;
    p.f = in[e0 + (long long)i * px];
    p.cR = in[e0 + (long long)CS_CR * px];
    p.rho = in[e0 + (long long)CS_RHO * px];
    p.gx = in[e0 + (long long)CS_G * px];
    p.gy = in[e0 + (long long)(CS_G + 1) * px];
    p.gz = in[e0 + (long long)(CS_G + 2) * px];
    return p;
}

// t_k of pair k chosen at run time (per lane): the operations of rc_of / pair_t for that pair
__device__ __forceinline__ double pair_t_rt(int k, double beta, double cR, double rho, double gx, double gy, double gz)
{
    const double cB = 1.0 - cR;
    const double fator = beta * cR * cB;
    const double fw = fator * (k < 3 ? W1 : W2) * rho;
    const double mod = sqrt(gx * gx + gy * gy + gz * gz);
    const double den = (k < 3) ? mod : mod * SQRT2;
    const int base = (k <= 6) ? 2 * k + 1 : 2 * k + 2;   // first velocity of pair k
    const double prod = cdot_rt(base, gx, gy, gz);
    return fw * ((mod != 0.0) ? prod / den : 0.0);
}

// f_R,i and f_B,i of the site whose slot 0 is element e0 (velocity i chosen at run time), or its stored
// values when it is raw
__device__ __forceinline__ void cpop_split(gcdouble* __restrict__ in, long long e0, int px, double beta, int i,
                                           const CPop& p, double& r, double& b)
{
    if (p.rho <= 1.0e-30) {
        r = p.f;
        b = in[e0 + (long long)cs_raw_b(i) * px];
        return;
    }
    const double cR = p.cR, cB = 1.0 - cR;
    if (i == 0) {
        r = cR * p.f;
        b = cB * p.f;
        return;
    }
    const double t = pair_t_rt((i - 1) >> 1, beta, cR, p.rho, p.gx, p.gy, p.gz);
    if (pair_first(i)) {
        r = cR * p.f - t;
        b = cB * p.f + t;
    } else {
        r = cR * p.f - (-t);
        b = cB * p.f + (-t);
    }
}

// site_sums() given f_i = fR_i + fB_i and the colour sums rR, rB already formed
__device__ __forceinline__ void site_sums_f(Site& s, const double f[Q], double rR, double rB)
{
    s.rho = rR + rB;
    s.skip = (s.rho <= 1.0e-30);
    s.cR = rR / s.rho;
    s.cB = 1.0 - s.cR;
#pragma unroll
    for (int i = 0; i < Q; ++i) s.f[i] = f[i];
    double rho2 = f[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) rho2 = rho2 + f[i];
    double mx = f[1] - f[2];
    mx = mx + f[7]; mx = mx - f[8]; mx = mx + f[9]; mx = mx - f[10];
    mx = mx + f[11]; mx = mx - f[12]; mx = mx + f[13]; mx = mx - f[14];
    double my = f[3] - f[4];
    my = my + f[7]; my = my - f[8]; my = my - f[9]; my = my + f[10];
    my = my - f[15]; my = my + f[16]; my = my - f[17]; my = my + f[18];
    double mz = f[5] - f[6];
    mz = mz + f[11]; mz = mz - f[12]; mz = mz - f[13]; mz = mz + f[14];
    mz = mz - f[15]; mz = mz + f[16]; mz = mz + f[17]; mz = mz - f[18];
    s.rho2 = rho2; s.mx = mx; s.my = my; s.mz = mz;
}

// The recolouring terms of the three stored rows a pulled row reads (y - 1, y, y + 1), in an LDS ring:
// field 0 cR, fields 1 + k the t_k, columns x0 - 1 (0) .. x0 + 64 (65).  Each lane writes its own
// column's terms once per stored row and reads its own and its x-neighbours' for every population it
// pulls; the two halo columns' terms (the pairs with c_x != 0 only) are formed by lanes 0 - 9, one pair
// per lane.
struct Ring1cLds {
    double rc[3][10][66];
    double hbuf[2][2 * Q];
};
constexpr int HALO_PAIRS[5] = {0, 3, 4, 5, 6};   // pairs (1,2) (7,8) (9,10) (11,12) (13,14)

// The pulling side of the compact form for one tile (an aligned 64-column strip): loads of the stored
// scalars and of the pulled totals, the recolouring terms of three stored rows in an LDS ring
// (rc[slot][field][column]), and a pulled row's populations formed from them.
struct CPull {
    struct Scal {
        double v[5];   // cR, rho, gx, gy, gz
    };
    gcdouble* in;
    double (*rc)[10][66];
    Lane L;
    long long rb;      // doubles per row block
    unsigned bytes;    // bytes per row block
    unsigned ooff;     // byte offset of the lane's column
    unsigned hoff;     // halo lanes: byte offset of the halo column, else past the block
    int ny, px, hk, hcol, first;
    bool hlane;
    double beta;

    // the halo columns' scalars: lanes 0 - 4 read column x0 - 1, lanes 5 - 9 column x0 + 64, one pair of
    // their terms each; the other lanes read past the buffer resource's extent, which returns 0 without
    // a memory access.  `first` is the first stored row the ring holds.
    __device__ __forceinline__ CPull(const TileCtx& t, const Lane& lane_, double (*ring)[10][66], int first_row)
        : in(t.in), rc(ring), L(lane_), first(first_row)
    {
        const int nx = t.cc.nx;
        ny = t.cc.ny;
        px = t.cc.px;
        rb = cblock_doubles(px);
        bytes = cblock_bytes(px);
        ooff = (unsigned)L.col * 8u;
        hlane = L.lane < 10;
        hk = HALO_PAIRS[L.lane % 5];
        hcol = L.lane < 5 ? 0 : 65;
        hoff = hlane ? (unsigned)wrapi(t.strip * 64 + (L.lane < 5 ? -1 : 64), nx) * 8u : 0x80000000u;
        beta = t.cc.beta;
    }
    __device__ __forceinline__ int slot(int r) const
    {
        const int m = (r - first) % 3;   // rows on either side of `first` (a tile may march up)
        return m < 0 ? m + 3 : m;
    }

    // live false: nothing is read, through an empty buffer resource; the loads are still issued, so every
    // path leaves the same loads outstanding
    __device__ __forceinline__ void scal_load(int r, Scal& o, Scal& hs, bool live = true) const
    {
        const rsrc_t rr = make_rsrc(in + wrapi(r, ny) * rb, live ? bytes : 0u);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            o.v[q] = buf_load<double>(rr, ooff, cslot_off(CS_CR + q, px));
            hs.v[q] = buf_load<double>(rr, hoff, cslot_off(CS_CR + q, px));
        }
    }
    // terms of stored row r into the ring; returns whether a site of the row (strip or halo columns) is raw
    __device__ __forceinline__ bool rc_write(int r, const Scal& o, const Scal& hs) const
    {
        const Rc a = rc_of(beta, o.v[0], o.v[1], o.v[2], o.v[3], o.v[4]);
        auto& R = rc[slot(r)];
        R[0][L.lane + 1] = a.cR;
#pragma unroll
        for (int k = 0; k < 9; ++k) R[1 + k][L.lane + 1] = a.t[k];
        if (hlane) {
            R[1 + hk][hcol] = pair_t_rt(hk, beta, hs.v[0], hs.v[1], hs.v[2], hs.v[3], hs.v[4]);
            R[0][hcol] = hs.v[0];
        }
        return __ballot(o.v[1] <= 1.0e-30 || (hlane && hs.v[1] <= 1.0e-30)) != 0;
    }
    // pulled totals of row Y (the x-shifted loads of load_row, one value per velocity)
    __device__ __forceinline__ void load_tot(int Y, double (&f)[Q], bool live = true) const
    {
        const unsigned nb = live ? bytes : 0u;
        const rsrc_t rm = make_rsrc(in + wrapi(Y - 1, ny) * rb, nb);
        const rsrc_t r0 = make_rsrc(in + wrapi(Y, ny) * rb, nb);
        const rsrc_t rp = make_rsrc(in + wrapi(Y + 1, ny) * rb, nb);
        const unsigned om = (unsigned)L.xm * 8u, op = (unsigned)L.xp * 8u;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const unsigned xs = CX[i] > 0 ? om : (CX[i] < 0 ? op : ooff);
            const rsrc_t r = CY[i] > 0 ? rm : (CY[i] < 0 ? rp : r0);
            f[i] = buf_load<double>(r, xs, cslot_off(i, px));
        }
    }
    // Row Y's pulled populations, formed in index order from the totals and the ring (or, when a sender
    // is raw, from loads): their totals f_i = fR_i + fB_i and the colour sums rR, rB, as dens() sums them
    __device__ __forceinline__ void pull_sums(int Y, const double (&tot)[Q], bool raw, double (&f)[Q], double& rR,
                                              double& rB) const
    {
        const auto& Rm = rc[slot(Y - 1)];
        const auto& R0 = rc[slot(Y)];
        const auto& Rp = rc[slot(Y + 1)];
        if (!raw) {
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                double r, b;
                const auto& Rs = CY[i] > 0 ? Rm : (CY[i] < 0 ? Rp : R0);
                const int c = L.lane + 1 - CX[i];
                const double cR = Rs[0][c];
                const double tk = (i == 0) ? 0.0 : Rs[1 + ((i - 1) >> 1)][c];
                switch (i) {   // csplit needs the velocity as a constant
#define ITOLBM_CS(n) case n: csplit<n>(tot[i], cR, 1.0 - cR, tk, r, b); break;
                ITOLBM_CS(0) ITOLBM_CS(1) ITOLBM_CS(2) ITOLBM_CS(3) ITOLBM_CS(4) ITOLBM_CS(5) ITOLBM_CS(6)
                ITOLBM_CS(7) ITOLBM_CS(8) ITOLBM_CS(9) ITOLBM_CS(10) ITOLBM_CS(11) ITOLBM_CS(12) ITOLBM_CS(13)
                ITOLBM_CS(14) ITOLBM_CS(15) ITOLBM_CS(16) ITOLBM_CS(17) default: csplit<18>(tot[i], cR, 1.0 - cR, tk, r, b);
#undef ITOLBM_CS
                }
                if (i == 0) {
                    rR = r;
                    rB = b;
                } else {
                    rR = rR + r;
                    rB = rB + b;
                }
                f[i] = r + b;
            }
        } else {
            // a raw sender among the rows (rare): each population from its sender's stored values, in a
            // rolled loop (unrolled, its 19 copies made the kernel several times larger, and launches of
            // a few wavefronts, which run its code once from a cold instruction cache, slower)
            const int xm = L.xm, xp = L.xp, xc = L.col;
#pragma unroll 1
            for (int i = 0; i < Q; ++i) {
                const int xs = cx_of(i) > 0 ? xm : (cx_of(i) < 0 ? xp : xc);
                const long long e0 = (long long)wrapi(Y - cy_of(i), ny) * rb + xs;
                const CPop p = cpop_load(in, e0, px, i);
                double r, b;
                cpop_split(in, e0, px, beta, i, p, r, b);
                rR = (i == 0) ? r : rR + r;
                rB = (i == 0) ? b : rB + b;
                const double v = r + b;
#pragma unroll
                for (int k = 0; k < Q; ++k) f[k] = (k == i) ? v : f[k];   // f[i] = v, f kept in registers
            }
        }
    }
    // the site's own pulled populations stored raw (a site the collision leaves untouched); a rolled
    // loop: unrolled, its loads took the registers of a second row
    __device__ __forceinline__ void store_raw(gdouble* out, int y) const
    {
        // (the loop reads locals: with the members it kept the whole object in scratch memory)
        gcdouble* const src = in;
        const long long b0 = rb;
        const int xm = L.xm, xp = L.xp, xc = L.col, n_y = ny, p_x = px;
        const unsigned oo = ooff;
        const double be = beta;
        const rsrc_t ro = make_rsrc(out + (long long)y * b0, bytes);
#pragma unroll 1
        for (int i = 0; i < Q; ++i) {
            const int xs = cx_of(i) > 0 ? xm : (cx_of(i) < 0 ? xp : xc);
            const long long e0 = (long long)wrapi(y - cy_of(i), n_y) * b0 + xs;
            const CPop p = cpop_load(src, e0, p_x, i);
            double r, b;
            cpop_split(src, e0, p_x, be, i, p, r, b);
            buf_store<double>(r, ro, oo, cslot_off(i, p_x));
            buf_store<double>(b, ro, oo, cslot_off(cs_raw_b(i), p_x));
        }
        buf_store<double>(CS_RAW, ro, oo, cslot_off(CS_RHO, p_x));
    }
    // a collided site: its totals and the scalars its recolouring is formed from
    __device__ __forceinline__ void store(gdouble* out, int y, const Site& s, double gx, double gy, double gz) const
    {
        const rsrc_t ro = make_rsrc(out + (long long)y * rb, bytes);
#pragma unroll
        for (int i = 0; i < Q; ++i) buf_store<double>(s.f[i], ro, ooff, cslot_off(i, px));
        buf_store<double>(s.cR, ro, ooff, cslot_off(CS_CR, px));
        buf_store<double>(s.rho, ro, ooff, cslot_off(CS_RHO, px));
        buf_store<double>(gx, ro, ooff, cslot_off(CS_G, px));
        buf_store<double>(gy, ro, ooff, cslot_off(CS_G + 1, px));
        buf_store<double>(gz, ro, ooff, cslot_off(CS_G + 2, px));
    }
};

template <bool F3>
__device__ __forceinline__ void ring1c_tile(const TileCtx& t, Ring1cLds& lds)
{
    const CaseConst& cc = t.cc;
    const int nx = cc.nx, ny = cc.ny;
    const Lane L = make_lane_aligned(nx, t.strip);
    const Halo1 h = make_halo1(nx, t.strip);
    const int lane = L.lane;
    const CPull P(t, L, lds.rc, t.y0 - 2);
    const int px = P.px;
    const long long rb = P.rb;
    auto lds_order = [] { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); };
    using Scal = CPull::Scal;
    // the halo columns' colour field: lane j < 38 pulls velocity h.i of halo site j / 19 from its
    // sender's stored values; the pairs go through LDS to lanes 0 and 63, which sum them in index order
    auto halo_e0 = [&](int Y) { return (long long)wrap_row3(Y, h.dy, ny) * rb + h.x; };
    auto halo_phase = [&](int Y, const CPop& hp, double& l, double& r) {
        double vr, vb;
        cpop_split(t.in, halo_e0(Y), px, cc.beta, h.i, hp, vr, vb);
        if (lane < 2 * Q) {
            lds.hbuf[0][lane] = vr;
            lds.hbuf[1][lane] = vb;
        }
        lds_order();
        double v = 0.0;
        if (lane == 0 || lane == 63) {
            const double* pr = lds.hbuf[0] + (lane ? Q : 0);
            const double* pb = lds.hbuf[1] + (lane ? Q : 0);
            double rR = pr[0], rB = pb[0];
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                rR = rR + pr[i];
                rB = rB + pb[i];
            }
            v = phase(rR, rB);
        }
        lds_order();   // hbuf is written again by the next row
        l = v;
        r = v;
    };

    // Even segments march down from their first row, odd ones up from their last: two neighbouring
    // segments then read the rows beside their shared boundary both at the start or both at the end
    // of their march, while those rows are in their common L2.  The rows are visited in march order
    // (d = +1 or -1); the colour-field window is kept in march order and put back in row order for
    // the stencils, so every site's arithmetic is the same in both directions.
    const int d = (t.seg & 1) ? -1 : 1;
    const int s0 = d > 0 ? t.y0 : t.y1 - 1;
    Scal sc, hsc;
    double tot[Q], fn[Q];
    double rRn, rBn;
    double ph[3], hl[3], hr[3];   // colour field of the rows behind, at and ahead of the current one
    CPop hp;
    bool Wm, W0, Wp;   // a raw site in the ring's stored rows (oldest, middle, newest in march order)
    // The loop's first two passes (k = -2, -1) only pull rows s0 - d and s0, the rows the first
    // collision needs besides its own next one: one copy of the pulling code serves them all.
    // Stored rows s0 - 2d and s0 - d go into the ring first; every pass adds the next one.
    {
        Scal sa, ha;
        P.scal_load(s0 - 2 * d, sc, hsc);
        P.scal_load(s0 - d, sa, ha);
        Wm = false;
        W0 = P.rc_write(s0 - 2 * d, sc, hsc);
        Wp = P.rc_write(s0 - d, sa, ha);
    }
    for (int k = -2, y = s0 - 2 * d; k < t.y1 - t.y0; ++k, y += d) {
        // the next row's loads go out first (its pulled totals, the scalars of stored row y + 2d and the
        // halo pulls); the gradient-independent part of this row's collision runs while they are in flight
        P.load_tot(y + d, tot);
        P.scal_load(y + 2 * d, sc, hsc);
        hp = cpop_load(t.in, halo_e0(y + d), px, h.i);
        const bool collide = k >= 0;
        Site s;
        if (collide) {
            site_sums_f(s, fn, rRn, rBn);
            if constexpr (!F3) bgk<false>(cc, s, 0.0, 0.0, 0.0);
        }

        Wm = W0; W0 = Wp;
        lds_order();   // the ring slot of stored row y - d is written again
        Wp = P.rc_write(y + 2 * d, sc, hsc);
        lds_order();
        P.pull_sums(y + d, tot, Wm || W0 || Wp, fn, rRn, rBn);
        ph[2] = phase(rRn, rBn);
        halo_phase(y + d, hp, hl[2], hr[2]);
        if (!collide) {
            ph[0] = ph[1]; ph[1] = ph[2];
            hl[0] = hl[1]; hl[1] = hl[2];
            hr[0] = hr[1]; hr[1] = hr[2];
            continue;
        }
        // rows y - 1, y, y + 1
        const double pr[3] = {d > 0 ? ph[0] : ph[2], ph[1], d > 0 ? ph[2] : ph[0]};
        const double lr[3] = {d > 0 ? hl[0] : hl[2], hl[1], d > 0 ? hl[2] : hl[0]};
        const double rr[3] = {d > 0 ? hr[0] : hr[2], hr[1], d > 0 ? hr[2] : hr[0]};
        const Nine p9 = gather9h(pr, lr, rr);
        double gx, gy, gz;
        mediator_gradient(p9, gx, gy, gz);
        const double mod2 = gx * gx + gy * gy + gz * gz;
        const double mod = sqrt(mod2);

        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
        if constexpr (F3) {
            // mu grad phi (Operadores_tensao.cpp:273-287, 325-330)
            double g0, g1, g2;
            phase_gradient(p9, g0, g1, g2);
            const double phc = ph[1];
            double lap = 0.0;
#define ITOLBM_LAP(i) lap += (i < 7 ? W1 : W2) * (p9.at(CX[i], CY[i]) - phc);
            ITOLBM_LAP(1) ITOLBM_LAP(2) ITOLBM_LAP(3) ITOLBM_LAP(4) ITOLBM_LAP(5) ITOLBM_LAP(6)
            ITOLBM_LAP(7) ITOLBM_LAP(8) ITOLBM_LAP(9) ITOLBM_LAP(10) ITOLBM_LAP(11) ITOLBM_LAP(12)
            ITOLBM_LAP(13) ITOLBM_LAP(14) ITOLBM_LAP(15) ITOLBM_LAP(16) ITOLBM_LAP(17) ITOLBM_LAP(18)
#undef ITOLBM_LAP
            lap *= 6.0;
            const double mu = 4.0 * cc.bf * phc * (phc * phc - 1.0) - cc.kf * lap;
            Fx = mu * g0; Fy = mu * g1; Fz = mu * g2;
            bgk<true>(cc, s, Fx, Fy, Fz);
        } else {
            perturb(cc, s.f, gx, gy, gz, mod2, mod);
        }

        if (L.interior) {
            const unsigned site = (unsigned)(y * nx + L.col);
            if (!s.skip) P.store(t.out, y, s, gx, gy, gz);
            else P.store_raw(t.out, y);   // left untouched (rare): the pulled populations, raw
            if (t.aux) store_aux(t.aux, t.S, site, gx, gy, gz, Fx, Fy, Fz);
        }
        ph[0] = ph[1]; ph[1] = ph[2];
        hl[0] = hl[1]; hl[1] = hl[2];
        hr[0] = hr[1]; hr[1] = hr[2];
    }
}

template <bool F3>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) step_ring1c(StepLaunch launch)
{
    __shared__ Ring1cLds lds;
    TileCtx t;
    if (!tile_ctx(launch, t)) return;
    ring1c_tile<F3>(t, lds);
}

// ------------------------------------------------------------------ F1 / F2 (2-ring)
struct Aux6 { double a[6]; };

template <int OP>
__device__ __forceinline__ Aux6 aux_of(const CaseConst& cc, double g0, double g1, double g2)
{
    Aux6 r;
    const double m = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    if constexpr (OP == OP_F1) {
        const double inv = (m > 1.0e-12) ? 1.0 / m : 0.0;
        r.a[0] = g0 * inv; r.a[1] = g1 * inv; r.a[2] = g2 * inv;
        r.a[3] = r.a[4] = r.a[5] = 0.0;
    } else {
        const double h = cc.hs;
        const double inv = (m > 1.0e-30) ? 1.0 / m : 0.0;
        r.a[0] = h * (m - g0 * g0 * inv);
        r.a[1] = h * (m - g1 * g1 * inv);
        r.a[2] = h * (m - g2 * g2 * inv);
        r.a[3] = -h * g0 * g1 * inv;
        r.a[4] = -h * g0 * g2 * inv;
        r.a[5] = -h * g1 * g2 * inv;
    }
    return r;
}

// Aligned 64-column strips for F1 / F2: the colour field is needed two columns beyond the strip on
// each side; those four halo columns come from two extra loads per colour and row (Halo2), summed in
// the same order as every other site's, so the halo values are identical to the owning lanes'.
//
// The fields derived from the populations live in LDS rings, not in registers:
//   phi  colour field, 5 rows (y - 2 .. y + 2) x 68 columns (x0 - 2 .. x0 + 65);
//   n/T  F1's unit normal n or F2's tensor T (NA doubles), 3 rows (y - 1 .. y + 1) x 66 columns
//        (x0 - 1 .. x0 + 64).
// Each row iteration forms the colour field of row y + 2 from its pulled populations, n / T of row
// y + 1 from the colour-field ring, and the force and mediator gradient of row y from the rings;
// then it collides row y, whose populations are read again (read two iterations ago for the colour
// field).  Holding the 3 x 3 neighbourhoods of every component and the five colour-field rows in
// registers took 256 VGPRs plus accumulation registers (one wavefront per SIMD); with the rings it
// takes 169 (F1) / 209 (F2), two wavefronts per SIMD, and summing the halo sites' densities through
// LDS instead of 304 lane reads per row halved the instructions of the row loop.  Eight 512^2 cases:
// F1 347 -> 309 us, F2 376 -> 318 us per launch.  What bounds them now is the re-read of row y: the
// rows between a strip's colour-field read and its re-read (3 x 19.5 KB per wavefront) exceed an
// XCD's 4 MB L2 at two wavefronts per SIMD, so the re-read comes from the last-level cache (L2 fetch
// 2.2x the algorithmic reads).  Every value is computed with the same expressions as before, so the
// results are unchanged bit for bit.
template <int OP>
struct Ring2Lds {
    static constexpr int NA = (OP == OP_F1) ? 3 : 6;
    double phir[5][68];      // colour-field ring columns: x0 - 2 (0) .. x0 + 65 (67)
    double ring[3][NA][66];  // n / T ring columns: x0 - 1 (0) .. x0 + 64 (65)
    double hbuf[2][4 * Q];   // pulled populations of the four halo sites, by colour
};

// LDS of the cached form (step_ring2c): Ring2Lds and the totals and colour sums of two rows,
// [slot][value][lane] (values 0 .. Q - 1 the totals f_i, Q and Q + 1 the colour sums); each lane reads
// back only what it wrote.
template <int OP>
struct Ring2cLds : Ring2Lds<OP> {
    double kept[2][Q + 2][64];
};

// CACHED: what the collision of a row needs from its pulled populations (the totals f_i = fR_i + fB_i
// and the colour sums) is kept from its colour field until its collision, two iterations later,
// instead of reading the populations again: rows y and y + 1 in LDS (lds.kept), row y + 2 in registers
// until row y's slot is free.  The loads of row y + 3 are issued before row y is collided, so they are
// in flight over the collision.  This takes one wavefront per SIMD.
template <int OP, bool CACHED, class Lds>
__device__ __forceinline__ void ring2_tile(const TileCtx& t, Lds& lds)
{
    constexpr int NA = Ring2Lds<OP>::NA;
    constexpr int NP = 68;
    constexpr int NC = 66;
    auto& phir = lds.phir;
    auto& ring = lds.ring;
    auto& hbuf = lds.hbuf;
    const CaseConst& cc = t.cc;
    const int nx = cc.nx, ny = cc.ny;
    const Lane L = make_lane_aligned(nx, t.strip);
    const Halo2 h = make_halo2(nx, t.strip);
    const int lane = L.lane;
    // ring slots: colour field of row yy (yy >= y0 - 2), n / T of row yy (yy >= y0 - 1)
    auto pslot = [&](int yy) { return (yy - t.y0 + 2) % 5; };
    auto nslot = [&](int yy) { return (yy - t.y0 + 1) % 3; };
    // LDS written by some lanes is read by others below: the wavefront's LDS operations run in order,
    // and this keeps the compiler from moving reads above writes (it waits on LDS only)
    auto lds_order = [] { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); };

    // a row's pulled populations: the own lanes' 38 and this lane's values of the two halo loads
    struct Pulled { double fR[Q], fB[Q], ra, ba, rb, bb; };
    // with live = false the own lanes' loads read nothing (empty resources) and the halo loads read row
    // yy's halo values (valid addresses)
    auto issue = [&](int yy, Pulled& p, bool live) {
        load_row(t.in, nx, ny, yy, L, p.fR, p.fB, live);
        const unsigned long long e0 = elem_index(nx, wrap_row3(yy, h.dy[0], ny), 0, h.i[0], h.x[0]);
        const unsigned long long e1 = elem_index(nx, wrap_row3(yy, h.dy[1], ny), 0, h.i[1], h.x[1]);
        const unsigned cb = (unsigned)(Q * nx);
        p.ra = t.in[e0]; p.ba = t.in[e0 + cb]; p.rb = t.in[e1]; p.bb = t.in[e1 + cb];
    };
    // colour field of row yy into its ring slot: own lanes, and the four halo columns.  The halo sites'
    // 76 pulled values (two loads per colour, value g = 19 s + i in lane g of load A or g - 64 of load B)
    // go through LDS to lanes 0 .. 3, each of which sums one site's values in index order, as dens().
    // rR / rB receive the own lanes' colour sums.
    auto consume = [&](int yy, const Pulled& p, double& rR, double& rB) {
        const double (&fR)[Q] = p.fR;
        const double (&fB)[Q] = p.fB;
        const double ra = p.ra, ba = p.ba, rb = p.rb, bb = p.bb;
        double* row = phir[pslot(yy)];
        rR = dens(fR);
        rB = dens(fB);
        row[lane + 2] = phase(rR, rB);
        hbuf[0][lane] = ra;
        hbuf[1][lane] = ba;
        if (lane < 4 * Q - 64) {
            hbuf[0][64 + lane] = rb;
            hbuf[1][64 + lane] = bb;
        }
        lds_order();
        if (lane < 4) {
            const double* vr = hbuf[0] + Q * lane;
            const double* vb = hbuf[1] + Q * lane;
            double rR = vr[0], rB = vb[0];
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                rR = rR + vr[i];
                rB = rB + vb[i];
            }
            row[lane < 2 ? lane : NP - 4 + lane] = phase(rR, rB);
        }
        lds_order();   // hbuf is written again by the next row
    };
    auto phi_row = [&](int yy) {
        Pulled p;
        double rR, rB;
        issue(yy, p, true);
        consume(yy, p, rR, rB);
    };
    // CACHED: the row's totals f_i = fR_i + fB_i and colour sums, all its collision needs
    auto keep = [&](int yy, const Pulled& p, double (&f)[Q], double& rR, double& rB) {
        consume(yy, p, rR, rB);
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = p.fR[i] + p.fB[i];
    };
    // kept values of row yy in and out of its LDS slot
    auto put = [&](int yy, const double (&f)[Q], double rR, double rB) {
        if constexpr (CACHED) {
            auto& k = lds.kept[(yy - t.y0) & 1];
#pragma unroll
            for (int i = 0; i < Q; ++i) k[i][lane] = f[i];
            k[Q][lane] = rR;
            k[Q + 1][lane] = rB;
        }
    };
    auto get = [&](int yy, double (&f)[Q], double& rR, double& rB) {
        if constexpr (CACHED) {
            const auto& k = lds.kept[(yy - t.y0) & 1];
#pragma unroll
            for (int i = 0; i < Q; ++i) f[i] = k[i][lane];
            rR = k[Q][lane];
            rB = k[Q + 1][lane];
        }
    };
    // 3 x 3 colour field around ring column c (colour-field ring coordinates) of row yy
    auto phi9 = [&](int yy, int c) {
        Nine n;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double* row = phir[pslot(yy - 1 + j)];
#pragma unroll
            for (int i = 0; i < 3; ++i) n.v[j][i] = row[c - 1 + i];
        }
        return n;
    };
    // n / T of row yy into its ring slot (columns x0 - 1 .. x0 + 64); g receives the own lanes'
    // colour-field gradient at row yy
    auto nt_row = [&](int yy, double& g0, double& g1, double& g2) {
        const int s = nslot(yy);
        phase_gradient(phi9(yy, lane + 2), g0, g1, g2);
        const Aux6 a = aux_of<OP>(cc, g0, g1, g2);
#pragma unroll
        for (int q = 0; q < NA; ++q) ring[s][q][lane + 1] = a.a[q];
        // the two halo columns, on lanes 0 (x0 - 1) and 1 (x0 + 64)
        if (lane < 2) {
            double h0, h1, h2;
            phase_gradient(phi9(yy, lane == 0 ? 1 : NP - 2), h0, h1, h2);
            const Aux6 b = aux_of<OP>(cc, h0, h1, h2);
#pragma unroll
            for (int q = 0; q < NA; ++q) ring[s][q][lane == 0 ? 0 : NC - 1] = b.a[q];
        }
    };

    Pulled ahead;   // CACHED: row y + 2's pulled populations, loaded during row y - 1's iteration
    // one row of populations in flight at a time (the scheduler would otherwise issue all four rows'
    // loads together and size the kernel's registers for them)
    for (int k = -2; k <= 1; ++k) {
        if (CACHED && k >= 0) {
            Pulled p;
            double f[Q], rR, rB;
            issue(t.y0 + k, p, true);
            keep(t.y0 + k, p, f, rR, rB);
            put(t.y0 + k, f, rR, rB);
        } else {
            phi_row(t.y0 + k);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (CACHED) issue(t.y0 + 2, ahead, true);
    lds_order();
    double g00, g01, g02;   // colour-field gradient at row y (F1 scales it by the curvature)
    nt_row(t.y0 - 1, g00, g01, g02);
    nt_row(t.y0, g00, g01, g02);

    // one row: the colour field of row y + 2 (its totals and colour sums kept when CACHED), n / T of
    // row y + 1, then the force, mediator gradient and collision of row y (from its kept values when
    // CACHED, its populations read again otherwise)
    auto row_step = [&](int y) {
        double nf[Q], nr, nb;   // row y + 2's kept values (CACHED)
        if constexpr (CACHED) {
            keep(y + 2, ahead, nf, nr, nb);
            __builtin_amdgcn_sched_barrier(0);
            issue(y + 3, ahead, y + 1 < t.y1);   // in flight over the rest of this row
            __builtin_amdgcn_sched_barrier(0);
        } else {
            phi_row(y + 2);
        }
        lds_order();
        double gp0, gp1, gp2;
        nt_row(y + 1, gp0, gp1, gp2);
        lds_order();
        const int s0 = nslot(y);
        // neighbour value (dx, dy) of component q from the n / T ring (the slot by arithmetic: a choice
        // among three slot variables became a private-memory lookup table)
        auto A = [&](int q, int dx, int dy) { return ring[(s0 + dy + 3) % 3][q][lane + 1 + dx]; };
        double Fx, Fy, Fz;
        if constexpr (OP == OP_F1) {
            // kappa = -div_iso(n) (Operadores_tensao.cpp:187-200, 300-304)
            // (a loop over the directions kept rolled needs 25 (F1) / 71 (F2) fewer VGPRs, three waves per
            // SIMD instead of two, but waits on each direction's ring reads in turn: 10-20 % slower)
            double d = 0.0;
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                const int dx = cx_of(i), dy = cy_of(i);
                d += (i < 7 ? W1 : W2) * cdot_rt(i, A(0, dx, dy), A(1, dx, dy), A(2, dx, dy));
            }
            const double kappa = -(3.0 * d);
            Fx = cc.hs * kappa * g00; Fy = cc.hs * kappa * g01; Fz = cc.hs * kappa * g02;
        } else {
            // F_a = d_b T_ab (Operadores_tensao.cpp:306-324); T = xx yy zz xy xz yz.  The three sums are
            // independent chains over the directions; each is formed in turn, reading only the three
            // components it needs (the order within each sum is the reference's)
            double sx = 0.0, sy = 0.0, sz = 0.0;
            auto chain = [&](double& acc, int a, int b, int c) {
#pragma unroll
                for (int i = 1; i < Q; ++i) {
                    const int dx = cx_of(i), dy = cy_of(i);
                    acc += (i < 7 ? W1 : W2) * cdot_rt(i, A(a, dx, dy), A(b, dx, dy), A(c, dx, dy));
                }
            };
            chain(sx, 0, 3, 4);
            chain(sy, 3, 1, 5);
            chain(sz, 4, 5, 2);
            Fx = 3.0 * sx; Fy = 3.0 * sy; Fz = 3.0 * sz;
        }
        double gx, gy, gz;
        mediator_gradient(phi9(y, lane + 2), gx, gy, gz);
        const double mod = sqrt(gx * gx + gy * gy + gz * gz);
        // populations of row y again (read two iterations ago for the colour field, now in L2), issued
        // once the force is formed: the scheduler would otherwise hoist them above the force's ring
        // reads and hold both at once
        __builtin_amdgcn_sched_barrier(0);
        Site s;
        if constexpr (CACHED) {
            double cf[Q], cr, cb;
            get(y, cf, cr, cb);
            site_sums_f(s, cf, cr, cb);
        } else {
            double fR[Q], fB[Q];
            load_row(t.in, nx, ny, y, L, fR, fB);
            site_sums(s, fR, fB, dens(fR), dens(fB));
        }

        if (L.interior) {
            const unsigned site = (unsigned)(y * nx + L.col);
            const rsrc_t ro = make_rsrc(t.out + (long long)y * block_doubles(nx), block_bytes(nx));
            const unsigned xoff = (unsigned)L.col * 8u;
            if (!s.skip) {
                bgk<true>(cc, s, Fx, Fy, Fz);
                recolour(cc, s, gx, gy, gz, mod, [=](int i, double r, double b) { store_pop(ro, nx, xoff, i, r, b); });
            } else {   // left untouched (rare): the raw populations again
                double uR[Q], uB[Q];
                load_row(t.in, nx, ny, y, L, uR, uB);
                store_row(ro, nx, xoff, uR, uB);
            }
            if (t.aux) store_aux(t.aux, t.S, site, gx, gy, gz, Fx, Fy, Fz);
        }
        g00 = gp0; g01 = gp1; g02 = gp2;
        put(y + 2, nf, nr, nb);   // into row y's slot
    };

    for (int y = t.y0; y < t.y1; ++y) row_step(y);
}

template <int OP>
__global__ void __launch_bounds__(64) step_ring2(StepLaunch launch)
{
    __shared__ Ring2Lds<OP> lds;
    TileCtx t;
    if (!tile_ctx(launch, t)) return;
    ring2_tile<OP, false>(t, lds);
}

// the rows' totals and colour sums kept: one wavefront per SIMD (the register file's 512 per lane)
template <int OP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) step_ring2c(StepLaunch launch)
{
    __shared__ Ring2cLds<OP> lds;
    TileCtx t;
    if (!tile_ctx(launch, t)) return;
    ring2_tile<OP, true>(t, lds);
}

// All four families in one launch: each wavefront takes the tiles blockIdx.x, blockIdx.x + gridDim.x,
// ... of a table covering every case (gridDim.x a multiple of 8, so a wavefront's tiles all lie in
// its XCD's contiguous range), and runs the family kernel of each tile's case.  With a grid of one
// block per tile every wavefront takes one tile; with a grid of the resident capacity the launch is
// persistent.  The bodies are the family kernels' own, so the results are the same bit for bit.
union FusedLds {
    double ring1[2][2 * Q];
    Ring1cLds ring1c;
    Ring2Lds<OP_F1> f1;
    Ring2Lds<OP_F2> f2;
};

// CMP: the P1/P3 and F3 cases hold the compact form
template <bool CMP>
__global__ void __launch_bounds__(64) step_fused(StepLaunch launch)
{
    __shared__ FusedLds lds;
    for (int vb = blockIdx.x; vb < launch.ntiles; vb += gridDim.x) {
        TileCtx t;
        if (!tile_ctx_at(launch, xcd_tile(vb, launch.ntiles), t)) continue;
        switch (t.kind) {
        case KIND_P:
            if constexpr (CMP) ring1c_tile<false>(t, lds.ring1c);
            else ring1_tile<false>(t, lds.ring1);
            break;
        case KIND_F3:
            if constexpr (CMP) ring1c_tile<true>(t, lds.ring1c);
            else ring1_tile<true>(t, lds.ring1);
            break;
        case KIND_F1: ring2_tile<OP_F1, false>(t, lds.f1); break;
        default: ring2_tile<OP_F2, false>(t, lds.f2); break;
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // the next tile writes the LDS again
    }
}

// ------------------------------------------------------------------ boxes with solid sites
// A case with solid sites (the reference library's general geometry: read_geo's 0 sites) steps with
// three site-parallel kernels instead of a strip kernel; one thread per site, x fastest, so every
// load of a velocity row is coalesced.  The fields the strip kernels keep in registers and LDS
// (colour field, its gradient, n / T / mu) go through per-case arrays of the padded stride S.
//   geo_phase   colour field phi of every fluid site from its pulled populations;
//   geo_fields  (F class) the isotropic gradient of phi and n (F1), T (F2) or mu (F3);
//   geo_collide pull, mediator gradient, force, collision and recolouring; the site's post-collision
//               populations are stored at the site, as in the strip kernels.
// Bounce-back (def_dir_prop, LBM_functions.cpp:509-853: a population pushed towards a solid site
// comes back in the opposite direction at its own site) becomes, in the pull: f_i(x) = f*_opp(i)(x)
// when x - c_i is solid.  A mediator that would come from a solid site is the wall's, w_i wett 3
// (propag_site_med :1331).  The force stencils read the site itself where a neighbour is solid
// (aloca_campos_F, Operadores_tensao.cpp:170-178).  Every value is formed with the same expressions
// as in the strip kernels, so an all-fluid neighbourhood gives their bits.
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ __forceinline__ void geo_pull(const GeoLaunch& a, int x, int y, double fR[Q], double fB[Q])
{
    const int nx = a.cc.nx, ny = a.cc.ny;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int sx = wrap1(x - CX[i], nx), sy = wrap1(y - CY[i], ny);
        const bool back = (i != 0) && a.solid[(long long)sy * nx + sx];
        const int yy = back ? y : sy, xx = back ? x : sx, ii = back ? OPP[i] : i;
        fR[i] = a.in[elem_index(nx, yy, 0, ii, (unsigned)xx)];
        fB[i] = a.in[elem_index(nx, yy, 1, ii, (unsigned)xx)];
    }
}

// 3 x 3 values of a site field around (x, y); a solid neighbour reads the site's own value
__device__ __forceinline__ Nine geo_nine(const GeoLaunch& a, const double* __restrict__ f, int x, int y)
{
    const int nx = a.cc.nx, ny = a.cc.ny;
    const long long p = (long long)y * nx + x;
    Nine n;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            const long long q = (long long)wrap1(y + dy, ny) * nx + wrap1(x + dx, nx);
            n.v[dy + 1][dx + 1] = a.solid[q] ? f[p] : f[q];
        }
    return n;
}

__global__ void __launch_bounds__(256) geo_phase(GeoLaunch a)
{
    const long long n = (long long)a.cc.nx * a.cc.ny;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (a.solid[p]) {
        a.phi[p] = 0.0;
        return;
    }
    const int y = (int)(p / a.cc.nx), x = (int)(p - (long long)y * a.cc.nx);
    double fR[Q], fB[Q];
    geo_pull(a, x, y, fR, fB);
    a.phi[p] = phase(dens(fR), dens(fB));
}

// gradient of phi (calcula_forca :219-235) into g[3][S]; n (F1) / T (F2) into nt[q][S] (aux_of), or
// mu (F3, :273-287) into nt[0][S]
template <int OP>
__global__ void __launch_bounds__(256) geo_fields(GeoLaunch a)
{
    const long long n = (long long)a.cc.nx * a.cc.ny;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || a.solid[p]) return;
    const int y = (int)(p / a.cc.nx), x = (int)(p - (long long)y * a.cc.nx);
    const CaseConst& cc = a.cc;
    const Nine p9 = geo_nine(a, a.phi, x, y);
    double g0, g1, g2;
    phase_gradient(p9, g0, g1, g2);
    a.g[p] = g0; a.g[a.S + p] = g1; a.g[2 * a.S + p] = g2;
    if constexpr (OP == OP_F3) {
        const double phc = p9.at(0, 0);
        double lap = 0.0;
#define ITOLBM_LAP(i) lap += (i < 7 ? W1 : W2) * (p9.at(CX[i], CY[i]) - phc);
        ITOLBM_LAP(1) ITOLBM_LAP(2) ITOLBM_LAP(3) ITOLBM_LAP(4) ITOLBM_LAP(5) ITOLBM_LAP(6)
        ITOLBM_LAP(7) ITOLBM_LAP(8) ITOLBM_LAP(9) ITOLBM_LAP(10) ITOLBM_LAP(11) ITOLBM_LAP(12)
        ITOLBM_LAP(13) ITOLBM_LAP(14) ITOLBM_LAP(15) ITOLBM_LAP(16) ITOLBM_LAP(17) ITOLBM_LAP(18)
#undef ITOLBM_LAP
        lap *= 6.0;
        a.nt[p] = 4.0 * cc.bf * phc * (phc * phc - 1.0) - cc.kf * lap;
    } else {
        const Aux6 r = aux_of<OP>(cc, g0, g1, g2);
        constexpr int NA = (OP == OP_F1) ? 3 : 6;
#pragma unroll
        for (int q = 0; q < NA; ++q) a.nt[q * a.S + p] = r.a[q];
    }
}

// mediator momentum (as mediator_gradient) with the wall's mediators where the source site is solid
__device__ __forceinline__ void geo_mediator_gradient(const GeoLaunch& a, int x, int y, double& gx, double& gy,
                                                      double& gz)
{
    const int nx = a.cc.nx, ny = a.cc.ny;
    auto M = [&](double w, int dx, int dy) {
        const long long q = (long long)wrap1(y + dy, ny) * nx + wrap1(x + dx, nx);
        return a.solid[q] ? w * a.wett * 3.0 : w * (a.phi[q] * 3.0);
    };
    const double M1 = M(W1, -1, 0), M2 = M(W1, 1, 0), M3 = M(W1, 0, -1), M4 = M(W1, 0, 1);
    const double M5 = M(W1, 0, 0);
    const double M7 = M(W2, -1, -1), M8 = M(W2, 1, 1), M9 = M(W2, -1, 1), M10 = M(W2, 1, -1);
    const double M11 = M(W2, -1, 0), M12 = M(W2, 1, 0), M15 = M(W2, 0, 1), M16 = M(W2, 0, -1);
    gx = M1 - M2;
    gx = gx + M7; gx = gx - M8; gx = gx + M9; gx = gx - M10;
    gx = gx + M11; gx = gx - M12; gx = gx + M11; gx = gx - M12;
    gy = M3 - M4;
    gy = gy + M7; gy = gy - M8; gy = gy - M9; gy = gy + M10;
    gy = gy - M15; gy = gy + M16; gy = gy - M15; gy = gy + M16;
    gz = M5 - M5;
    gz = gz + M11; gz = gz - M12; gz = gz - M11; gz = gz + M12;
    gz = gz - M15; gz = gz + M16; gz = gz + M15; gz = gz - M16;
}

template <int KIND>
__global__ void __launch_bounds__(256) geo_collide(GeoLaunch a)
{
    const int nx = a.cc.nx;
    const long long n = (long long)nx * a.cc.ny;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int y = (int)(p / nx), x = (int)(p - (long long)y * nx);
    double* out = a.out;
    auto put = [&](int i, double r, double b) {
        out[elem_index(nx, y, 0, i, (unsigned)x)] = r;
        out[elem_index(nx, y, 1, i, (unsigned)x)] = b;
    };
    if (a.solid[p]) {   // a solid site holds no populations
#pragma unroll
        for (int i = 0; i < Q; ++i) put(i, 0.0, 0.0);
        if (a.aux)
            for (int q = 0; q < 6; ++q) a.aux[q * a.S + p] = 0.0;
        return;
    }
    const CaseConst& cc = a.cc;
    double fR[Q], fB[Q];
    geo_pull(a, x, y, fR, fB);
    Site s;
    site_sums(s, fR, fB, dens(fR), dens(fB));
    double gx, gy, gz;
    geo_mediator_gradient(a, x, y, gx, gy, gz);
    const double mod2 = gx * gx + gy * gy + gz * gz;
    const double mod = sqrt(mod2);
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    if constexpr (KIND == KIND_P) {
        bgk<false>(cc, s, 0.0, 0.0, 0.0);
        perturb(cc, s.f, gx, gy, gz, mod2, mod);
    } else {
        const double g0 = a.g[p], g1 = a.g[a.S + p], g2 = a.g[2 * a.S + p];
        if constexpr (KIND == KIND_F3) {
            const double mu = a.nt[p];
            Fx = mu * g0; Fy = mu * g1; Fz = mu * g2;
        } else {
            // the neighbour (dx, dy) of component q of n / T, the site itself where it is solid
            auto A = [&](int q, int dx, int dy) {
                const long long r = (long long)wrap1(y + dy, a.cc.ny) * nx + wrap1(x + dx, nx);
                return a.nt[q * a.S + (a.solid[r] ? p : r)];
            };
            if constexpr (KIND == KIND_F1) {
                double d = 0.0;
#pragma unroll
                for (int i = 1; i < Q; ++i) {
                    const int dx = cx_of(i), dy = cy_of(i);
                    d += (i < 7 ? W1 : W2) * cdot_rt(i, A(0, dx, dy), A(1, dx, dy), A(2, dx, dy));
                }
                const double kappa = -(3.0 * d);
                Fx = cc.hs * kappa * g0; Fy = cc.hs * kappa * g1; Fz = cc.hs * kappa * g2;
            } else {
                double sx = 0.0, sy = 0.0, sz = 0.0;
                auto chain = [&](double& acc, int qa, int qb, int qc) {
#pragma unroll
                    for (int i = 1; i < Q; ++i) {
                        const int dx = cx_of(i), dy = cy_of(i);
                        acc += (i < 7 ? W1 : W2) * cdot_rt(i, A(qa, dx, dy), A(qb, dx, dy), A(qc, dx, dy));
                    }
                };
                chain(sx, 0, 3, 4);
                chain(sy, 3, 1, 5);
                chain(sz, 4, 5, 2);
                Fx = 3.0 * sx; Fy = 3.0 * sy; Fz = 3.0 * sz;
            }
        }
        bgk<true>(cc, s, Fx, Fy, Fz);
    }
    if (!s.skip) {
        recolour(cc, s, gx, gy, gz, mod, put);
    } else {   // left untouched (rare): the pulled populations
#pragma unroll
        for (int i = 0; i < Q; ++i) put(i, fR[i], fB[i]);
    }
    if (a.aux) {
        a.aux[p] = gx; a.aux[a.S + p] = gy; a.aux[2 * a.S + p] = gz;
        a.aux[3 * a.S + p] = Fx; a.aux[4 * a.S + p] = Fy; a.aux[5 * a.S + p] = Fz;
    }
}

}  // namespace

int step_halo(int op) { return (op == OP_F1 || op == OP_F2) ? 2 : 1; }

hipError_t launch_step_geo(int kind, const GeoLaunch& a, hipStream_t stream)
{
    const long long n = (long long)a.cc.nx * a.cc.ny;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipLaunchKernelGGL(geo_phase, grid, block, 0, stream, a);
    switch (kind) {
    case KIND_P: break;
    case KIND_F3: hipLaunchKernelGGL(geo_fields<OP_F3>, grid, block, 0, stream, a); break;
    case KIND_F1: hipLaunchKernelGGL(geo_fields<OP_F1>, grid, block, 0, stream, a); break;
    case KIND_F2: hipLaunchKernelGGL(geo_fields<OP_F2>, grid, block, 0, stream, a); break;
    default: return hipErrorInvalidValue;
    }
    switch (kind) {
    case KIND_P: hipLaunchKernelGGL(geo_collide<KIND_P>, grid, block, 0, stream, a); break;
    case KIND_F3: hipLaunchKernelGGL(geo_collide<KIND_F3>, grid, block, 0, stream, a); break;
    case KIND_F1: hipLaunchKernelGGL(geo_collide<KIND_F1>, grid, block, 0, stream, a); break;
    default: hipLaunchKernelGGL(geo_collide<KIND_F2>, grid, block, 0, stream, a); break;
    }
    return hipGetLastError();
}



namespace {
// row blocks [y][colour][19][nx] -> [colour][19][S] (the diagnostics' and transfers' view)
__global__ void __launch_bounds__(256) f64_rows_to_soa(const double* __restrict__ pop, int nx, long long n, long long S,
                                                       double* __restrict__ soa)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long y = p / nx, x = p - y * nx;
        const double* row = pop + y * (2LL * Q * nx) + x;
#pragma unroll
        for (int j = 0; j < 2 * Q; ++j) soa[j * S + p] = row[(long long)j * nx];
    }
}

// the compact form -> [colour][19][S]: each site's 38 populations formed from its own stored values
__global__ void __launch_bounds__(256) f64c_rows_to_soa(const double* __restrict__ pop, double beta, int nx, int px,
                                                        long long n, long long S, double* __restrict__ soa)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long y = p / nx, x = p - y * nx;
        const long long e0 = y * cblock_doubles(px) + x;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const CPop c = cpop_load((gcdouble*)pop, e0, px, i);
            double r, b;
            cpop_split((gcdouble*)pop, e0, px, beta, i, c, r, b);
            soa[i * S + p] = r;
            soa[(Q + i) * S + p] = b;
        }
    }
}
}  // namespace

namespace {
// One case's state from the full FP64 form to the compact one (every site raw) or back (each site's 38
// populations formed from its stored values): the populations are unchanged, only their layout.
__global__ void __launch_bounds__(256) f64_convert(const double* __restrict__ in, double* __restrict__ out,
                                                   int to_compact, double beta, int nx, int px, long long n)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long y = p / nx, x = p - y * nx;
        const long long f0 = y * (2LL * Q * nx) + x;   // full form: row blocks [y][colour][19][nx]
        const long long c0 = y * cblock_doubles(px) + x;
        if (to_compact) {
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                out[c0 + (long long)i * px] = in[f0 + (long long)i * nx];
                out[c0 + (long long)cs_raw_b(i) * px] = in[f0 + (long long)(Q + i) * nx];
            }
            out[c0 + (long long)CS_RHO * px] = CS_RAW;
        } else {
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const CPop c = cpop_load((gcdouble*)in, c0, px, i);
                double r, b;
                cpop_split((gcdouble*)in, c0, px, beta, i, c, r, b);
                out[f0 + (long long)i * nx] = r;
                out[f0 + (long long)(Q + i) * nx] = b;
            }
        }
    }
}
}  // namespace

hipError_t launch_f64_convert(const double* in, double* out, bool to_compact, double beta, int nx, int px, long long n,
                              hipStream_t stream)
{
    const int nb = (int)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(f64_convert, dim3(nb), dim3(256), 0, stream, in, out, to_compact ? 1 : 0, beta, nx, px, n);
    return hipGetLastError();
}

hipError_t launch_f64_rows_to_soa(const double* pop, bool compact, double beta, int nx, int px, long long n,
                                  long long S, double* soa, hipStream_t stream)
{
    const int nb = (int)std::min<long long>((n + 255) / 256, 8192);
    if (compact) hipLaunchKernelGGL(f64c_rows_to_soa, dim3(nb), dim3(256), 0, stream, pop, beta, nx, px, n, S, soa);
    else hipLaunchKernelGGL(f64_rows_to_soa, dim3(nb), dim3(256), 0, stream, pop, nx, n, S, soa);
    return hipGetLastError();
}

hipError_t launch_step_f64_fused(const StepLaunch& a, int grid, hipStream_t stream)
{
    if (a.ntiles <= 0) return hipSuccess;
    if (a.compact) hipLaunchKernelGGL(step_fused<true>, dim3(grid), dim3(64), 0, stream, a);
    else hipLaunchKernelGGL(step_fused<false>, dim3(grid), dim3(64), 0, stream, a);
    return hipGetLastError();
}

int step_f64_fused_capacity(bool compact)
{
    static int cap[2] = {0, 0};
    int& c = cap[compact ? 1 : 0];
    if (c == 0) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t prop{};
        const hipError_t e = compact ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, step_fused<true>, 64, 0)
                                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, step_fused<false>, 64, 0);
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess || e != hipSuccess ||
            per_cu < 1)
            return 0;
        c = (prop.multiProcessorCount * per_cu + 7) / 8 * 8;
    }
    return c;
}

hipError_t launch_step_f64(int kind, const StepLaunch& a, int ntiles, hipStream_t stream)
{
    if (ntiles <= 0) return hipSuccess;
    const dim3 grid(ntiles), block(64);
    switch (kind) {
    case KIND_P:
        if (a.compact) hipLaunchKernelGGL(step_ring1c<false>, grid, block, 0, stream, a);
        else hipLaunchKernelGGL(step_ring1<false>, grid, block, 0, stream, a);
        break;
    case KIND_F3:
        if (a.compact) hipLaunchKernelGGL(step_ring1c<true>, grid, block, 0, stream, a);
        else hipLaunchKernelGGL(step_ring1<true>, grid, block, 0, stream, a);
        break;
    case KIND_F1:
        if (a.depth >= 2) hipLaunchKernelGGL(step_ring2c<OP_F1>, grid, block, 0, stream, a);
        else hipLaunchKernelGGL(step_ring2<OP_F1>, grid, block, 0, stream, a);
        break;
    case KIND_F2:
        if (a.depth >= 2) hipLaunchKernelGGL(step_ring2c<OP_F2>, grid, block, 0, stream, a);
        else hipLaunchKernelGGL(step_ring2<OP_F2>, grid, block, 0, stream, a);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace itolbm
