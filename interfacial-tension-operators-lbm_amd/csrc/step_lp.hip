// FP32-compute step kernels with FP32 or FP16 population storage: the same fused pass as the FP64
// kernels (pull-stream, colour field, mediator gradient, tension operator, BGK + Guo, recolouring),
// reformulated in deviation variables so that single and half precision keep their digits.
//
// Storage: one pair (h_i, e_i) per site and velocity i, in rows: pair(i, x, y) at ((y * 19 + i) px + x), px
// >= nx the slot pitch (CaseConst::px).
// A row of one velocity is contiguous (coalesced pulls and stores) and the 19 velocities of a row sit
// together (one buffer resource per row block, few pages per wavefront):
//   s_i = fR_i + fB_i - w_i                         deviation of the total population from w_i (rho0 = 1)
//   h_i = 2^k s_i - c_k r(e_i)                      stored: s_i scaled exactly (lattice.h), less the
//                                                   case's reference density
//   e_i = (fR_i - fB_i) / (fR_i + fB_i)             colour fraction of the population, exactly +-1 in a
//                                                   pure phase
// From these a site has f_i = w_i + s_i, rho = 1 + sum_i s_i, rho_R - rho_B = sum_i e_i f_i and
// phi = rho^N = (rho_R - rho_B) / rho; the colour populations are fR_i = f_i (1 + e_i) / 2 and
// fB_i = f_i (1 - e_i) / 2.  Collision and recolouring (colisao_comum, Operadores_tensao.cpp:340-373;
// coll_BGK, Collision_operators.cpp:11-49; recolloring, LBM_functions.cpp:1712-1746) act on s_i:
//   s_i += (s_i^eq - s_i) / tau + S_i (+ A|g|(w_i (c_i.n)^2 - B_i) for P1/P3),
//   s_i^eq = w_i [ (rho - 1) + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2) ],
//   e_i = phi - 2 beta cR cB w_i rho cos(c_i, gm) / f_i            (f_i post-collision).
// On the nz = 1 slab the isotropic stencils reduce to in-plane differences with weights
// w_axis = W1 + 2 W2 = 1/9 and w_diag = W2 = 1/36, and their z components vanish exactly.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "lattice.h"
#include "step_kernels.h"
#include "tile.h"

namespace itolbm {

namespace {

constexpr float WF[Q] = {WL0, WL1, WL1, WL1, WL1, WL1, WL1, WL2, WL2, WL2, WL2, WL2, WL2, WL2, WL2, WL2, WL2, WL2, WL2};
constexpr float WAX = (float)(W1 + 2.0 * W2);   // in-plane axis weight of the folded stencil
constexpr float WDG = (float)W2;                // in-plane diagonal weight
constexpr float ISQ2 = 0.70710678118654752f;    // 1 / |c| for the diagonals

template <typename T>
using gcst = __attribute__((address_space(1))) const T;
template <typename T>
using gst = __attribute__((address_space(1))) T;
using gdouble = __attribute__((address_space(1))) double;

// Stored pair of one population, (h_i, e_i) with h_i = 2^k s_i - c_k r(e_i) (lattice.h: s_i = f_i - w_i,
// 2^k = 4 / 16 / 32 by velocity class, c_k = 2^k w_k), where r(e) = ref_mid + ref_half e is the case's
// reference density deviation (1 + r(+1) for pure red, 1 + r(-1) for pure blue, set from measured
// phase densities; zero by default), so that the stored number is small in both bulk phases and
// keeps its relative digits.  The scaling by powers of two makes encoding and decoding exact apart
// from the one rounding of each sum.
//   FP32 form, 8 bytes: h and e as floats.
//   FP16 form, 4 bytes: h as an IEEE half in the low 16 bits; in the high 16 bits the minority colour
//   fraction m = (1 - |e|) / 2 as a half whose sign bit is that of e, e = copysign(1 - 2m, e).  A
//   pure phase (m = +-0) decodes to e = +-1 exactly and the thin colour tails of an interface keep
//   their relative precision, so rounding does not bias the colour masses.  |e| > 1 (a negative
//   colour population) is stored as |e| = 1.
template <typename ST> struct PairOf;
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <> struct PairOf<float> {
    using raw = f32x2;
    __device__ static __forceinline__ float a(raw r) { return r.x; }
    __device__ static __forceinline__ float e(raw r) { return r.y; }
    __device__ static __forceinline__ raw make(float a, float e) { return raw{a, e}; }
};
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
template <> struct PairOf<_Float16> {
    using raw = unsigned int;
    __device__ static __forceinline__ float a(raw r) { return (float)__builtin_bit_cast(f16x2, r).x; }
    // 1 - 2|m| >= 0, so the sign bit of the pair (that of m) is or-ed in
    __device__ static __forceinline__ float e(raw r)
    {
        const float t = __builtin_fmaf(-2.0f, __builtin_fabsf((float)__builtin_bit_cast(f16x2, r).y), 1.0f);
        return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, t) | (r & 0x80000000u));
    }
    __device__ static __forceinline__ raw make(float a, float e)
    {
        const float m = __builtin_fmaxf(__builtin_fmaf(-0.5f, __builtin_fabsf(e), 0.5f), 0.0f);
        const float sm = __builtin_bit_cast(float, __builtin_bit_cast(unsigned, m) | (__builtin_bit_cast(unsigned, e) & 0x80000000u));
        return __builtin_bit_cast(raw, f16x2{(_Float16)a, (_Float16)sm});
    }
};

// one stored pair through a buffer resource (8-byte FP32 pairs, 4-byte FP16 pairs)
template <typename R>
__device__ __forceinline__ R pair_load(rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(R) == 8) return buf_load64<R>(r, voff, soff);
    else return buf_load<unsigned>(r, voff, soff);
}
template <typename R>
__device__ __forceinline__ void pair_store(R v, rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(R) == 8) buf_store64(v, r, voff, soff);
    else buf_store<unsigned>(v, r, voff, soff);
}

// pulled populations of one site in deviation form
struct Row {
    float s[Q];   // s_i = f_i - w_i
    float e[Q];   // colour fraction
};

// Issue the pulls of site (col, y): raw pair i from row y - c_iy, column col - c_ix.  With live =
// false the loads are still issued but through empty resources: they read no memory and return 0.
// A row loop that issues its look-ahead loads on every path has the same loads outstanding on each,
// so the compiler's waits for the current row (vector memory counts retire in issue order) do not
// also wait for the rows ahead.
template <typename ST, bool SYM>
__device__ __forceinline__ void load_raw(gcst<typename PairOf<ST>::raw>* __restrict__ in, int px, int ny, int y,
                                         const Lane& L, typename PairOf<ST>::raw r[Q], bool live = true)
{
    using R = typename PairOf<ST>::raw;
    const long long pitch = (long long)n_slots(SYM) * px;   // pairs per row block (px: the slot pitch)
    const unsigned bytes = live ? (unsigned)(pitch * (long long)sizeof(R)) : 0u;
    const rsrc_t rowm = make_rsrc(in + wrapi(y - 1, ny) * pitch, bytes);
    const rsrc_t row0 = make_rsrc(in + wrapi(y, ny) * pitch, bytes);
    const rsrc_t rowp = make_rsrc(in + wrapi(y + 1, ny) * pitch, bytes);
    // the lane part is one of three byte offsets; the velocity row's offset is uniform
    const unsigned om = (unsigned)L.xm * (unsigned)sizeof(R), o0 = (unsigned)L.col * (unsigned)sizeof(R),
                   op = (unsigned)L.xp * (unsigned)sizeof(R);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (!is_stored(SYM, i)) continue;
        const unsigned xs = CX[i] > 0 ? om : (CX[i] < 0 ? op : o0);
        r[i] = pair_load<R>(CY[i] > 0 ? rowm : (CY[i] < 0 ? rowp : row0), xs,
                            (unsigned)(slot_of(SYM, i) * px * (int)sizeof(R)));
    }
#pragma unroll
    for (int i = 0; i < Q; ++i)
        if (!is_stored(SYM, i)) r[i] = r[MIRROR_OF[i]];   // the z mirror pulls from the same site
}

using Ref = LpRef;

// deviation s of a stored pair (a, e) of velocity class k: (a + c_k r(e)) 2^-k, written as
// a 2^-k + (c_k r(e)) 2^-k, which rounds identically because the scalings are exact
__device__ __forceinline__ float decode_s(float a, float e, int k, const Ref& rf)
{
    const float mid = rf.mid(k) * ISCL[k], half = rf.half(k) * ISCL[k];
    return __builtin_fmaf(a, ISCL[k], __builtin_fmaf(half, e, mid));
}

template <typename ST>
__device__ __forceinline__ void decode(const typename PairOf<ST>::raw r[Q], Row& row, const Ref& rf)
{
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        row.e[i] = PairOf<ST>::e(r[i]);
        row.s[i] = decode_s(PairOf<ST>::a(r[i]), row.e[i], vclass(i), rf);
    }
}

template <typename ST>
__device__ __forceinline__ typename PairOf<ST>::raw encode(int i, float s, float e, const Ref& rf)
{
    const int k = vclass(i);
    return PairOf<ST>::make(__builtin_fmaf(s, SCL[k], -__builtin_fmaf(rf.half(k), e, rf.mid(k))), e);
}

// density deviation, rho and phi of a site.  phi = sum e_i f_i / sum f_i is formed in FP64 (the
// products of two floats are exact there), so it carries the colour content of the stored row to
// 1e-16: the recolouring hands that content on unchanged apart from the rounding of each stored e_i,
// which is independent between populations.  An FP32 sum would change a site's colour by a rounding
// that is the same every step in a steady state, i.e. a colour source at every interface site.  A pure
// phase (all e_i = +-1) gives phi = +-1 exactly.  The density used by the collision comes from the
// deviations, which keep more digits.
__device__ __forceinline__ void row_phase(const Row& r, float& ds, float& rho, double& phi)
{
    float a = 0.f;
    double D = 0.0, F = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const float f = r.s[i] + WF[i];
        a += r.s[i];
        // the product of two floats is exact in FP64, so the fused form rounds as product-then-sum
        D = __builtin_fma((double)r.e[i], (double)f, D);
        F += (double)f;
    }
    ds = a;
    rho = 1.f + a;
    phi = D / F;
}

// colour field of a site (the gradients use it rounded to FP32)
template <typename ST>
__device__ __forceinline__ double raw_phase(const typename PairOf<ST>::raw r[Q], const Ref& rf)
{
    Row row;
    decode<ST>(r, row, rf);
    float ds, rho;
    double phi;
    row_phase(row, ds, rho, phi);
    return phi;
}

// density deviation and rho of a decoded row (the summation of row_phase)
__device__ __forceinline__ void row_density(const Row& r, float& ds, float& rho)
{
    float a = 0.f;
#pragma unroll
    for (int i = 0; i < Q; ++i) a += r.s[i];
    ds = a;
    rho = 1.f + a;
}

// 64-bit value of the lane that DPP control CTRL selects (rows outside ROWS give 0)
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_value(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sums of v over lanes 0..31 and over lanes 32..63 in a fixed tree order (pairs, quads, half rows,
// rows of 16, then row 0 into row 1 and row 2 into row 3); the totals end in lanes 31 and 63.
__device__ __forceinline__ double half_wave_sum(double v)
{
    v += dpp_value<0xB1>(v);        // quad_perm [1, 0, 3, 2]
    v += dpp_value<0x4E>(v);        // quad_perm [2, 3, 0, 1]
    v += dpp_value<0x141>(v);       // row_half_mirror
    v += dpp_value<0x140>(v);       // row_mirror
    v += dpp_value<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    return v;
}

// Halo columns of an aligned strip in the pair layout: the 19 pulled values of a halo site sit in
// lanes 0..18 (first site) or 32..50 (second site) of one wavefront load; lanes 19..31 and 51..63
// re-read velocity 0 of their site (a valid address) and add nothing.
struct HaloLaneLP {
    int i, dy;
    unsigned x;
    bool on;
};
__device__ __forceinline__ HaloLaneLP halo_lane_lp(int nx, int col_first, int col_second)
{
    HaloLaneLP h;
    const int lane = threadIdx.x;
    const int site = lane >> 5, j = lane & 31;
    h.on = j < Q;
    h.i = h.on ? j : 0;
    const int cx = (int)((CXP_MASK >> h.i) & 1u) - (int)((CXN_MASK >> h.i) & 1u);
    const int cy = (int)((CYP_MASK >> h.i) & 1u) - (int)((CYN_MASK >> h.i) & 1u);
    h.x = (unsigned)wrapi((site ? col_second : col_first) - cx, nx);
    h.dy = -cy;
    return h;
}

// Colour field of the two halo sites of one halo load: each lane decodes its pair as raw_phase does
// and the site's f_i and e_i f_i are summed in FP64 over its lanes.  The sum order differs from the
// owning lane's, but the terms are the same and the FP64 sums agree with it to ~1e-16, far below the
// FP32 rounding of the result.
template <typename ST>
__device__ __forceinline__ void halo_pair_phase(typename PairOf<ST>::raw v, const Ref& rf, float& first,
                                                float& second)
{
    const int j = threadIdx.x & 31;
    const bool on = j < Q;
    const int k = j == 0 ? 0 : (j < 7 ? 1 : 2);
    const float w = k == 0 ? WL0 : (k == 1 ? WL1 : WL2);
    const float isc = k == 0 ? ISCL[0] : (k == 1 ? ISCL[1] : ISCL[2]);
    const float mid = rf.mid(k) * isc, half = rf.half(k) * isc;
    const float e = PairOf<ST>::e(v);
    const float si = __builtin_fmaf(PairOf<ST>::a(v), isc, __builtin_fmaf(half, e, mid));
    const float f = si + w;
    const double D = half_wave_sum(on ? (double)e * (double)f : 0.0);
    const double F = half_wave_sum(on ? (double)f : 0.0);
    const float ph = (float)(D / F);   // meaningful in lanes 31 and 63
    first = lane_value(ph, 31);
    second = lane_value(ph, 63);
}

template <typename ST>
__device__ __forceinline__ void halo_phase(typename PairOf<ST>::raw v, const Ref& rf, float& left, float& right)
{
    halo_pair_phase<ST>(v, rf, left, right);
}

struct Nine {
    float v[3][3];
    __device__ __forceinline__ float at(int dx, int dy) const { return v[dy + 1][dx + 1]; }
};

// 3 x 3 neighbourhood of each lane's site from the lanes of rows y-1, y, y+1 (a row per lane), with the
// columns left of lane 0 and right of lane 63 taken from the strip's halo values (hl, hr per row), for
// aligned 64-column strips.
__device__ __forceinline__ Nine gather9h(float rm, float r0, float rp, const float hl[3], const float hr[3])
{
    Nine n;
    const float rows[3] = {rm, r0, rp};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        n.v[k][0] = from_left(rows[k], hl[k]);
        n.v[k][1] = rows[k];
        n.v[k][2] = from_right(rows[k], hr[k]);
    }
    return n;
}

// isotropic in-plane derivative 3 sum_i w_i c_i f(x + c_i), folded onto the slab
__device__ __forceinline__ float dx_iso(const Nine& f)
{
    return 3.f * (WAX * (f.at(1, 0) - f.at(-1, 0)) +
                  WDG * ((f.at(1, 1) - f.at(-1, -1)) + (f.at(1, -1) - f.at(-1, 1))));
}
__device__ __forceinline__ float dy_iso(const Nine& f)
{
    return 3.f * (WAX * (f.at(0, 1) - f.at(0, -1)) +
                  WDG * ((f.at(1, 1) - f.at(-1, -1)) - (f.at(1, -1) - f.at(-1, 1))));
}
// 6 sum_i w_i (f(x + c_i) - f(x))
__device__ __forceinline__ float lap_iso(const Nine& f)
{
    const float c = f.at(0, 0);
    return 6.f * (WAX * ((f.at(1, 0) + f.at(-1, 0)) + (f.at(0, 1) + f.at(0, -1)) - 4.f * c) +
                  WDG * ((f.at(1, 1) + f.at(-1, -1)) + (f.at(1, -1) + f.at(-1, 1)) - 4.f * c));
}

struct ConstF {
    float oot, A, B0, B1, B2, beta, hs, kf, bf, sf0, sf1, sf2;
};

__device__ __forceinline__ ConstF to_float(const CaseConst& c)
{
    ConstF f;
    f.oot = (float)c.one_over_tau; f.A = (float)c.A; f.B0 = (float)c.B0; f.B1 = (float)c.B1; f.B2 = (float)c.B2;
    f.beta = (float)c.beta; f.hs = (float)c.hs; f.kf = (float)c.kf; f.bf = (float)c.bf;
    f.sf0 = (float)c.sfac0; f.sf1 = (float)c.sfac1; f.sf2 = (float)c.sfac2;
    return f;
}

// Collision of one site given its pulled row, phi, the mediator gradient gm = -grad(phi) and the
// force (class F); writes the post-collision s_i and e_i through out(i, s_i, e_i).
// c_i . (x, y, z) for a velocity index that is a compile-time constant (unrolled loops): only the
// velocity's nonzero components enter (a product with a zero component cannot be folded away in
// IEEE arithmetic), added in x, y, z order, so the value is that of sum_a c_ia v_a rounded once.
__device__ __forceinline__ float cdot3(int i, float x, float y, float z)
{
    float r = 0.f;
    bool any = false;
    auto add = [&](int c, float v) {
        if (c == 0) return;
        const float t = c > 0 ? v : -v;
        r = any ? r + t : t;
        any = true;
    };
    add(CX[i], x);
    add(CY[i], y);
    add(CZ[i], z);
    return r;
}
// the in-plane part c_ix x + c_iy y (fields with no z component)
__device__ __forceinline__ float cdot2(int i, float x, float y)
{
    float r = 0.f;
    bool any = false;
    auto add = [&](int c, float v) {
        if (c == 0) return;
        const float t = c > 0 ? v : -v;
        r = any ? r + t : t;
        any = true;
    };
    add(CX[i], x);
    add(CY[i], y);
    return r;
}
// z moment sum_i c_iz v_i as a sum of z-mirror pair differences (i and i' differ in the sign of c_z
// only): exactly 0 when the values are mirror-symmetric
__device__ __forceinline__ float zmoment(const float v[Q])
{
    static_assert(CZ[5] == 1 && CZ[6] == -1 && CZ[11] == 1 && CZ[13] == -1 && CZ[14] == 1 && CZ[12] == -1 &&
                      CZ[16] == 1 && CZ[18] == -1 && CZ[17] == 1 && CZ[15] == -1,
                  "z-mirror pairs of the D3Q19 ordering");
    static_assert(CX[11] == CX[13] && CX[14] == CX[12] && CY[16] == CY[18] && CY[17] == CY[15], "mirror pairs");
    return (((v[5] - v[6]) + (v[11] - v[13])) + ((v[14] - v[12]) + (v[16] - v[18]))) + (v[17] - v[15]);
}

template <bool FORCED, bool ZSYM, class Out>
__device__ __forceinline__ void collide_lp(const ConstF& k, const Row& r, float ds, float rho, double phid, float gx,
                                           float gy, float Fx, float Fy, Out out)
{
    // momentum (sum_i c_i w_i = 0, so it is the momentum of the deviations).  The z component is 0
    // in every state that is mirror-symmetric in z (all states the programs create: one node in z,
    // so gradients and forces have no z component); it is kept for uploaded states that are not, and
    // summed over z-mirror pairs, so that it is exactly 0 in a symmetric state.
    float jx = 0.f, jy = 0.f;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        if (CX[i] > 0) jx += r.s[i];
        if (CX[i] < 0) jx -= r.s[i];
        if (CY[i] > 0) jy += r.s[i];
        if (CY[i] < 0) jy -= r.s[i];
    }
    // z-symmetric storage: the state is symmetric by construction and the z moment is 0
    const float jz = ZSYM ? 0.f : zmoment(r.s);
    const float irho = __builtin_amdgcn_rcpf(rho);
    float ux = jx * irho, uy = jy * irho;
    const float uz = ZSYM ? 0.f : jz * irho;
    if constexpr (FORCED) {
        ux += 0.5f * Fx * irho;
        uy += 0.5f * Fy * irho;
    }
    const float usq = ZSYM ? ux * ux + uy * uy : ux * ux + uy * uy + uz * uz;
    const float vF = ux * Fx + uy * Fy;
    float s[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const float cu = ZSYM ? cdot2(i, ux, uy) : cdot3(i, ux, uy, uz);
        const float seq = WF[i] * (ds + rho * (3.f * cu + 4.5f * cu * cu - 1.5f * usq));
        s[i] = r.s[i] + (seq - r.s[i]) * k.oot;
        if constexpr (FORCED) {
            const float cF = cdot2(i, Fx, Fy);
            const float sf = (i == 0) ? k.sf0 : (i < 7 ? k.sf1 : k.sf2);
            s[i] += sf * (cF + 3.f * cF * cu - vF);
        }
    }
    const float mod2 = gx * gx + gy * gy;
    const bool has_g = mod2 > 0.f;
    const float imod = has_g ? rsqrtf(mod2) : 0.f;
    const float nx_ = gx * imod, ny_ = gy * imod;
    if constexpr (!FORCED) {
        // perturbation: A |g| (w_i (c_i.n)^2 - B_i)
        if (has_g && k.A != 0.f) {
            const float Am = k.A * mod2 * imod;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const float cn = cdot2(i, nx_, ny_);
                const float B = (i == 0) ? k.B0 : (i < 7 ? k.B1 : k.B2);
                s[i] += Am * (WF[i] * cn * cn - B);
            }
        }
    }
    // Mass and momentum restored to what the update conserves exactly (mass; momentum + F).  With
    // the FP32 weights the second moment 3 sum_i w_i c_ix^2 is 1 - 6e-8 (1/3 has no binary
    // representation), so the equilibrium lost a fraction 6e-8 / tau of the momentum every step
    // (3e-3 of a translating drop's velocity over 40000 steps) and gained mass in proportion to
    // u^2.  The residuals go to the rest population (mass) and to the +-x, +-y, +-z pairs (momentum).
    {
        float m = 0.f, px = 0.f, py = 0.f;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            m += s[i];
            if (CX[i] > 0) px += s[i];
            if (CX[i] < 0) px -= s[i];
            if (CY[i] > 0) py += s[i];
            if (CY[i] < 0) py -= s[i];
        }
        const float pz = ZSYM ? 0.f : zmoment(s);
        const float dx = 0.5f * ((FORCED ? jx + Fx : jx) - px);
        const float dy = 0.5f * ((FORCED ? jy + Fy : jy) - py);
        const float dz = 0.5f * (jz - pz);
        s[0] += ds - m;
        s[1] += dx;
        s[2] -= dx;
        s[3] += dy;
        s[4] -= dy;
        if constexpr (!ZSYM) {
            s[5] += dz;
            s[6] -= dz;
        }
    }
    // recolouring: e_i = phi - 2 beta cR cB w_i rho cos_i / f_i, the difference taken in FP64 and
    // rounded once
    const float phi = (float)phid;
    const float cR = 0.5f * (1.f + phi), cB = 0.5f * (1.f - phi);
    const float kk = 2.f * k.beta * cR * cB * rho;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        float e = phi;
        if (i > 0) {   // without a gradient n = 0, so the correction is 0 (no branch per lane)
            const float f = s[i] + WF[i];
            const float cn = cdot2(i, nx_, ny_);
            const float cosp = (i < 7) ? cn : cn * ISQ2;
            const float corr = kk * WF[i] * cosp * __builtin_amdgcn_rcpf(f);
            e = (float)(phid - (double)((f != 0.f) ? corr : 0.f));
        }
        out(i, s[i], e);
    }
}

struct TileCtxLP {
    CaseConst cc;
    const void* in;
    void* out;
    gdouble* aux;
    long long S;
    int strip, y0, y1;
    bool rev;   // odd segments march from y1 - 1 down to y0 (see lp_pipe)
};

__device__ __forceinline__ bool tile_ctx(const StepLaunch& L, TileCtxLP& t)
{
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int c = L.tile_case[tile];
    const CaseDev& d = L.tab[c];
    t.cc = d.cc;
    t.in = d.pop[L.parity];
    t.out = d.pop[L.parity ^ 1];
    t.S = d.S;
    t.aux = L.record ? (gdouble*)d.aux : nullptr;
    // Within a case the tiles run segment row by segment row.  A case of a multiple of 32 strips
    // (width >= 2048) is dealt in bands of 32 strips, every segment row of a band before the next
    // band: the wavefronts an XCD runs at once then cover fewer columns and more segment rows, so
    // more of the rows and columns they share are read while in its L2 (8192^2, one case, the
    // tuner's shape: FP32 3790 -> 3734 us, FP32 z-symmetric 2795 -> 2757 us, FP16 2044 -> 2040 us;
    // bands of 8 strips were slower, 16 as fast as 32).
    constexpr int BAND = 32;
    const int local = tile - d.tile0;
    int seg;
    if (d.nstrip % BAND == 0) {
        const int band = local / (BAND * d.nseg), r = local % (BAND * d.nseg);
        seg = r / BAND;
        t.strip = band * BAND + r % BAND;
    } else {
        t.strip = local % d.nstrip;
        seg = local / d.nstrip;
    }
    t.y0 = seg * d.ly;
    t.y1 = min(t.y0 + d.ly, d.cc.ny);
    t.rev = (seg & 1) != 0;
    return t.y0 < d.cc.ny;
}

__device__ __forceinline__ void store_aux(gdouble* __restrict__ aux, long long S, unsigned site, float gx, float gy,
                                          float Fx, float Fy)
{
    (aux + 0 * S)[site] = gx; (aux + 1 * S)[site] = gy; (aux + 2 * S)[site] = 0.0;
    (aux + 3 * S)[site] = Fx; (aux + 4 * S)[site] = Fy; (aux + 5 * S)[site] = 0.0;
}

// in-plane field the class-F force differentiates: n = g/|g| (F1) or T = (sigma/2)(|g| I - g g/|g|)
// with components xx, yy, xy (F2); g = grad(phi)
struct AuxF { float g0, g1, a0, a1, a2; };

template <int KIND>
__device__ __forceinline__ AuxF aux_of(const ConstF& k, const Nine& p)
{
    AuxF a;
    a.g0 = dx_iso(p);
    a.g1 = dy_iso(p);
    // |g| and 1/|g| from one reciprocal square root (thresholds as the reference's on |g|:
    // 1e-12 for F1, 1e-30 for F2, whose square underflows in FP32 so that the test is |g|^2 > 0)
    const float m2 = a.g0 * a.g0 + a.g1 * a.g1;
    const float r = __builtin_amdgcn_rsqf(m2);
    if constexpr (KIND == KIND_F1) {
        const float inv = (m2 > 1.0e-24f) ? r : 0.f;
        a.a0 = a.g0 * inv; a.a1 = a.g1 * inv; a.a2 = 0.f;
    } else {
        const float inv = (m2 > 0.f) ? r : 0.f;
        const float m = m2 * inv;
        a.a0 = k.hs * (m - a.g0 * a.g0 * inv);
        a.a1 = k.hs * (m - a.g1 * a.g1 * inv);
        a.a2 = -k.hs * a.g0 * a.g1 * inv;
    }
    return a;
}

// Two-ring halo of an aligned strip (two-row stencils): load A holds columns c0-2 and c0-1, load B
// columns c0+64 and c0+65, in the lane layout of halo_lane_lp.
template <typename ST, bool SYM>
struct Halo2LP {
    using R = typename PairOf<ST>::raw;
    unsigned long long off[2];   // element offset of (velocity i, column x) within a row block
    int dy[2];
    __device__ __forceinline__ void init(int nx, int px, int strip)
    {
        const int c0 = strip * 64;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const HaloLaneLP h = s ? halo_lane_lp(nx, c0 + 64, c0 + 65) : halo_lane_lp(nx, c0 - 2, c0 - 1);
            off[s] = (unsigned long long)slot_of(SYM, h.i) * (unsigned)px + h.x;
            dy[s] = h.dy;
        }
    }
    __device__ __forceinline__ void load(gcst<R>* in, int px, int ny, int y, R& a, R& b) const
    {
        const unsigned long long pitch = (unsigned long long)n_slots(SYM) * (unsigned)px;
        a = in[(unsigned long long)wrap_row3(y, dy[0], ny) * pitch + off[0]];
        b = in[(unsigned long long)wrap_row3(y, dy[1], ny) * pitch + off[1]];
    }
};

// colour field of the four halo columns c0-2, c0-1, c0+64, c0+65
template <typename ST>
__device__ __forceinline__ void halo2_phase(typename PairOf<ST>::raw va, typename PairOf<ST>::raw vb, const Ref& rf,
                                            float (&hw)[4])
{
    halo_pair_phase<ST>(va, rf, hw[0], hw[1]);
    halo_pair_phase<ST>(vb, rf, hw[2], hw[3]);
}

// Collision of one site row of an aligned strip and the stores of its 19 pairs (lanes inside the
// case; a site without fluid, rho <= 1e-30, is stored unchanged), and the aux field on record steps.
template <typename ST, bool FORCED, bool SYM>
__device__ __forceinline__ void collide_store(const TileCtxLP& t, const ConstF& k, const Ref& rf, const Lane& L,
                                              gst<typename PairOf<ST>::raw>* out, int y,
                                              const typename PairOf<ST>::raw (&cur)[Q], double phi, float gx,
                                              float gy, float Fx, float Fy)
{
    using R = typename PairOf<ST>::raw;
    if (!L.interior) return;
    const int nx = t.cc.nx, px = t.cc.px;
    Row row;
    decode<ST>(cur, row, rf);
    float ds, rho;
    row_density(row, ds, rho);
    constexpr int NS = n_slots(SYM);
    const rsrc_t orow = make_rsrc(out + (long long)y * ((long long)NS * px), (unsigned)(NS * px * (int)sizeof(R)));
    const unsigned ocol = (unsigned)L.col * (unsigned)sizeof(R);
    auto st = [&](int i, float s, float e) {
        if (!is_stored(SYM, i)) return;   // equal to its z mirror's
        pair_store<R>(encode<ST>(i, s, e, rf), orow, ocol, (unsigned)(slot_of(SYM, i) * px * (int)sizeof(R)));
    };
    if (rho > 1.0e-30f) {
        collide_lp<FORCED, SYM>(k, row, ds, rho, phi, gx, gy, Fx, Fy, st);
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) st(i, row.s[i], row.e[i]);
    }
    if (t.aux) store_aux(t.aux, t.S, (unsigned)(y * nx + L.col), gx, gy, Fx, Fy);
}

// collide_store for a row whose deviations s_i were kept from its colour field (the collision uses
// only them); a site without fluid reads its pairs again to store them unchanged
template <typename ST, bool SYM>
__device__ __forceinline__ void collide_store_kept(const TileCtxLP& t, const ConstF& k, const Ref& rf, const Lane& L,
                                                   gcst<typename PairOf<ST>::raw>* in, gst<typename PairOf<ST>::raw>* out,
                                                   int y, const float (&sv)[Q], double phi, float gx, float gy,
                                                   float Fx, float Fy)
{
    using R = typename PairOf<ST>::raw;
    if (!L.interior) return;
    const int nx = t.cc.nx, ny = t.cc.ny, px = t.cc.px;
    Row row;
#pragma unroll
    for (int i = 0; i < Q; ++i) row.s[i] = sv[i];
    float ds, rho;
    row_density(row, ds, rho);
    constexpr int NS = n_slots(SYM);
    const rsrc_t orow = make_rsrc(out + (long long)y * ((long long)NS * px), (unsigned)(NS * px * (int)sizeof(R)));
    const unsigned ocol = (unsigned)L.col * (unsigned)sizeof(R);
    auto st = [&](int i, float s, float e) {
        if (!is_stored(SYM, i)) return;   // equal to its z mirror's
        pair_store<R>(encode<ST>(i, s, e, rf), orow, ocol, (unsigned)(slot_of(SYM, i) * px * (int)sizeof(R)));
    };
    if (rho > 1.0e-30f) {
        collide_lp<true, SYM>(k, row, ds, rho, phi, gx, gy, Fx, Fy, st);
    } else {
        R cur[Q];
        load_raw<ST, SYM>(in, px, ny, y, L, cur);
        decode<ST>(cur, row, rf);
#pragma unroll
        for (int i = 0; i < Q; ++i) st(i, row.s[i], row.e[i]);
    }
    if (t.aux) store_aux(t.aux, t.S, (unsigned)(y * nx + L.col), gx, gy, Fx, Fy);
}

// Aligned 64-column strip kernel for the two-row stencils (F1, F2): rows y-2 .. y+2 of the colour
// field are kept for the own lanes (P) and the four halo columns (HW); the gradient and the n / T
// field at the halo columns c0-1 and c0+64 are evaluated on the wavefront's uniform values with the
// same functions as every lane's.  Row y's populations are read again (from L2) for its collision once
// its force is known.  DEPTH sets the loads in flight: 1 loads row y + 2 in the iteration that needs
// it; 2 and 3 issue its loads 1 and 2 iterations ahead (rotating buffers, the row loop unrolled so
// every buffer role is a compile-time constant).  The re-read of row y is issued before the wait on
// row y + 2, so that the collision's wait leaves the rows ahead in flight (vector memory counts
// retire in issue order).
// REV marches from y1 - 1 down to y0 (odd segments), so that adjacent segments reach the rows they
// share at the same time (as in lp_pipe).  The window P / HW / pd holds rows in march order ("ahead"
// is y - 1 .. y - 2 then); every three-row stencil reads it through G, which gives the window index
// of geometric rows y - 1, y, y + 1, so the arithmetic per site is that of the forward march.
template <typename ST, int KIND, int DEPTH, bool SYM, bool REV>
__device__ __forceinline__ void f12_rows(const TileCtxLP& t)
{
    constexpr int NA = DEPTH - 1;         // rows loaded ahead of the iteration that uses them
    constexpr int NU = NA > 0 ? NA : 1;   // unroll of the row loop
    using R = typename PairOf<ST>::raw;
    const int nx = t.cc.nx, ny = t.cc.ny, px = t.cc.px;
    const ConstF k = to_float(t.cc);
    const Ref rf = make_lp_ref(t.cc.ref_mid, t.cc.ref_half);
    const Lane L = make_lane_aligned(nx, t.strip);
    gcst<R>* in = (gcst<R>*)t.in;
    gst<R>* out = (gst<R>*)t.out;
    Halo2LP<ST, SYM> h;
    h.init(nx, px, t.strip);
    const int y0 = t.y0, y1 = t.y1, n = y1 - y0;
    auto row = [&](int kk) { return REV ? y1 - 1 - kk : y0 + kk; };
    auto G = [](int r) { return REV ? 2 - r : r; };

    R cur[Q];
    R ahead[NU][Q];   // row(k + 2) in buffer k mod NU (DEPTH > 1)
    R ha[NU], hb[NU];
    float P[5], HW[5][4];   // rows row(k - 2) .. row(k + 2)
    double pd[3];   // FP64 colour field of rows row(k), row(k + 1), row(k + 2) (for their collision)
    auto phi_row = [&](int y, float (&hw)[4]) {
        load_raw<ST, SYM>(in, px, ny, y, L, cur);
        R va, vb;
        h.load(in, px, ny, y, va, vb);
        halo2_phase<ST>(va, vb, rf, hw);
        return raw_phase<ST>(cur, rf);
    };
    auto aux_lanes = [&](int j) {   // n / T at the centre row of window rows j .. j+2
        const float hl[3] = {HW[j + G(0)][1], HW[j + 1][1], HW[j + G(2)][1]};
        const float hr[3] = {HW[j + G(0)][2], HW[j + 1][2], HW[j + G(2)][2]};
        return aux_of<KIND>(k, gather9h(P[j + G(0)], P[j + 1], P[j + G(2)], hl, hr));
    };
    auto aux_halo = [&](int j, int side) {   // n / T at c0-1 (side 0) or c0+64 (side 1)
        Nine nn;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int w = j + G(r);
            if (side == 0) {
                nn.v[r][0] = HW[w][0];
                nn.v[r][1] = HW[w][1];
                nn.v[r][2] = lane_value(P[w], 0);
            } else {
                nn.v[r][0] = lane_value(P[w], 63);
                nn.v[r][1] = HW[w][2];
                nn.v[r][2] = HW[w][3];
            }
        }
        return aux_of<KIND>(k, nn);
    };
    P[0] = (float)phi_row(row(-2), HW[0]);
    P[1] = (float)phi_row(row(-1), HW[1]);
    pd[1] = phi_row(row(0), HW[2]);
    P[2] = (float)pd[1];
    pd[2] = phi_row(row(1), HW[3]);
    P[3] = (float)pd[2];
    AuxF am = aux_lanes(0), a0 = aux_lanes(1);
    // n / T at the halo columns of rows row(k-1) and row(k) (carried along; row(k+1)'s is new each row)
    AuxF hl[3], hr[3];
    hl[0] = aux_halo(0, 0); hr[0] = aux_halo(0, 1);
    hl[1] = aux_halo(1, 0); hr[1] = aux_halo(1, 1);
#pragma unroll
    for (int r = 0; r < NA; ++r) {
        load_raw<ST, SYM>(in, px, ny, row(2 + r), L, ahead[r]);
        h.load(in, px, ny, row(2 + r), ha[r], hb[r]);
    }

    for (int k0 = 0; k0 < n; k0 += NU) {
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const int kk = k0 + j;
            if (kk >= n) break;
            const int yy = row(kk);
            pd[0] = pd[1];
            pd[1] = pd[2];
            if constexpr (NA > 0) {
                load_raw<ST, SYM>(in, px, ny, yy, L, cur);   // row yy again (from L2), for its collision
                __builtin_amdgcn_sched_barrier(0);      // issued before the wait on row(kk + 2)
                pd[2] = raw_phase<ST>(ahead[j], rf);    // row(kk + 2) (waits for its loads)
                halo2_phase<ST>(ha[j], hb[j], rf, HW[4]);
                // row(kk + 2 + NA), used NA iterations on; past the segment's last row the loads are
                // issued empty, so every path leaves the same loads outstanding
                const bool live = kk + NA < n;
                load_raw<ST, SYM>(in, px, ny, row(kk + 2 + NA), L, ahead[j], live);
                h.load(in, px, ny, live ? row(kk + 2 + NA) : yy, ha[j], hb[j]);
                __builtin_amdgcn_sched_barrier(0);
            } else {
                pd[2] = phi_row(row(kk + 2), HW[4]);
                load_raw<ST, SYM>(in, px, ny, yy, L, cur);   // row yy again (from L2), for its collision
            }
            P[4] = (float)pd[2];
            const AuxF ap = aux_lanes(2);
            hl[2] = aux_halo(2, 0);
            hr[2] = aux_halo(2, 1);
            auto field = [&](float AuxF::*m) {
                const float l3[3] = {hl[G(0)].*m, hl[1].*m, hl[G(2)].*m};
                const float r3[3] = {hr[G(0)].*m, hr[1].*m, hr[G(2)].*m};
                return REV ? gather9h(ap.*m, a0.*m, am.*m, l3, r3)
                          : gather9h(am.*m, a0.*m, ap.*m, l3, r3);
            };
            float Fx, Fy;
            if constexpr (KIND == KIND_F1) {
                const float divn = dx_iso(field(&AuxF::a0)) + dy_iso(field(&AuxF::a1));
                Fx = -k.hs * divn * a0.g0;
                Fy = -k.hs * divn * a0.g1;
            } else {
                const Nine txy = field(&AuxF::a2);
                Fx = dx_iso(field(&AuxF::a0)) + dy_iso(txy);
                Fy = dx_iso(txy) + dy_iso(field(&AuxF::a1));
            }
            const float hl1[3] = {HW[1 + G(0)][1], HW[2][1], HW[1 + G(2)][1]};
            const float hr0[3] = {HW[1 + G(0)][2], HW[2][2], HW[1 + G(2)][2]};
            const Nine p9 = gather9h(P[1 + G(0)], P[2], P[1 + G(2)], hl1, hr0);
            const float gx = -dx_iso(p9), gy = -dy_iso(p9);   // mediator momentum = -grad(phi)
            collide_store<ST, true, SYM>(t, k, rf, L, out, yy, cur, pd[0], gx, gy, Fx, Fy);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                P[r] = P[r + 1];
#pragma unroll
                for (int q = 0; q < 4; ++q) HW[r][q] = HW[r + 1][q];
            }
            am = a0;
            a0 = ap;
            hl[0] = hl[1]; hl[1] = hl[2];
            hr[0] = hr[1]; hr[1] = hr[2];
        }
    }
}

// Depth-1 form of f12_rows without the re-read: the deviations s_i decoded for a row's colour field
// (all its collision uses) are kept in registers until that row is collided two iterations later, in
// three rotating buffers (the row loop unrolled three times, so every buffer role is a compile-time
// constant), and the next row's pairs are loaded one iteration ahead, in flight over the collision.
// The re-read of f12_rows mostly missed L2 (fetch 1.8-2.8x the algorithmic reads over eight 512^2
// cases): the rows between a strip's two reads of a row exceed an XCD's 4 MB L2 at the wavefronts in
// flight.  Same arithmetic per site as f12_rows.
template <typename ST, int KIND, bool SYM, bool REV>
__device__ __forceinline__ void f12_rows_cached(const TileCtxLP& t)
{
    using R = typename PairOf<ST>::raw;
    const int nx = t.cc.nx, ny = t.cc.ny, px = t.cc.px;
    const ConstF k = to_float(t.cc);
    const Ref rf = make_lp_ref(t.cc.ref_mid, t.cc.ref_half);
    const Lane L = make_lane_aligned(nx, t.strip);
    gcst<R>* in = (gcst<R>*)t.in;
    gst<R>* out = (gst<R>*)t.out;
    Halo2LP<ST, SYM> h;
    h.init(nx, px, t.strip);
    const int y0 = t.y0, y1 = t.y1, n = y1 - y0;
    auto row = [&](int kk) { return REV ? y1 - 1 - kk : y0 + kk; };
    auto G = [](int r) { return REV ? 2 - r : r; };

    float sk[3][Q];         // deviations of row row(kk) in buffer kk mod 3, from its colour field to its collision
    float P[5], HW[5][4];   // rows row(k - 2) .. row(k + 2)
    double pd[3];           // FP64 colour field of rows row(k), row(k + 1), row(k + 2) (for their collision)
    R ahead[Q], aha, ahb;   // pairs of the next row to consume (own lanes, halo loads)
    // with live = false the own lanes' loads read nothing; the halo loads read valid (wrapped) rows
    auto issue = [&](int y, bool live) {
        load_raw<ST, SYM>(in, px, ny, y, L, ahead, live);
        h.load(in, px, ny, y, aha, ahb);
    };
    auto consume = [&](float (&sv)[Q], float (&hw)[4]) {
        halo2_phase<ST>(aha, ahb, rf, hw);
        Row row;
        decode<ST>(ahead, row, rf);
        float ds, rho;
        double phi;
        row_phase(row, ds, rho, phi);
#pragma unroll
        for (int i = 0; i < Q; ++i) sv[i] = row.s[i];
        return phi;
    };
    auto aux_lanes = [&](int j) {   // n / T at the centre row of window rows j .. j+2
        const float hl[3] = {HW[j + G(0)][1], HW[j + 1][1], HW[j + G(2)][1]};
        const float hr[3] = {HW[j + G(0)][2], HW[j + 1][2], HW[j + G(2)][2]};
        return aux_of<KIND>(k, gather9h(P[j + G(0)], P[j + 1], P[j + G(2)], hl, hr));
    };
    auto aux_halo = [&](int j, int side) {   // n / T at c0-1 (side 0) or c0+64 (side 1)
        Nine nn;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int w = j + G(r);
            if (side == 0) {
                nn.v[r][0] = HW[w][0];
                nn.v[r][1] = HW[w][1];
                nn.v[r][2] = lane_value(P[w], 0);
            } else {
                nn.v[r][0] = lane_value(P[w], 63);
                nn.v[r][1] = HW[w][2];
                nn.v[r][2] = HW[w][3];
            }
        }
        return aux_of<KIND>(k, nn);
    };
    {
        float tmp[Q];
        issue(row(-2), true);
        P[0] = (float)consume(tmp, HW[0]);
        __builtin_amdgcn_sched_barrier(0);
        issue(row(-1), true);
        P[1] = (float)consume(tmp, HW[1]);
        __builtin_amdgcn_sched_barrier(0);
    }
    issue(row(0), true);
    pd[1] = consume(sk[0], HW[2]);
    P[2] = (float)pd[1];
    __builtin_amdgcn_sched_barrier(0);
    issue(row(1), true);
    pd[2] = consume(sk[1], HW[3]);
    P[3] = (float)pd[2];
    __builtin_amdgcn_sched_barrier(0);
    issue(row(2), true);
    AuxF am = aux_lanes(0), a0 = aux_lanes(1);
    // n / T at the halo columns of rows row(k-1) and row(k) (carried along; row(k+1)'s is new each row)
    AuxF hl[3], hr[3];
    hl[0] = aux_halo(0, 0); hr[0] = aux_halo(0, 1);
    hl[1] = aux_halo(1, 0); hr[1] = aux_halo(1, 1);

    for (int k0 = 0; k0 < n; k0 += 3) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int kk = k0 + j;
            if (kk >= n) break;
            const int yy = row(kk);
            pd[0] = pd[1];
            pd[1] = pd[2];
            pd[2] = consume(sk[(j + 2) % 3], HW[4]);
            __builtin_amdgcn_sched_barrier(0);
            issue(row(kk + 3), kk + 1 < n);   // in flight over the rest of this row
            __builtin_amdgcn_sched_barrier(0);
            P[4] = (float)pd[2];
            const AuxF ap = aux_lanes(2);
            hl[2] = aux_halo(2, 0);
            hr[2] = aux_halo(2, 1);
            auto field = [&](float AuxF::*m) {
                const float l3[3] = {hl[G(0)].*m, hl[1].*m, hl[G(2)].*m};
                const float r3[3] = {hr[G(0)].*m, hr[1].*m, hr[G(2)].*m};
                return REV ? gather9h(ap.*m, a0.*m, am.*m, l3, r3)
                          : gather9h(am.*m, a0.*m, ap.*m, l3, r3);
            };
            float Fx, Fy;
            if constexpr (KIND == KIND_F1) {
                const float divn = dx_iso(field(&AuxF::a0)) + dy_iso(field(&AuxF::a1));
                Fx = -k.hs * divn * a0.g0;
                Fy = -k.hs * divn * a0.g1;
            } else {
                const Nine txy = field(&AuxF::a2);
                Fx = dx_iso(field(&AuxF::a0)) + dy_iso(txy);
                Fy = dx_iso(txy) + dy_iso(field(&AuxF::a1));
            }
            const float hl1[3] = {HW[1 + G(0)][1], HW[2][1], HW[1 + G(2)][1]};
            const float hr0[3] = {HW[1 + G(0)][2], HW[2][2], HW[1 + G(2)][2]};
            const Nine p9 = gather9h(P[1 + G(0)], P[2], P[1 + G(2)], hl1, hr0);
            const float gx = -dx_iso(p9), gy = -dy_iso(p9);   // mediator momentum = -grad(phi)
            collide_store_kept<ST, SYM>(t, k, rf, L, in, out, yy, sk[j], pd[0], gx, gy, Fx, Fy);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                P[r] = P[r + 1];
#pragma unroll
                for (int q = 0; q < 4; ++q) HW[r][q] = HW[r + 1][q];
            }
            am = a0;
            a0 = ap;
            hl[0] = hl[1]; hl[1] = hl[2];
            hr[0] = hr[1]; hr[1] = hr[2];
        }
    }
}

template <typename ST, int KIND, int DEPTH, bool SYM>
__global__ void __launch_bounds__(64) lp_f12(StepLaunch launch)
{
    static_assert(KIND == KIND_F1 || KIND == KIND_F2, "two-row stencils only");
    static_assert(DEPTH >= 1 && DEPTH <= 3, "pipeline depth 1..3");
    TileCtxLP t;
    if (!tile_ctx(launch, t)) return;
    // FP16: the second body's registers cost more than the shared rows save (measured), so FP16
    // segments all march forward; the sites' results are the same either way
    if constexpr (DEPTH == 1) {
        if (std::is_same<ST, float>::value && t.rev) f12_rows_cached<ST, KIND, SYM, std::is_same<ST, float>::value>(t);
        else f12_rows_cached<ST, KIND, SYM, false>(t);
    } else {
        if (std::is_same<ST, float>::value && t.rev) f12_rows<ST, KIND, DEPTH, SYM, std::is_same<ST, float>::value>(t);
        else f12_rows<ST, KIND, DEPTH, SYM, false>(t);
    }
}

// Halo of an aligned strip for the pair layout: one wavefront load per row (halo_lane_lp).  The address is
// computed from a per-lane row index, apart from the strip's own loads, whose row bases stay uniform
// (scalar registers) that way.
template <typename ST, bool SYM>
struct HaloLP {
    using R = typename PairOf<ST>::raw;
    unsigned long long off;   // element offset of (velocity i, column x) within a row block
    int dy;                   // -c_iy
    bool on;
    __device__ __forceinline__ void init(int nx, int px, int strip)
    {
        const HaloLaneLP h = halo_lane_lp(nx, strip * 64 - 1, strip * 64 + 64);
        off = (unsigned long long)slot_of(SYM, h.i) * (unsigned)px + h.x;
        dy = h.dy;
        on = h.on;
    }
    __device__ __forceinline__ R load(gcst<R>* in, int px, int ny, int y) const
    {
        const int yy = wrap_row3(y, dy, ny);
        // every lane loads (the unused lanes re-read an in-range value): no branch
        return in[(unsigned long long)yy * ((unsigned long long)n_slots(SYM) * (unsigned)px) + off];
    }
};

// Collision of row y of an aligned strip, given the colour field of rows y-1, y, y+1 (own lanes and
// the two halo columns).
template <typename ST, int KIND, bool SYM>
__device__ __forceinline__ void collide_row(const TileCtxLP& t, const ConstF& k, const Ref& rf, const Lane& L,
                                            gst<typename PairOf<ST>::raw>* out, int y,
                                            const typename PairOf<ST>::raw (&cur)[Q], const float (&p)[3],
                                            const float (&hl)[3], const float (&hr)[3], double phi)
{
    const Nine p9 = gather9h(p[0], p[1], p[2], hl, hr);
    const float gx = -dx_iso(p9), gy = -dy_iso(p9);   // mediator momentum = -grad(phi)
    float Fx = 0.f, Fy = 0.f;
    if constexpr (KIND == KIND_F3) {
        const float phc = p[1];
        const float mu = 4.f * k.bf * phc * (phc * phc - 1.f) - k.kf * lap_iso(p9);
        Fx = -mu * gx;
        Fy = -mu * gy;
    }
    collide_store<ST, KIND != KIND_P, SYM>(t, k, rf, L, out, y, cur, phi, gx, gy, Fx, Fy);
}

__device__ __forceinline__ void opaque(f32x2& v)
{
    float a = v.x, b = v.y;
    asm volatile("" : "+v"(a), "+v"(b));
    v = f32x2{a, b};
}
__device__ __forceinline__ void opaque(unsigned& v) { asm volatile("" : "+v"(v)); }

// Aligned-strip kernel for the one-row stencils (P1, P3, F3), pipelined DEPTH rows deep: DEPTH + 1
// raw row buffers rotate, so while row y is collided the loads of rows up to y + DEPTH are in
// flight.  Row y + 1's colour field (own lanes and the two halo columns) completes row y's gradient.
// The row loop is unrolled DEPTH + 1 times, so every buffer role is a compile-time constant (no
// register copies).  DEPTH trades registers (occupancy) for bytes in flight per wavefront; the
// library's tuner picks it with the row-segment length.
template <typename ST, int KIND, int DEPTH, bool SYM>
__global__ void __launch_bounds__(64) lp_pipe(StepLaunch launch)
{
    static_assert(KIND == KIND_P || KIND == KIND_F3, "one-row stencils only");
    static_assert(DEPTH >= 1 && DEPTH <= 3, "pipeline depth 1..3");
    constexpr int NB = DEPTH + 1;
    using R = typename PairOf<ST>::raw;
    TileCtxLP t;
    if (!tile_ctx(launch, t)) return;
    const int nx = t.cc.nx, ny = t.cc.ny, px = t.cc.px;
    const ConstF k = to_float(t.cc);
    const Ref rf = make_lp_ref(t.cc.ref_mid, t.cc.ref_half);
    const Lane L = make_lane_aligned(nx, t.strip);
    gcst<R>* in = (gcst<R>*)t.in;
    gst<R>* out = (gst<R>*)t.out;
    HaloLP<ST, SYM> h;
    h.init(nx, px, t.strip);
    const int y0 = t.y0, y1 = t.y1, n = y1 - y0;
    // Segments march in alternating directions: even ones from y0 to y1 - 1, odd ones from y1 - 1 to
    // y0.  Two adjacent segments then reach the rows they share (each reads one row beyond its own)
    // at the same time, both at the start of their march or both at its end, while the rows are in
    // the L2 both use (adjacent tiles run on one XCD); marching the same way, one read them at its
    // start and the other a whole march later.  The k-th row visited is row(k); row(-1) and row(n)
    // are the rows just outside the segment.
    const bool rev = t.rev;
    auto row = [&](int kk) { return rev ? y1 - 1 - kk : y0 + kk; };

    // visited row k lives in buffer k mod NB; row(-1) in buffer NB - 1
    R buf[NB][Q];
    R hv[NB];
    float p[3], hl[3], hr[3];   // colour field of rows row(k-1), row(k), row(k+1)
    load_raw<ST, SYM>(in, px, ny, row(-1), L, buf[NB - 1]);
    hv[NB - 1] = h.load(in, px, ny, row(-1));
#pragma unroll
    for (int r = 0; r < DEPTH; ++r) {
        load_raw<ST, SYM>(in, px, ny, row(r), L, buf[r]);
        hv[r] = h.load(in, px, ny, row(r));
    }
    p[1] = (float)raw_phase<ST>(buf[NB - 1], rf);
    halo_phase<ST>(hv[NB - 1], rf, hl[1], hr[1]);
    double phn = raw_phase<ST>(buf[0], rf);   // phi of the next row to collide
    p[2] = (float)phn;
    halo_phase<ST>(hv[0], rf, hl[2], hr[2]);
    for (int k0 = 0; k0 < n; k0 += NB) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int kk = k0 + j;
            if (kk < n) {
                R(&cur)[Q] = buf[j];
                R(&nxt)[Q] = buf[(j + 1) % NB];
                R(&far)[Q] = buf[(j + DEPTH) % NB];   // held row(kk - 1), free now
                // rows up to row(n) feed the gradients; past it the loads are issued empty, so every
                // path leaves the same loads outstanding
                const bool live = kk + DEPTH <= n;
                load_raw<ST, SYM>(in, px, ny, row(kk + DEPTH), L, far, live);
                hv[(j + DEPTH) % NB] = h.load(in, px, ny, live ? row(kk + DEPTH) : row(kk));
                __builtin_amdgcn_sched_barrier(0);
                p[0] = p[1]; p[1] = p[2];
                hl[0] = hl[1]; hl[1] = hl[2];
                hr[0] = hr[1]; hr[1] = hr[2];
                const double phc = phn;
                phn = raw_phase<ST>(nxt, rf);   // row(kk + 1) (waits for it)
                p[2] = (float)phn;
                halo_phase<ST>(hv[(j + 1) % NB], rf, hl[2], hr[2]);
                if constexpr (DEPTH > 1) {
#pragma unroll
                    for (int i = 0; i < Q; ++i) opaque(cur[i]);   // re-decode rather than keep decoded values alive
                }
                // the stencil takes rows y - 1, y, y + 1
                const float pg[3] = {rev ? p[2] : p[0], p[1], rev ? p[0] : p[2]};
                const float hlg[3] = {rev ? hl[2] : hl[0], hl[1], rev ? hl[0] : hl[2]};
                const float hrg[3] = {rev ? hr[2] : hr[0], hr[1], rev ? hr[0] : hr[2]};
                collide_row<ST, KIND, SYM>(t, k, rf, L, out, row(kk), cur, pg, hlg, hrg, phc);
            }
        }
    }
}

template <typename ST, int KIND, bool SYM>
void launch_pipe(int depth, const StepLaunch& a, dim3 grid, hipStream_t stream)
{
    const dim3 block(64);
    if (depth <= 1) hipLaunchKernelGGL((lp_pipe<ST, KIND, 1, SYM>), grid, block, 0, stream, a);
    else if (depth == 2) hipLaunchKernelGGL((lp_pipe<ST, KIND, 2, SYM>), grid, block, 0, stream, a);
    else hipLaunchKernelGGL((lp_pipe<ST, KIND, 3, SYM>), grid, block, 0, stream, a);
}

template <typename ST, int KIND, bool SYM>
void launch_f12(int depth, const StepLaunch& a, dim3 grid, hipStream_t stream)
{
    const dim3 block(64);
    if (depth <= 1) hipLaunchKernelGGL((lp_f12<ST, KIND, 1, SYM>), grid, block, 0, stream, a);
    else if (depth == 2) hipLaunchKernelGGL((lp_f12<ST, KIND, 2, SYM>), grid, block, 0, stream, a);
    else hipLaunchKernelGGL((lp_f12<ST, KIND, 3, SYM>), grid, block, 0, stream, a);
}

template <typename ST, bool SYM>
hipError_t launch_typed(int kind, const StepLaunch& a, int ntiles, hipStream_t stream)
{
    const dim3 grid(ntiles);
    // pipeline depth: the launch's (chosen by the tuner); 0 = default, 2 rows for FP32 storage and 1
    // for FP16 storage (measured best on MI355X without tuning)
    const int depth = a.depth > 0 ? a.depth : (sizeof(typename PairOf<ST>::raw) == 8 ? 2 : 1);
    switch (kind) {
    case KIND_P: launch_pipe<ST, KIND_P, SYM>(depth, a, grid, stream); break;
    case KIND_F3: launch_pipe<ST, KIND_F3, SYM>(depth, a, grid, stream); break;
    case KIND_F1: launch_f12<ST, KIND_F1, SYM>(depth, a, grid, stream); break;
    case KIND_F2: launch_f12<ST, KIND_F2, SYM>(depth, a, grid, stream); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ conversions for transfers / diagnostics
// (h, e) pairs in rows -> fR, fB in FP64, [19][S] each (own site, post-collision); s is decoded as
// the step kernels decode it, f = w + s in FP64 (a z mirror of the symmetric storage from its slot)
template <typename ST, bool SYM>
__global__ void __launch_bounds__(256) lp_to_f64(const typename PairOf<ST>::raw* __restrict__ pop, int nx, int px,
                                                 long long S, long long n, Ref rf, double* __restrict__ fR,
                                                 double* __restrict__ fB)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long y = p / nx, x = p - y * nx;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const auto r = pop[(y * n_slots(SYM) + slot_of(SYM, i)) * px + x];
            const float e = PairOf<ST>::e(r);
            const float s = decode_s(PairOf<ST>::a(r), e, vclass(i), rf);
            const double f = (double)WF[i] + (double)s;
            fR[i * S + p] = 0.5 * f * (1.0 + (double)e);
            fB[i * S + p] = 0.5 * f * (1.0 - (double)e);
        }
    }
}

// Re-express a stored state against a new reference density, in place:
// h' = h + c_k r_old(e) - c_k r_new(e).  Pair p is slot (p / px) mod (19 or 14) of its row block (the
// pitch's padding pairs, never read, are re-expressed too).
template <typename ST, bool SYM>
__global__ void __launch_bounds__(256) lp_recentre(typename PairOf<ST>::raw* __restrict__ pop, long long npairs,
                                                   int px, Ref from, Ref to)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npairs;
         p += (long long)gridDim.x * blockDim.x) {
        const int sl = (int)((p / px) % n_slots(SYM));
        const int i = SYM ? VEL_SYM[sl] : sl;
        const int k = vclass(i);
        const auto r = pop[p];
        const float e = PairOf<ST>::e(r);
        const float A = PairOf<ST>::a(r) + __builtin_fmaf(from.half(k), e, from.mid(k));
        pop[p] = PairOf<ST>::make(A - __builtin_fmaf(to.half(k), e, to.mid(k)), e);
    }
}

// Drop initial condition written directly in the stored form (initial_conditions_drop,
// Initial_conditions.cpp:37-169, EQ_CLASSICO, cylinder, rho_R = rho_B = 1): f_R = f^eq(phi rho,
// (v0, 0, 0)), f_B = f^eq((1 - phi) rho, (v0, 0, 0)).  The stored value at site x for velocity i is
// the population that streams into x, i.e. the initial value at x + c_i (the kernel pulls).
template <typename ST, bool SYM>
__global__ void __launch_bounds__(256) lp_init_drop(typename PairOf<ST>::raw* __restrict__ pop, int nx, int px,
                                                    int ny, double R, double kp, int x0, int y0, double v0, Ref rf)
{
    const long long n = (long long)nx * ny;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const int y = (int)(p / nx), x = (int)(p - (long long)y * nx);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (!is_stored(SYM, i)) continue;   // its z mirror's value
            const int xs = wrapi(x + CX[i], nx), ys = wrapi(y + CY[i], ny);
            const double dx = (double)(xs - x0), dy = (double)(ys - y0);
            const double r = sqrt(dx * dx + dy * dy);
            double phi;
            if (kp > 0.0) {
                const double arg = fmin(fmax(kp * (r - R), -60.0), 60.0);
                phi = 1.0 / (1.0 + exp(arg));
            } else {
                phi = (r < R) ? 1.0 : 0.0;
            }
            const double w = (double)WF[i];
            const double cv = (double)CX[i] * v0;
            const double g = w * (1.0 + 3.0 * cv + 4.5 * cv * cv - 1.5 * v0 * v0);
            const double fR = g * phi, fB = g * (1.0 - phi), f = fR + fB;
            const float e = (float)((f != 0.0) ? (fR - fB) / f : 0.0);
            const int k = vclass(i);
            const float h = (float)((f - w) * (double)SCL[k]) - __builtin_fmaf(rf.half(k), e, rf.mid(k));
            pop[((long long)y * n_slots(SYM) + slot_of(SYM, i)) * px + x] = PairOf<ST>::make(h, e);
        }
    }
}

}  // namespace

hipError_t launch_step_lp(int elem, bool sym, int kind, const StepLaunch& a, int ntiles, hipStream_t stream)
{
    if (ntiles <= 0) return hipSuccess;
    if (elem == 4) return sym ? launch_typed<float, true>(kind, a, ntiles, stream) : launch_typed<float, false>(kind, a, ntiles, stream);
    if (elem == 2)
        return sym ? launch_typed<_Float16, true>(kind, a, ntiles, stream) : launch_typed<_Float16, false>(kind, a, ntiles, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_lp_to_f64(int elem, bool sym, const void* pop, int nx, int px, long long S, long long n, double rmid,
                            double rhalf, double* fR, double* fB, hipStream_t stream)
{
    const int nb = (int)std::min<long long>((n + 255) / 256, 4096);
    const Ref rf = make_lp_ref(rmid, rhalf);
    const dim3 g(nb), b(256);
    if (elem == 4) {
        if (sym) hipLaunchKernelGGL((lp_to_f64<float, true>), g, b, 0, stream, (const f32x2*)pop, nx, px, S, n, rf, fR, fB);
        else hipLaunchKernelGGL((lp_to_f64<float, false>), g, b, 0, stream, (const f32x2*)pop, nx, px, S, n, rf, fR, fB);
    } else {
        if (sym) hipLaunchKernelGGL((lp_to_f64<_Float16, true>), g, b, 0, stream, (const unsigned*)pop, nx, px, S, n, rf, fR, fB);
        else hipLaunchKernelGGL((lp_to_f64<_Float16, false>), g, b, 0, stream, (const unsigned*)pop, nx, px, S, n, rf, fR, fB);
    }
    return hipGetLastError();
}

hipError_t launch_lp_recentre(int elem, bool sym, void* pop, long long npairs, int px, double mid0, double half0,
                              double mid1, double half1, hipStream_t stream)
{
    const int nb = (int)std::min<long long>((npairs + 255) / 256, 8192);
    const Ref from = make_lp_ref(mid0, half0), to = make_lp_ref(mid1, half1);
    const dim3 g(nb), b(256);
    if (elem == 4) {
        if (sym) hipLaunchKernelGGL((lp_recentre<float, true>), g, b, 0, stream, (f32x2*)pop, npairs, px, from, to);
        else hipLaunchKernelGGL((lp_recentre<float, false>), g, b, 0, stream, (f32x2*)pop, npairs, px, from, to);
    } else {
        if (sym) hipLaunchKernelGGL((lp_recentre<_Float16, true>), g, b, 0, stream, (unsigned*)pop, npairs, px, from, to);
        else hipLaunchKernelGGL((lp_recentre<_Float16, false>), g, b, 0, stream, (unsigned*)pop, npairs, px, from, to);
    }
    return hipGetLastError();
}

hipError_t launch_lp_init_drop(int elem, bool sym, void* pop, int nx, int px, int ny, double R, double kp, int x0, int y0,
                               double v0, double rmid, double rhalf, hipStream_t stream)
{
    const long long n = (long long)nx * ny;
    const int nb = (int)std::min<long long>((n + 255) / 256, 8192);
    const Ref rf = make_lp_ref(rmid, rhalf);
    const dim3 g(nb), b(256);
    if (elem == 4) {
        if (sym) hipLaunchKernelGGL((lp_init_drop<float, true>), g, b, 0, stream, (f32x2*)pop, nx, px, ny, R, kp, x0, y0, v0, rf);
        else hipLaunchKernelGGL((lp_init_drop<float, false>), g, b, 0, stream, (f32x2*)pop, nx, px, ny, R, kp, x0, y0, v0, rf);
    } else {
        if (sym)
            hipLaunchKernelGGL((lp_init_drop<_Float16, true>), g, b, 0, stream, (unsigned*)pop, nx, px, ny, R, kp, x0, y0, v0, rf);
        else
            hipLaunchKernelGGL((lp_init_drop<_Float16, false>), g, b, 0, stream, (unsigned*)pop, nx, px, ny, R, kp, x0, y0, v0, rf);
    }
    return hipGetLastError();
}

}  // namespace itolbm
