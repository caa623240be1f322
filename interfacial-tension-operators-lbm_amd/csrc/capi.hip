// C ABI (include/itolbm.h): context, device memory, state transfer, stepping and the host side
// of the diagnostics.  Host arithmetic here is compiled with -ffp-contract=off and keeps the
// reference's expressions, so the closing formulas of the diagnostics round as the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "diag_kernels.h"
#include "itolbm.h"
#include "lattice.h"
#include "step_kernels.h"

using namespace itolbm;

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int DIAG_THREADS = 256;
constexpr int N_ANG = 256;

struct CaseState {
    lbm_case_desc d{};
    CaseConst cc{};
    long long n = 0, S = 0;
    double* pop[2] = {nullptr, nullptr};   // FP64: row blocks [y][colour][19][nx] or the compact form [y][CSLOTS][nx];
                                            // FP32/FP16: pairs ((y*NS + slot)*nx + x)
    int cur = 0;                            // buffer holding the latest post-collision state
    bool compact = false;                   // FP64 compact form (P1/P3 and F3 cases without solid sites)
    double* aux = nullptr;                  // [6][S] gm, F
    bool aux_valid = false;
    int kind = 0;                           // kernel family (KIND_*)
    int elem = 8;                           // bytes per stored value: 8 (FP64), 4 (FP32), 2 (FP16)
    bool sym = false;                       // FP32/FP16 z-symmetric storage: 14 of the 19 populations
    int nslot = Q;                          // stored pairs per site (19, or 14 with sym)
    bool active = true;                     // advanced by lbm_step
    double* f64 = nullptr;                  // [2][19][S] FP64 copy of a low-precision state (diagnostics)
    // diagnostics scratch (allocated on first use)
    double* gal = nullptr;
    double* dsum = nullptr;
    double* partial = nullptr;
    double* hist = nullptr;
    unsigned long long* histn = nullptr;
    double* tab = nullptr;                  // cosx[nx] sinx[nx] cosy[ny] siny[ny] ct[256] st[256] r[256]
    double* ordered = nullptr;              // [S + 8]: per-site phi terms, then the site-ordered sums
    double* prof = nullptr;
    int n_casca = 0;
    int nblk = 1;
    // solid sites (lbm_set_geometry): host and device masks (1 = solid), the wall's wettability and
    // the site-parallel kernels' field arrays phi [S], g [3][S], nt [6][S]
    std::vector<unsigned char> solid_h;
    unsigned char* solid = nullptr;
    double wett = 0.0;
    double* geo_f = nullptr;
};

// The group of every strip-kernel case of an FP64 context, stepped by one launch (step_fused).
constexpr int KIND_FUSED = 100;

// A set of cases advanced by one launch per step (cases of one kernel family).
struct Group {
    int kind = 0;
    std::vector<int> members;
    CaseDev* dtab = nullptr;       // device copy of the case table
    int* dtile = nullptr;          // tile -> index into members
    std::vector<int> ly;           // rows per segment of each member in this group
    std::vector<int> cur0;         // members' state buffer index when the table was written
    std::vector<double*> aux0;     // members' aux pointer when the table was written
    int k = 0;                     // launches since the table was written
    int ntiles = 0;
    int cap = 0;                   // entries dtab can hold
    int tcap = 0;                  // entries dtile can hold
    bool written = false;
    bool tuned = false;            // tile shape chosen by timing (online_tune)
    int depth = 0;                 // pipeline depth of the FP32/FP16 kernels (0 = default)
    long long version = 0;         // bumped whenever the table is rewritten
};

// A captured sequence of `steps` steps of all active families (lbm_step), replayed with one
// hipGraphLaunch.  It is valid while every family it covers still has the same table (version,
// tile count) and parity; steps is even, so a replay returns every family to its parity.
struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int steps = 0;
    std::vector<long long> key;
};

}  // namespace

struct lbm_ctx {
    std::vector<CaseState> cases;
    std::string err;
    hipStream_t stream = nullptr;      // stream in use
    hipStream_t own_stream = nullptr;  // the context's own stream
    hipEvent_t ev[2] = {nullptr, nullptr};   // autotuner timing
    hipEvent_t fork = nullptr;                 // lbm_step: families run on side streams
    hipStream_t side[N_KINDS] = {};
    hipEvent_t join[N_KINDS] = {};
    bool autotune = true;
    int f64_compact = -1;              // FP64 P1/P3 and F3 cases' storage form: -1 chosen per launch by its size,
                                       // 0 always the full form, 1 always the compact one (ITOLBM_F64_COMPACT)
    int graph_steps = 32;              // steps per captured graph in lbm_step (0: no graphs)
    StepGraph graph;
    long long bytes = 0;
    int device = 0;                    // the device current at lbm_create
    void* scratch = nullptr;           // that device's static scratch array (lbm_step_case_noalloc)
    Group family[N_KINDS];             // all cases of each kernel family (lbm_step)
    std::vector<Group> single;         // one group per case (lbm_step_case)
    // lbm_step's launch form for the strip families (FP64): -1 not chosen yet, 0 one launch per family
    // (on forked streams), 1 one launch of all of them (a wavefront per tile), 2 the same launch with
    // a grid of the resident capacity (persistent); chosen by timing, or ITOLBM_FUSED
    int fuse_mode = -1;
    Group fused;                       // the table of that one launch (kind KIND_FUSED)
};

namespace {

int fail(lbm_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, LBM_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

int dev_alloc(lbm_ctx* c, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(c, LBM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->bytes += (long long)bytes;
    return LBM_OK;
}

bool valid_op(int op) { return op == OP_P1 || op == OP_P3 || op == OP_F1 || op == OP_F2 || op == OP_F3; }

// Constants of one case, formed with the reference's expressions (coef_B Operadores_tensao.cpp:94-109,
// calcula_forca :239-241, source LBM_functions.cpp:1641, coll_BGK Collision_operators.cpp:24).
CaseConst make_const(const lbm_case_desc& d)
{
    CaseConst c{};
    c.op = d.op;
    c.nx = d.nx;
    c.ny = d.ny;
    c.one_over_tau = 1.0 / d.tau;
    c.A = d.A;
    if (d.op == OP_P1) {
        c.B0 = W0 / 3.0; c.B1 = W1 / 3.0; c.B2 = W2 / 3.0;
    } else {
        const double chi = d.chi;
        const double den = 6.0 * chi + 24.0;
        c.B0 = -(2.0 + 2.0 * chi) / (3.0 * chi + 12.0);
        c.B1 = chi / den;
        c.B2 = 1.0 / den;
    }
    c.beta = d.beta;
    const double sigma = d.A;
    c.hs = 0.5 * sigma;
    c.kf = 0.75 * sigma * d.xi;
    c.bf = 3.0 * sigma / (8.0 * d.xi);
    const double factor = (1. - 1. / (2. * d.tau));
    c.sfac0 = factor * W0 * 3.0;
    c.sfac1 = factor * W1 * 3.0;
    c.sfac2 = factor * W2 * 3.0;
    return c;
}

// Elements added to the stride S of the [k][S] site arrays (the aux fields and the FP64 [19][S]
// views of transfers and diagnostics).  0 by default: the step kernels use row blocks, whose
// streams the row pitch already spreads over HBM channels.  ITOLBM_PAD sets it: a whole number of
// elements in [0, 2^20]; anything else (negative, not a number, trailing text) is refused, and
// lbm_create returns LBM_E_ARG.  Returns -1 for a refused value.
constexpr long long MAX_STRIDE_PAD = 1LL << 20;
long long stride_pad()
{
    const char* e = std::getenv("ITOLBM_PAD");
    if (!e || !*e) return 0;
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(e, &end, 10);
    if (errno != 0 || end == e || *end != '\0' || v < 0 || v > MAX_STRIDE_PAD) {
        std::fprintf(stderr, "itolbm: ITOLBM_PAD=\"%s\" refused (a whole number of elements in [0, %lld])\n", e,
                     MAX_STRIDE_PAD);
        return -1;
    }
    return v;
}

// A launch of FP64 P1/P3 or F3 cases covering this many sites or more runs them in the compact form
// (24 stored doubles per site instead of 38), a smaller one in the full form.  The compact kernel moves
// 37 % fewer bytes but makes a longer chain of dependent work per row, which a launch of few wavefronts
// cannot hide: measured per step, eight 192^2 cases (295 k sites) 39.2 us compact against 44.3 us full,
// one 448^2 case (201 k) 32.1 against 33.7 us, one 384^2 case (147 k) 28.5 against 24.8 us, one 64^2
// case 13.6 against 7.3 us.
constexpr long long COMPACT_MIN_SITES = 196608;

// Elements added to the slot pitch of the FP64 compact form (px = nx + pad), from ITOLBM_CPAD: a whole
// number in [0, 4096] (-1: refused, lbm_create fails), or -2 when unset (compact_pad_of decides).
long long compact_pad()
{
    const char* e = std::getenv("ITOLBM_CPAD");
    if (!e || !*e) return -2;
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(e, &end, 10);
    if (errno != 0 || end == e || *end != '\0' || v < 0 || v > 4096) {
        std::fprintf(stderr, "itolbm: ITOLBM_CPAD=\"%s\" refused (a whole number of elements in [0, 4096])\n", e);
        return -1;
    }
    return v;
}

// The pad of the slot pitch of an nx-column case whose slots hold `slot_bytes`-byte elements (FP64
// compact and FP32 forms 8, FP16 4): ITOLBM_CPAD if set, else 64 elements when the pitch is a multiple
// of 64 KiB, 0 otherwise.  With such a pitch every slot of a row block falls on the same offset of a
// 64 KiB window: the 8192^2 FP64 compact case ran 5.67-5.95 ms per step unpadded and 5.06-5.37 ms with
// the pad (8- to 128-row segments), while 512^2 (187.2 us per eight-case launch unpadded, 187.6-191.8 us
// padded by 8-128), 2048^2 (0.316 against 0.333 ms) and 4096^2 (1.256 against 1.289 ms) ran best
// unpadded.
int compact_pad_of(long long env_pad, int nx, int slot_bytes)
{
    if (env_pad >= 0) return (int)env_pad;
    return ((long long)nx * slot_bytes % 65536 == 0) ? 64 : 0;
}

bool want_compact(const lbm_ctx* c, long long launch_sites)
{
    if (c->f64_compact >= 0) return c->f64_compact == 1;
    return launch_sites >= COMPACT_MIN_SITES;
}

// Whether the context's stream is capturing a HIP graph (a caller's capture).  Work that allocates,
// frees, copies from host memory or synchronises must not be recorded into a graph: a captured
// host-to-device copy keeps the address of a host buffer that is gone when the graph replays, and a
// captured free leaves the graph holding memory that is no longer the context's.  An error from the
// query counts as capturing.
bool capturing(lbm_ctx* c)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &st) != hipSuccess) return true;
    return st != hipStreamCaptureStatusNone;
}
int refuse_capture(lbm_ctx* c, const char* what)
{
    return fail(c, LBM_E_STATE, std::string(what) + " is not allowed while the context's stream is capturing a graph");
}

// Strips per row: aligned 64-column strips (halo columns come from separate loads).
int strips_of(const lbm_case_desc& d) { return (d.nx + 63) / 64; }

// Rows per wavefront segment for a group before tuning: enough segments for >= 4096 wavefronts in
// the launch, within 1..64 rows (fewer rows = more re-read halo rows).  ITOLBM_LY overrides.
int group_rows(const std::vector<CaseState>& cs, const std::vector<int>& members)
{
    if (const char* e = std::getenv("ITOLBM_LY")) {
        const int v = std::atoi(e);
        if (v > 0) return v;
    }
    long long strip_rows = 0;
    for (int id : members) strip_rows += (long long)strips_of(cs[id].d) * cs[id].d.ny;
    long long ly = strip_rows / 4096;
    if (ly < 1) ly = 1;
    if (ly > 64) ly = 64;
    return (int)ly;
}

const int CXh[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int CYh[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1};
const int CZh[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1};
const int OPPh[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};   // c_opp = -c_i

inline long long nbr(int x, int y, int dx, int dy, int nx, int ny)
{
    int xx = (x + dx) % nx, yy = (y + dy) % ny;
    if (xx < 0) xx += nx;
    if (yy < 0) yy += ny;
    return (long long)xx + (long long)nx * yy;
}

int check_case(lbm_ctx* c, int id)
{
    if (!c) return LBM_E_ARG;
    if (id < 0 || id >= (int)c->cases.size()) return fail(c, LBM_E_ARG, "case id out of range");
    return LBM_OK;
}

int ensure_aux(lbm_ctx* c, CaseState& k)
{
    if (k.aux) return LBM_OK;
    if (capturing(c)) return refuse_capture(c, "allocating the aux fields (the first LBM_STEP_RECORD_AUX step)");
    int r = dev_alloc(c, (void**)&k.aux, sizeof(double) * 6 * (size_t)k.S);
    if (r) return r;
    HIPCHK(c, hipMemsetAsync(k.aux, 0, sizeof(double) * 6 * (size_t)k.S, c->stream));
    return LBM_OK;
}

// The drop's area and centroid sums run in the reference's site order on one wavefront, whatever the
// box size: bitwise equal sums, hence the same centre, radius and in / out region masks as the
// reference (a block-tree order moves them by ~1e-16 relative, enough to move a site that lies on a
// region boundary to the other side and flip delta_p).  Cost ~1 ns per site (67 ms at 8192^2, once
// per diagnostic).

int ensure_diag(lbm_ctx* c, CaseState& k)
{
    if (k.gal) return LBM_OK;
    const int nx = k.d.nx, ny = k.d.ny;
    k.nblk = (int)std::min<long long>((k.n + DIAG_THREADS - 1) / DIAG_THREADS, 1024);
    k.n_casca = (int)(std::sqrt((double)(nx * nx + ny * ny + 1)) / 0.5) + 2;
    const int np = std::max(std::max(P1_NP, P2_NP), I1_NP);
    int r;
    if ((r = dev_alloc(c, (void**)&k.gal, sizeof(double) * 4 * (size_t)k.S))) return r;
    if ((r = dev_alloc(c, (void**)&k.dsum, sizeof(double) * 2 * (size_t)k.S))) return r;
    if ((r = dev_alloc(c, (void**)&k.ordered, sizeof(double) * ((size_t)k.S + 8)))) return r;
    if ((r = dev_alloc(c, (void**)&k.partial, sizeof(double) * (size_t)k.nblk * np))) return r;
    if ((r = dev_alloc(c, (void**)&k.hist, sizeof(double) * (size_t)k.n_casca))) return r;
    if ((r = dev_alloc(c, (void**)&k.histn, sizeof(unsigned long long) * (size_t)k.n_casca))) return r;
    const size_t ntab = 2 * (size_t)nx + 2 * (size_t)ny + 3 * N_ANG;
    if ((r = dev_alloc(c, (void**)&k.tab, sizeof(double) * ntab))) return r;
    // trigonometric tables from the host's libm, as the reference evaluates them
    // (Other_functions.cpp:72-73,175-176; Operadores_tensao.cpp:419)
    std::vector<double> t(ntab, 0.0);
    const double tx = 2.0 * PI / (double)nx, ty = 2.0 * PI / (double)ny;
    for (int x = 0; x < nx; ++x) { t[x] = std::cos(tx * x); t[nx + x] = std::sin(tx * x); }
    for (int y = 0; y < ny; ++y) { t[2 * nx + y] = std::cos(ty * y); t[2 * nx + ny + y] = std::sin(ty * y); }
    for (int kk = 0; kk < N_ANG; ++kk) {
        const double th = 2.0 * M_PI * kk / N_ANG;
        t[2 * nx + 2 * ny + kk] = std::cos(th);
        t[2 * nx + 2 * ny + N_ANG + kk] = std::sin(th);
    }
    HIPCHK(c, hipMemcpy(k.tab, t.data(), sizeof(double) * ntab, hipMemcpyHostToDevice));
    return LBM_OK;
}

template <int NP>
int reduce_partials(lbm_ctx* c, const CaseState& k, std::vector<double>& out, int ns, int nxm, int nn)
{
    std::vector<double> h((size_t)k.nblk * NP);
    HIPCHK(c, hipMemcpyAsync(h.data(), k.partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out.assign(NP, 0.0);
    for (int j = 0; j < ns; ++j) {
        double s = 0.0;
        for (int b = 0; b < k.nblk; ++b) s = s + h[(size_t)b * NP + j];
        out[j] = s;
    }
    for (int j = ns; j < ns + nxm; ++j) {
        double m = h[j];
        for (int b = 1; b < k.nblk; ++b) { const double v = h[(size_t)b * NP + j]; m = (m < v) ? v : m; }
        out[j] = m;
    }
    for (int j = ns + nxm; j < ns + nxm + nn; ++j) {
        double m = h[j];
        for (int b = 1; b < k.nblk; ++b) { const double v = h[(size_t)b * NP + j]; m = (v < m) ? v : m; }
        out[j] = m;
    }
    return LBM_OK;
}

inline double gi_image(double dx, double length) { return dx - length * std::floor(dx / length + 0.5); }

long long gcdll(long long a, long long b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { const long long t = a % b; a = b; b = t; }
    return a;
}

// (fR, fB) [colour][19][S] FP64 at own site -> pairs (h_i, e_i) at ((y * NS + slot) nx + x) in the case's
// storage type (the z-symmetric storage keeps the mirror pairs' mean after checking that they agree): e_i = (fR - fB) / (fR + fB), h_i = 2^k (f_i - w_i) - c_k r(e_i) with the case's reference
// density deviation r(e) = ref_mid + ref_half e (lattice.h).  The FP16 form keeps h as a half and e as the signed
// minority fraction copysign((1 - |e|) / 2, e) as a half (the device encoding of the same pair).
int upload_lp(lbm_ctx* c, CaseState& k, const std::vector<double>& h)
{
    const int nx = k.d.nx, ns = k.nslot, px = k.cc.px;
    const LpRef rf = make_lp_ref(k.cc.ref_mid, k.cc.ref_half);
    std::vector<float> a(2 * (size_t)ns * px * k.d.ny, 0.0f);   // rows of the padded pitch px
    // z-symmetric storage: a state whose mirror populations differ beyond rounding cannot be held
    if (k.sym)
        for (int i = 0; i < Q; ++i) {
            const int m = MIRROR_OF[i];
            if (m == i) continue;
            for (int col = 0; col < 2; ++col)
                for (long long p = 0; p < k.n; ++p) {
                    const double u = h[(size_t)(col * Q + i) * k.S + p], v = h[(size_t)(col * Q + m) * k.S + p];
                    if (!(std::fabs(u - v) <= 1e-12 * std::fmax(std::fabs(u), std::fabs(v)) + 1e-300))
                        return fail(c, LBM_E_ARG, "state is not mirror-symmetric in z (z-symmetric storage holds 14 of "
                                                  "the 19 populations); use LBM_PREC_F32 or LBM_PREC_F16");
                }
        }
    for (long long p = 0; p < k.n; ++p) {
        const long long y = p / nx, x = p - y * nx;
        for (int i = 0; i < Q; ++i) {
            if (!is_stored(k.sym, i)) continue;
            const int cl = vclass(i);
            const double w = (i == 0) ? WL0 : (i < 7 ? WL1 : WL2);   // the weights the device decodes with
            double r = h[(size_t)i * k.S + p], b = h[(size_t)(Q + i) * k.S + p];
            if (k.sym) {
                for (int m = 0; m < Q; ++m)
                    if (m != i && MIRROR_OF[m] == i) {
                        r = 0.5 * (r + h[(size_t)m * k.S + p]);
                        b = 0.5 * (b + h[(size_t)(Q + m) * k.S + p]);
                    }
            }
            const double f = r + b;
            const size_t j = 2 * (size_t)((y * ns + slot_of(k.sym, i)) * px + x);
            const float e = (float)((f != 0.0) ? (r - b) / f : 0.0);
            a[j] = (float)((f - w) * (double)SCL[cl]) - std::fma(rf.half(cl), e, rf.mid(cl));   // as the device
            a[j + 1] = e;
        }
    }
    if (k.elem == 4) {
        HIPCHK(c, hipMemcpyAsync(k.pop[k.cur], a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<uint32_t> hh(a.size() / 2);
        for (size_t j = 0; j < hh.size(); ++j) {
            const float e = a[2 * j + 1];
            const float m = std::fmax(0.5f * (1.0f - std::fabs(e)), 0.0f);
            const _Float16 lo = (_Float16)a[2 * j], hi = (_Float16)std::copysign(m, e);
            uint16_t l16, h16;
            std::memcpy(&l16, &lo, 2);
            std::memcpy(&h16, &hi, 2);
            hh[j] = (uint32_t)l16 | ((uint32_t)h16 << 16);
        }
        HIPCHK(c, hipMemcpyAsync(k.pop[k.cur], hh.data(), sizeof(uint32_t) * hh.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return LBM_OK;
}

// FP64 view [colour][19][S] of the current post-collision state (device pointer)
int state_f64(lbm_ctx* c, CaseState& k, const double** out)
{
    if (!k.f64) {
        if (int r = dev_alloc(c, (void**)&k.f64, sizeof(double) * 2 * Q * (size_t)k.S)) return r;
    }
    if (k.elem == 8) {
        HIPCHK(c, launch_f64_rows_to_soa((const double*)k.pop[k.cur], k.compact, k.cc.beta, k.d.nx, k.cc.px, k.n, k.S,
                                         k.f64, c->stream));
    } else {
        HIPCHK(c, launch_lp_to_f64(k.elem, k.sym, k.pop[k.cur], k.d.nx, k.cc.px, k.S, k.n, k.cc.ref_mid, k.cc.ref_half, k.f64,
                                   k.f64 + Q * k.S, c->stream));
    }
    *out = k.f64;
    return LBM_OK;
}

}  // namespace

extern "C" {

const char* lbm_last_error(const lbm_ctx* ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

long long lbm_device_bytes(const lbm_ctx* ctx) { return ctx ? ctx->bytes : 0; }

int lbm_create(const lbm_case_desc* cases, int n_cases, int precision, lbm_ctx** out)
{
    if (!out) return LBM_E_ARG;
    *out = nullptr;
    if (!cases || n_cases < 1) return LBM_E_ARG;
    if (precision < LBM_PREC_F64 || precision > LBM_PREC_F16_ZSYM) return LBM_E_ARG;
    const int elem = (precision == LBM_PREC_F64) ? 8 : ((precision == LBM_PREC_F32 || precision == LBM_PREC_F32_ZSYM) ? 4 : 2);
    const bool sym = precision == LBM_PREC_F32_ZSYM || precision == LBM_PREC_F16_ZSYM;
    lbm_ctx* c = new (std::nothrow) lbm_ctx();
    if (!c) return LBM_E_NOMEM;
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) {
        delete c;
        return LBM_E_HIP;
    }
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->own_stream) != hipSuccess) {
        delete c;
        return LBM_E_HIP;
    }
    c->stream = c->own_stream;
    // the allocation-free step's scratch array on this device (its address, looked up once per device)
    if (scratch_setup(c->device, &c->scratch) != hipSuccess) {
        lbm_destroy(c);
        return LBM_E_HIP;
    }
    const long long pad = stride_pad(), cpad = compact_pad();
    if (pad < 0 || cpad == -1) {
        lbm_destroy(c);
        return LBM_E_ARG;
    }
    if (const char* e = std::getenv("ITOLBM_F64_COMPACT")) {
        if (!std::strcmp(e, "0")) c->f64_compact = 0;
        else if (!std::strcmp(e, "1")) c->f64_compact = 1;
    }
    c->cases.resize(n_cases);
    for (int i = 0; i < n_cases; ++i) {
        const lbm_case_desc& d = cases[i];
        if (d.nx < 1 || d.ny < 1 || !valid_op(d.op) || !(d.tau > 0.0) || !std::isfinite(d.A) ||
            !std::isfinite(d.beta) || (d.op == OP_F3 && !(d.xi != 0.0)) ||
            (long long)d.nx * d.ny > (1LL << 31) / Q ||
            ((long long)d.nx + compact_pad_of(cpad, d.nx, 8)) * F64_SLOTS * (long long)sizeof(double) >= (1LL << 31)) {   // a row block's byte
                                                                                       // offsets are 32-bit
            lbm_destroy(c);
            return LBM_E_ARG;
        }
        CaseState& k = c->cases[i];
        k.d = d;
        k.cc = make_const(d);
        k.n = (long long)d.nx * d.ny;
        k.S = (k.n + 63) / 64 * 64 + pad;
        k.kind = step_kind(d.op);
        k.elem = elem;
        k.sym = sym;
        k.nslot = n_slots(sym);
        k.compact = false;   // the form of the first launch is set below
        // slot pitch of the FP64 compact form (8-byte slots) and of the FP32 / FP16 forms (8- / 4-byte pairs)
        k.cc.px = d.nx + compact_pad_of(cpad, d.nx, elem == 2 ? 4 : 8);
        // state buffers hold 2 x nslot values per site in row blocks (the step kernels index them by
        // nx and ny only), so their size follows from n and not from the padded stride S; an FP64
        // buffer holds either form (F64_SLOTS doubles per site of the padded pitch px), so a case can
        // change form in place
        const size_t pop_bytes = elem == 8 ? sizeof(double) * F64_SLOTS * (size_t)k.cc.px * (size_t)d.ny
                                           : (size_t)elem * 2 * k.nslot * (size_t)k.cc.px * (size_t)d.ny;
        for (int b = 0; b < 2; ++b) {
            if (dev_alloc(c, (void**)&k.pop[b], pop_bytes) || hipMemsetAsync(k.pop[b], 0, pop_bytes, c->stream) != hipSuccess) {
                lbm_destroy(c);
                return LBM_E_NOMEM;
            }
        }
    }
    for (int kd = 0; kd < N_KINDS; ++kd) c->family[kd].kind = kd;
    c->fused.kind = KIND_FUSED;
    // FP64 P1/P3 and F3 cases start in the form lbm_step's launch of their family would use
    for (CaseState& k : c->cases) {
        if (k.elem != 8 || (k.kind != KIND_P && k.kind != KIND_F3)) continue;
        long long sites = 0;
        for (const CaseState& o : c->cases) sites += o.kind == k.kind ? o.n : 0;
        k.compact = want_compact(c, sites);
    }
    c->single.resize(n_cases);
    for (int i = 0; i < n_cases; ++i) {
        c->family[c->cases[i].kind].members.push_back(i);
        c->single[i].kind = c->cases[i].kind;
        c->single[i].members.push_back(i);
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        lbm_destroy(c);
        return LBM_E_HIP;
    }
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (c->stream && capturing(c)) return refuse_capture(c, "lbm_destroy");
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (CaseState& k : c->cases) {
        for (double* p : {k.pop[0], k.pop[1], k.aux, k.gal, k.dsum, k.ordered, k.partial, k.hist, k.tab, k.prof, k.f64,
                          k.geo_f})
            if (p) (void)hipFree(p);
        if (k.histn) (void)hipFree(k.histn);
        if (k.solid) (void)hipFree(k.solid);
    }
    auto free_group = [](Group& g) {
        if (g.dtab) (void)hipFree(g.dtab);
        if (g.dtile) (void)hipFree(g.dtile);
    };
    for (Group& g : c->family) free_group(g);
    for (Group& g : c->single) free_group(g);
    free_group(c->fused);
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    if (c->graph.graph) (void)hipGraphDestroy(c->graph.graph);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->fork) (void)hipEventDestroy(c->fork);
    for (int k = 0; k < N_KINDS; ++k) {
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
        if (c->join[k]) (void)hipEventDestroy(c->join[k]);
    }
    delete c;
    return LBM_OK;
}

int lbm_upload_state(lbm_ctx* c, int id, const double* fR, const double* fB)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_upload_state");
    if (!fR || !fB) return fail(c, LBM_E_ARG, "null population array");
    CaseState& k = c->cases[id];
    const int nx = k.d.nx, ny = k.d.ny;
    // device holds "post-collision at own site": f*_i(x) = f_i(x + c_i) for a post-stream state f;
    // with solid sites, f*_i(x) = f_opp(i)(x) where x + c_i is solid (it bounced back), and solid sites
    // hold nothing (0)
    std::vector<double> h(2 * Q * (size_t)k.S, 0.0);
    const unsigned char* solid = k.solid_h.empty() ? nullptr : k.solid_h.data();
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const long long site = (long long)x + (long long)nx * y;
            if (solid && solid[site]) continue;
            for (int i = 0; i < Q; ++i) {
                long long src = nbr(x, y, CXh[i], CYh[i], nx, ny);
                int j = i;
                if (solid && solid[src]) {
                    src = site;
                    j = OPPh[i];
                }
                h[(size_t)i * k.S + site] = fR[(size_t)src * Q + j];
                h[(size_t)(Q + i) * k.S + site] = fB[(size_t)src * Q + j];
            }
        }
    if (k.elem == 8 && k.compact) {
        // compact row blocks [y][CSLOTS][px], every site raw: f_R,i in slot i, f_B,i in slot cs_raw_b(i)
        const size_t px = (size_t)k.cc.px;
        std::vector<double> rows((size_t)CSLOTS * px * ny, 0.0);
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t p = (size_t)y * nx + x;
                double* blk = rows.data() + (size_t)y * CSLOTS * px + x;
                for (int i = 0; i < Q; ++i) {
                    blk[(size_t)i * px] = h[(size_t)i * k.S + p];
                    blk[(size_t)cs_raw_b(i) * px] = h[(size_t)(Q + i) * k.S + p];
                }
                blk[(size_t)CS_RHO * px] = CS_RAW;
            }
        HIPCHK(c, hipMemcpyAsync(k.pop[k.cur], rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else if (k.elem == 8) {
        // row blocks [y][colour][19][nx]
        std::vector<double> rows(2 * Q * (size_t)k.n);
        for (int y = 0; y < ny; ++y)
            for (int j = 0; j < 2 * Q; ++j)
                for (int x = 0; x < nx; ++x)
                    rows[((size_t)y * 2 * Q + j) * nx + x] = h[(size_t)j * k.S + (size_t)y * nx + x];
        HIPCHK(c, hipMemcpyAsync(k.pop[k.cur], rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        if (int r = upload_lp(c, k, h)) return r;
    }
    k.aux_valid = false;
    return LBM_OK;
}

namespace {

// dist_eq (LBM_functions.cpp:924-944), host side
void host_feq(double* feq, double vx, double vy, double vz, double rho)
{
    static const double Ch[Q][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                                    {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
                                    {-1, 1, 0}, {1, 0, 1},   {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                                    {0, -1, -1}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1}};
    const double v[3] = {vx, vy, vz};
    const double vquad = (vx * vx + vy * vy + vz * vz);
    const double o = 3.0;
    for (int i = 0; i < Q; i++) {
        double cv = 0.0;
        for (int a = 0; a < 3; a++) cv = cv + Ch[i][a] * v[a];
        const double w = (i == 0) ? W0 : (i < 7 ? W1 : W2);
        feq[i] = w * rho * (1.0 + cv * o + 0.5 * cv * cv * o * o - 0.5 * vquad * o);
    }
}

// calcula (LBM_functions.cpp:1099-1171) of one colour's populations: density and velocity, summed in
// index order (rec_velocity's per-colour moments, :1386-1404)
void host_moments(const double* f, double v[3], double& rho)
{
    double mx = 0.0, my = 0.0, mz = 0.0;
    rho = 0.0;
    for (int i = 0; i < Q; ++i) {
        rho = rho + f[i];
        mx = mx + CXh[i] * f[i];
        my = my + CYh[i] * f[i];
        mz = mz + CZh[i] * f[i];
    }
    if (rho != 0.0) {
        const double one_over_rho = 1.0 / rho;
        v[0] = mx * one_over_rho;
        v[1] = my * one_over_rho;
        v[2] = mz * one_over_rho;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
}

}  // namespace

int lbm_init_drop(lbm_ctx* c, int id, double R, double width, int x0, int y0, double v0)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_init_drop");
    if (!std::isfinite(R) || !std::isfinite(v0) || !std::isfinite(width)) return fail(c, LBM_E_ARG, "non-finite argument");
    CaseState& k = c->cases[id];
    const int nx = k.d.nx, ny = k.d.ny;
    const double k_perfil = (width > 0.0) ? 2.0 * std::log(9.0) / width : -1.0;
    if (k.elem != 8) {
        // reduced-precision forms: computed on the device in the stored form (no host arrays, so
        // boxes of 10^8 sites initialise in milliseconds); FP64 keeps the host path below, whose
        // libm exp gives the reference's bits
        HIPCHK(c, launch_lp_init_drop(k.elem, k.sym, k.pop[k.cur], nx, k.cc.px, ny, R, k_perfil, x0, y0, v0, k.cc.ref_mid,
                                      k.cc.ref_half, c->stream));
        k.aux_valid = false;
        return LBM_OK;
    }
    const double rho_ini_R = 1.0, rho_ini_B = 1.0, delta_rho = 0.0;
    std::vector<double> fR((size_t)k.n * Q), fB((size_t)k.n * Q);
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) {
            const double dx = (double)(x - x0), dy = (double)(y - y0), dz = 0.0;
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            double phi;
            if (k_perfil > 0.0) {
                double arg = k_perfil * (r - R);
                if (arg > 60.0) arg = 60.0;
                if (arg < -60.0) arg = -60.0;
                phi = 1.0 / (1.0 + std::exp(arg));
            } else {
                phi = (r < R) ? 1.0 : 0.0;
            }
            const double rho_tot = rho_ini_B + phi * (rho_ini_R - rho_ini_B) + phi * delta_rho;
            const size_t p = (size_t)x + (size_t)nx * y;
            host_feq(&fR[p * Q], v0, 0., 0., phi * rho_tot);
            host_feq(&fB[p * Q], v0, 0., 0., (1.0 - phi) * rho_tot);
        }
    return lbm_upload_state(c, id, fR.data(), fB.data());
}

int lbm_init_band(lbm_ctx* c, int id, int p, int q, double width)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_init_band");
    if (p == 0 && q == 0) return fail(c, LBM_E_ARG, "p = q = 0");
    const CaseState& k = c->cases[id];
    const int nx = k.d.nx, ny = k.d.ny;
    const long long M = gcdll((long long)p * nx, (long long)q * ny) / gcdll(p, q);
    const double m = std::sqrt((double)(p * p + q * q));
    const double kk = (width > 0) ? 2.0 * std::log(9.0) / width : -1.0;
    std::vector<double> fR((size_t)k.n * Q), fB((size_t)k.n * Q);
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            long long s = ((long long)p * x + (long long)q * y) % M;
            if (s < 0) s += M;
            const double a = (double)s, b = 0.5 * (double)M - (double)s;
            const double c1 = (double)s - 0.5 * (double)M, c2 = (double)(M - s);
            const double dred = (2 * s < M) ? std::min(a, b) / m : -std::min(c1, c2) / m;
            double cR;
            if (kk > 0) cR = 1.0 / (1.0 + std::exp(-kk * dred));
            else cR = (dred > 0) ? 1.0 : (dred < 0 ? 0.0 : 0.5);
            const size_t pto = (size_t)x + (size_t)nx * y;
            for (int i = 0; i < Q; i++) {
                const double w = (i == 0) ? W0 : (i < 7 ? W1 : W2);
                fR[pto * Q + i] = w * cR;
                fB[pto * Q + i] = w * (1.0 - cR);
            }
        }
    return lbm_upload_state(c, id, fR.data(), fB.data());
}

int lbm_download_state(lbm_ctx* c, int id, int which, double* fR, double* fB)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_download_state");
    if (!fR || !fB) return fail(c, LBM_E_ARG, "null population array");
    if (which != LBM_STATE_POST_STREAM && which != LBM_STATE_POST_COLLISION)
        return fail(c, LBM_E_ARG, "bad state selector");
    CaseState& k = c->cases[id];
    const int nx = k.d.nx, ny = k.d.ny;
    std::vector<double> h(2 * Q * (size_t)k.S);
    const double* src64 = nullptr;
    if (int r = state_f64(c, k, &src64)) return r;
    HIPCHK(c, hipMemcpyAsync(h.data(), src64, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // post-stream: f_i(x) = f*_i(x - c_i), or f*_opp(i)(x) where x - c_i is solid (bounce-back); a
    // solid site's populations are 0
    const unsigned char* solid = k.solid_h.empty() ? nullptr : k.solid_h.data();
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const long long site = (long long)x + (long long)nx * y;
            if (solid && solid[site]) {
                for (int i = 0; i < Q; ++i) fR[(size_t)site * Q + i] = fB[(size_t)site * Q + i] = 0.0;
                continue;
            }
            for (int i = 0; i < Q; ++i) {
                long long src = (which == LBM_STATE_POST_STREAM) ? nbr(x, y, -CXh[i], -CYh[i], nx, ny) : site;
                int j = i;
                if (solid && which == LBM_STATE_POST_STREAM && solid[src]) {
                    src = site;
                    j = OPPh[i];
                }
                fR[(size_t)site * Q + i] = h[(size_t)j * k.S + src];
                fB[(size_t)site * Q + i] = h[(size_t)(Q + j) * k.S + src];
            }
        }
    return LBM_OK;
}

}  // extern "C"

namespace {

void drop_graph(lbm_ctx* c);

// Converts an FP64 P1/P3 or F3 case to the compact form (to = true) or the full one on the device: the
// state moves to the case's other buffer, every population unchanged.  Launch tables see the buffer
// change and are written again; a captured graph of the old form is dropped.
int set_form(lbm_ctx* c, CaseState& k, bool to)
{
    if (k.elem != 8 || k.compact == to || (to && k.kind != KIND_P && k.kind != KIND_F3)) return LBM_OK;
    if (capturing(c)) return refuse_capture(c, "changing a case's FP64 storage form (a step of another launch size)");
    HIPCHK(c, launch_f64_convert(k.pop[k.cur], k.pop[k.cur ^ 1], to, k.cc.beta, k.d.nx, k.cc.px, k.n, c->stream));
    k.cur ^= 1;
    k.compact = to;
    drop_graph(c);
    return LBM_OK;
}

// The form a launch of the P1/P3 or F3 cases in `g` uses: compact when it covers COMPACT_MIN_SITES
// sites or more (or as ITOLBM_F64_COMPACT fixes); its members are converted to it.
int group_form(lbm_ctx* c, const Group& g)
{
    if (g.kind != KIND_P && g.kind != KIND_F3) return LBM_OK;
    long long sites = 0;
    for (int id : g.members) sites += c->cases[id].n;
    const bool to = want_compact(c, sites);
    for (int id : g.members)
        if (int r = set_form(c, c->cases[id], to)) return r;
    return LBM_OK;
}

// (Re)writes a group's case table when a member's state buffer or aux pointer moved since.
int group_sync(lbm_ctx* c, Group& g)
{
    bool stale = !g.written;
    for (size_t j = 0; !stale && j < g.members.size(); ++j) {
        const CaseState& k = c->cases[g.members[j]];
        stale = (k.cur != (g.cur0[j] ^ (g.k & 1))) || (k.aux != g.aux0[j]);
    }
    if (!stale) return LBM_OK;
    if (capturing(c)) return refuse_capture(c, "writing a family's case table (first step after a change of cases, shape or state)");
    const int nm = (int)g.members.size();
    if (g.kind == KIND_GEO) {
        // no table: each step's launches take the members' buffers from the host bookkeeping; the
        // version still changes, so a captured graph of the old buffers is not replayed
        g.cur0.assign(nm, 0);
        g.aux0.assign(nm, nullptr);
        for (int j = 0; j < nm; ++j) {
            g.cur0[j] = c->cases[g.members[j]].cur;
            g.aux0[j] = c->cases[g.members[j]].aux;
        }
        g.ntiles = 0;
        g.k = 0;
        g.written = true;
        g.version++;
        return LBM_OK;
    }
    if (g.ly.empty()) {
        const int ly = group_rows(c->cases, g.members);
        g.ly.assign(nm, ly);
    }
    std::vector<CaseDev> tab(nm);
    std::vector<int> tile_case;
    g.cur0.assign(nm, 0);
    g.aux0.assign(nm, nullptr);
    int tile0 = 0;
    for (int j = 0; j < nm; ++j) {
        const CaseState& k = c->cases[g.members[j]];
        CaseDev& d = tab[j];
        d.cc = k.cc;
        d.pop[0] = k.pop[k.cur];
        d.pop[1] = k.pop[k.cur ^ 1];
        d.aux = k.aux;
        d.S = k.S;
        d.ly = std::max(1, std::min(g.ly[j], k.d.ny));
        d.nstrip = strips_of(k.d);
        d.nseg = (k.d.ny + d.ly - 1) / d.ly;
        d.tile0 = tile0;
        d.kind = k.kind;
        const int nt = d.nstrip * d.nseg;
        tile_case.insert(tile_case.end(), nt, j);
        tile0 += nt;
        g.cur0[j] = k.cur;
        g.aux0[j] = k.aux;
    }
    if (g.dtab && nm > g.cap) {
        (void)hipFree(g.dtab);
        g.dtab = nullptr;
        c->bytes -= (long long)sizeof(CaseDev) * g.cap;
    }
    if (!g.dtab) {
        if (int r = dev_alloc(c, (void**)&g.dtab, sizeof(CaseDev) * nm)) return r;
        g.cap = nm;
    }
    if (g.dtile && tile0 > g.tcap) {
        (void)hipFree(g.dtile);
        g.dtile = nullptr;
        c->bytes -= (long long)sizeof(int) * g.tcap;
    }
    if (!g.dtile) {
        if (int r = dev_alloc(c, (void**)&g.dtile, sizeof(int) * std::max(tile0, 1))) return r;
        g.tcap = std::max(tile0, 1);
    }
    HIPCHK(c, hipMemcpyAsync(g.dtab, tab.data(), sizeof(CaseDev) * nm, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(g.dtile, tile_case.data(), sizeof(int) * tile0, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    g.ntiles = tile0;
    g.k = 0;
    g.written = true;
    g.version++;
    return LBM_OK;
}

int launch_group(lbm_ctx* c, Group& g, int parity, int record, hipStream_t st)
{
    if (g.kind == KIND_GEO) {   // cases with solid sites: the site-parallel kernels, case by case
        for (int id : g.members) {
            const CaseState& k = c->cases[id];
            GeoLaunch a{};
            a.cc = k.cc;
            a.in = k.pop[k.cur];
            a.out = k.pop[k.cur ^ 1];
            a.solid = k.solid;
            a.wett = k.wett;
            a.phi = k.geo_f;
            a.g = k.geo_f + k.S;
            a.nt = k.geo_f + 4 * k.S;
            a.aux = record ? k.aux : nullptr;
            a.S = k.S;
            HIPCHK(c, launch_step_geo(step_kind(k.d.op), a, st));
        }
        return LBM_OK;
    }
    StepLaunch a{};
    a.tab = g.dtab;
    a.tile_case = g.dtile;
    a.parity = parity;
    a.record = record;
    a.depth = g.depth;
    const CaseState& k0 = c->cases[g.members[0]];   // a context has one storage form
    a.ntiles = g.ntiles;
    a.compact = 0;   // the form of the launch's P1/P3 and F3 cases (all alike)
    for (int id : g.members)
        if (c->cases[id].kind == KIND_P || c->cases[id].kind == KIND_F3) {
            a.compact = c->cases[id].compact ? 1 : 0;
            break;
        }
    if (g.kind == KIND_FUSED) {   // every strip family of the FP64 context in one launch
        const int cap = step_f64_fused_capacity(a.compact != 0);
        const int grid = (c->fuse_mode == 2 && cap > 0) ? std::min(cap, (g.ntiles + 7) / 8 * 8) : g.ntiles;
        HIPCHK(c, launch_step_f64_fused(a, grid, st));
        return LBM_OK;
    }
    HIPCHK(c, k0.elem == 8 ? launch_step_f64(g.kind, a, g.ntiles, st)
                           : launch_step_lp(k0.elem, k0.sym, g.kind, a, g.ntiles, st));
    return LBM_OK;
}

// a tile shape and its measured time per step
struct Cand {
    int ly, depth;
    float ms;
};

// Tables, aux buffers and tile shape of a group, before its launches are queued.
int group_prepare(lbm_ctx* c, Group& g, bool rec)
{
    if (rec)
        for (int id : g.members)
            if (int r = ensure_aux(c, c->cases[id])) return r;
    return group_sync(c, g);
}

// Whether a group still needs its tile shape tuned: ITOLBM_LY (and ITOLBM_DEPTH, 0 = the kernel's
// default) fix the shape instead.
bool needs_tuning(lbm_ctx* c, Group& g)
{
    if (g.kind == KIND_GEO) g.tuned = true;   // site-parallel kernels: no tile shape
    if (g.tuned || !c->autotune) return false;
    if (std::getenv("ITOLBM_LY")) {
        if (const char* e = std::getenv("ITOLBM_DEPTH")) g.depth = std::max(0, std::min(3, std::atoi(e)));
        g.tuned = true;
        return false;
    }
    return true;
}

// Queues n steps of a prepared group on stream st (the last one recording aux when asked).
int group_step(lbm_ctx* c, Group& g, int n_steps, int flags, hipStream_t st)
{
    if (g.members.empty() || n_steps <= 0) return LBM_OK;
    const bool rec = (flags & LBM_STEP_RECORD_AUX) != 0;
    for (int s = 0; s < n_steps; ++s) {
        if (int r = launch_group(c, g, g.k & 1, (rec && s == n_steps - 1) ? 1 : 0, st)) return r;
        g.k++;
        for (int id : g.members) c->cases[id].cur ^= 1;
    }
    for (int id : g.members) c->cases[id].aux_valid = rec;
    return LBM_OK;
}

// Queues n steps of the active families (already prepared).  The families advance disjoint sets of
// cases, so with several of them each runs its n steps on a stream of its own, forked from and
// joined back to the context's stream: small launches of different families fill the GPU together.
int queue_steps(lbm_ctx* c, const std::vector<Group*>& act, int n_steps, int flags)
{
    if (act.size() == 1) return group_step(c, *act[0], n_steps, flags, c->stream);
    HIPCHK(c, hipEventRecord(c->fork, c->stream));
    for (size_t j = 0; j < act.size(); ++j) {
        hipStream_t st = c->side[j];
        HIPCHK(c, hipStreamWaitEvent(st, c->fork, 0));
        if (int r = group_step(c, *act[j], n_steps, flags, st)) return r;
        HIPCHK(c, hipEventRecord(c->join[j], st));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->join[j], 0));
    }
    return LBM_OK;
}

// Tile shapes of lbm_step's families, tuned on real steps.  Every shape computes every site the same
// way, so candidate shapes can be timed on steps the caller asked for: the simulation advances
// exactly as it would otherwise.  Real steps alternate the state buffers as stepping does (launches
// repeated on one buffer pair re-read an input that the last-level cache then partly holds, which
// favoured the wrong shapes), and all active families run together, as they do in a step, so a
// family's shape is chosen in the mix.  Coordinate search, one family at a time with the others
// fixed: the pipeline depth (FP32/FP16 kernels; FP64 F1 / F2: rows read again, or their totals and
// colour sums kept) at the current rows per segment, then the rows per segment (1 .. 64) at that depth,
// then the other depths at the best rows per segment, each candidate over W untimed + T timed steps; the
// two fastest are timed again over 2T steps.  Only families without a tuned shape are searched (a
// family whose members changed).  The steps used are subtracted from `left`; nothing is done unless
// `left` covers the whole search.
int online_tune(lbm_ctx* c, const std::vector<Group*>& act, int& left)
{
    constexpr int W = 1, T = 6;
    constexpr int LYS[] = {1, 2, 4, 8, 16, 32, 64};   // short segments suit small cases (more wavefronts)
    // FP32/FP16: rows in flight ahead 1 .. 3; FP64 F1 / F2: 1 reads a row's populations again for its
    // collision, 2 keeps them in registers (one wavefront per SIMD); other FP64 kernels: one form
    auto depths_of = [&](const Group& g) {
        const CaseState& k0 = c->cases[g.members[0]];
        if (k0.elem != 8) return 3;
        return (g.kind == KIND_F1 || g.kind == KIND_F2) && !k0.compact ? 2 : 1;
    };
    constexpr int COMMON[] = {4, 8, 16, 32, 64};
    long long budget = 0;
    int untuned = 0;
    for (const Group* g : act)
        if (!g->tuned) {
            budget += (long long)(2 * depths_of(*g) + 5) * (W + T) + 2 * (W + 2 * T);
            ++untuned;
        }
    if (untuned > 1) budget += (long long)(sizeof(COMMON) / sizeof(COMMON[0])) * (W + T);
    if (budget > left) return LBM_OK;
    if (!c->ev[0]) {
        HIPCHK(c, hipEventCreate(&c->ev[0]));
        HIPCHK(c, hipEventCreate(&c->ev[1]));
    }
    auto run = [&](int steps, float* ms) -> int {
        if (ms) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        if (int r = queue_steps(c, act, steps, 0)) return r;
        left -= steps;
        if (ms) {
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev[1]));
            HIPCHK(c, hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
            *ms /= (float)steps;
        }
        return LBM_OK;
    };
    // First one rows-per-segment for every untuned family together: the families share the GPU in a
    // step, and each family's search below starts from the others' shapes.  Starting from the
    // families' separate defaults, the coordinate search settled on mixes up to 3 % slower than the
    // best common shape (eight 512^2 cases of four families, FP64).
    if (untuned > 1) {
        float best = 0.f;
        int best_ly = 16;
        for (int v : COMMON) {
            for (Group* g : act)
                if (!g->tuned) {
                    g->ly.assign(g->members.size(), v);
                    g->depth = 0;
                    g->written = false;
                    if (int r = group_sync(c, *g)) return r;
                }
            float ms = 0.f;
            if (int r = run(W, nullptr)) return r;
            if (int r = run(T, &ms)) return r;
            if (v == COMMON[0] || ms < best) {
                best = ms;
                best_ly = v;
            }
        }
        for (Group* g : act)
            if (!g->tuned) {
                g->ly.assign(g->members.size(), best_ly);
                g->written = false;
                if (int r = group_sync(c, *g)) return r;
            }
    }
    for (Group* g : act) {
        if (g->tuned) continue;   // a family whose members did not change keeps its shape
        auto apply = [&](const Cand& cd) -> int {
            g->depth = cd.depth;
            g->ly.assign(g->members.size(), cd.ly);
            g->written = false;
            return group_sync(c, *g);
        };
        const char* log = std::getenv("ITOLBM_TUNE_LOG");
        auto time_of = [&](Cand& cd, int t) -> int {
            if (int r = apply(cd)) return r;
            if (int r = run(W, nullptr)) return r;
            const int r = run(t, &cd.ms);
            if (log && std::atoi(log) >= 2)
                std::fprintf(stderr, "itolbm step tuner: family %d trial %d rows, depth %d: %.1f us over %d steps\n",
                             g->kind, cd.ly, cd.depth, 1000.0 * cd.ms, t);
            return r;
        };
        auto by_time = [](const Cand& x, const Cand& y) { return x.ms < y.ms; };
        // starting rows per segment: the current one, rounded to a candidate
        int ly0 = g->ly.empty() ? 16 : g->ly[0];
        int start = LYS[0];
        for (int v : LYS)
            if (v <= ly0) start = v;
        std::vector<Cand> tried;
        const bool lp = depths_of(*g) > 1;
        for (int d = 1; d <= depths_of(*g); ++d) {
            Cand cd{start, lp ? d : 0, 0.f};
            if (int r = time_of(cd, T)) return r;
            tried.push_back(cd);
        }
        const int best_depth = std::min_element(tried.begin(), tried.end(), by_time)->depth;
        for (int v : LYS) {
            if (v == start) continue;
            Cand cd{v, best_depth, 0.f};
            if (int r = time_of(cd, T)) return r;
            tried.push_back(cd);
        }
        // the other depths again at the best segment length (the best depth can change with it)
        const int best_ly = std::min_element(tried.begin(), tried.end(), by_time)->ly;
        if (lp && best_ly != start)
            for (int d = 1; d <= depths_of(*g); ++d) {
                if (d == best_depth) continue;
                Cand cd{best_ly, d, 0.f};
                if (int r = time_of(cd, T)) return r;
                tried.push_back(cd);
            }
        std::sort(tried.begin(), tried.end(), by_time);
        for (int j = 0; j < 2; ++j)
            if (int r = time_of(tried[j], 2 * T)) return r;
        std::sort(tried.begin(), tried.begin() + 2, by_time);
        // the starting segment length stays unless another is faster by more than 1 % (run-to-run noise)
        if (tried[0].ly != start && tried[1].ly == start && tried[1].ms <= 1.01f * tried[0].ms) std::swap(tried[0], tried[1]);
        if (int r = apply(tried[0])) return r;
        g->tuned = true;
        if (log)
            std::fprintf(stderr, "itolbm step tuner: family %d, %zu cases, element %d B: %d rows per segment, depth %d, step %.1f us\n",
                         g->kind, g->members.size(), c->cases[g->members[0]].elem, tried[0].ly, tried[0].depth,
                         1000.0 * tried[0].ms);
    }
    return LBM_OK;
}

// The fused group's members: every active case of the strip families, in case order, each with its
// family's rows per segment.  Rewritten (table stale) only when that set or a shape changed.
void fused_members(lbm_ctx* c)
{
    std::vector<int> mem, ly;
    for (int id = 0; id < (int)c->cases.size(); ++id) {
        const CaseState& k = c->cases[id];
        if (!k.active || k.kind >= KIND_GEO) continue;
        const Group& f = c->family[k.kind];
        const auto it = std::find(f.members.begin(), f.members.end(), id);
        const size_t j = (size_t)(it - f.members.begin());
        mem.push_back(id);
        ly.push_back(j < f.ly.size() ? f.ly[j] : group_rows(c->cases, f.members));
    }
    Group& g = c->fused;
    if (mem != g.members || ly != g.ly) {
        g.members = mem;
        g.ly = ly;
        g.written = false;
    }
}

// The groups lbm_step launches: the families, or the one fused launch (plus cases with solids).
std::vector<Group*> run_groups(lbm_ctx* c, const std::vector<Group*>& act)
{
    if (c->fuse_mode <= 0 || c->fused.members.empty()) return act;
    std::vector<Group*> run{&c->fused};
    for (Group* g : act)
        if (g->kind == KIND_GEO) run.push_back(g);
    return run;
}

// Chooses lbm_step's launch form for an FP64 context once its families are tuned: one launch per
// family, one launch of all (step_fused, a wavefront per tile) or that launch persistent (a grid of
// the resident capacity), each timed over W untimed + 2T timed real steps; one launch per family
// stays unless another form is faster by more than 1 %.  ITOLBM_FUSED=0|1|2 fixes the form.  The
// forms give the same bits (the same per-tile code).  The steps used are subtracted from `left`.
int fuse_tune(lbm_ctx* c, const std::vector<Group*>& act, int& left)
{
    int strip = 0;
    for (const Group* g : act) strip += g->kind < KIND_GEO;
    // one launch runs one FP64 form for its P1/P3 and F3 cases: when their families' launch sizes put
    // them in different forms, each family keeps its own launch
    int forms = 0;
    for (const CaseState& k : c->cases)
        if (k.active && (k.kind == KIND_P || k.kind == KIND_F3)) forms |= k.compact ? 2 : 1;
    if (c->cases.empty() || c->cases[0].elem != 8 || strip == 0 || forms == 3) {
        c->fuse_mode = 0;
        return LBM_OK;
    }
    if (const char* e = std::getenv("ITOLBM_FUSED")) {
        c->fuse_mode = std::max(0, std::min(2, std::atoi(e)));
        return LBM_OK;
    }
    constexpr int W = 1, T = 12;
    if (!c->autotune || left < 3 * (W + T)) return LBM_OK;   // undecided: one launch per family for now
    fused_members(c);
    float ms[3] = {0.f, 0.f, 0.f};
    for (int mode = 0; mode < 3; ++mode) {
        c->fuse_mode = mode;
        const std::vector<Group*> run = run_groups(c, act);
        for (Group* g : run)   // the other form's steps moved the buffers: tables are rewritten
            if (int r = group_prepare(c, *g, false)) return r;
        if (int r = queue_steps(c, run, W, 0)) return r;
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        if (int r = queue_steps(c, run, T, 0)) return r;
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev[1]));
        HIPCHK(c, hipEventElapsedTime(&ms[mode], c->ev[0], c->ev[1]));
        left -= W + T;
    }
    int best = 0;
    for (int mode = 1; mode < 3; ++mode)
        if (ms[mode] < ms[best] && ms[mode] < 0.99f * ms[0]) best = mode;
    c->fuse_mode = best;
    if (std::getenv("ITOLBM_TUNE_LOG"))
        std::fprintf(stderr, "itolbm step tuner: launch form %d (per family %.1f us, one launch %.1f us, persistent %.1f us)\n",
                     best, 1000.0 * ms[0] / T, 1000.0 * ms[1] / T, 1000.0 * ms[2] / T);
    return LBM_OK;
}

// Host bookkeeping of n steps run by a graph replay (the launches themselves were captured).
void advance(lbm_ctx* c, Group& g, int n_steps)
{
    g.k += n_steps;
    for (int id : g.members) {
        c->cases[id].cur ^= (n_steps & 1);
        c->cases[id].aux_valid = false;
    }
}

void drop_graph(lbm_ctx* c)
{
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    if (c->graph.graph) (void)hipGraphDestroy(c->graph.graph);
    c->graph = StepGraph{};
}

std::vector<long long> graph_key(const lbm_ctx* c, const std::vector<Group*>& act)
{
    std::vector<long long> key{c->fuse_mode};
    for (const Group* g : act) {
        key.push_back(g->kind);
        key.push_back(g->version);
        key.push_back(g->ntiles);
        key.push_back(g->k & 1);
    }
    return key;
}

// Captures graph_steps steps of the active families from the context's stream (the side streams
// join the capture through the fork / join events), unless the current graph already covers them.
// Capturing queues nothing, so the host bookkeeping of the captured steps is undone afterwards.
int ensure_graph(lbm_ctx* c, const std::vector<Group*>& act, int steps)
{
    std::vector<long long> key = graph_key(c, act);
    if (c->graph.exec && c->graph.steps == steps && c->graph.key == key) return LBM_OK;
    drop_graph(c);
    std::vector<int> k0, cur0;
    for (const Group* g : act) k0.push_back(g->k);
    for (const CaseState& k : c->cases) cur0.push_back(k.cur);
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    const int r = queue_steps(c, act, steps, 0);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    for (size_t j = 0; j < act.size(); ++j) act[j]->k = k0[j];
    for (size_t i = 0; i < c->cases.size(); ++i) c->cases[i].cur = cur0[i];
    if (r != LBM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
    }
    if (e != hipSuccess) return fail(c, LBM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    c->graph.graph = graph;
    HIPCHK(c, hipGraphInstantiate(&c->graph.exec, graph, nullptr, nullptr, 0));
    c->graph.steps = steps;
    c->graph.key = std::move(key);
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_step_case(lbm_ctx* c, int id, int n_steps, int flags)
{
    if (int r = check_case(c, id)) return r;
    if (n_steps < 0) return fail(c, LBM_E_ARG, "negative step count");
    Group& g = c->single[id];
    if (n_steps == 0 || g.members.empty()) return LBM_OK;
    const bool cap = capturing(c);
    if (cap && (n_steps & 1)) return fail(c, LBM_E_ARG, "a captured step call must advance an even number of steps");
    const bool rec = (flags & LBM_STEP_RECORD_AUX) != 0;
    if (int r = group_form(c, g)) return r;
    if (int r = group_prepare(c, g, rec)) return r;
    int left = n_steps;
    if (!cap && needs_tuning(c, g)) {   // as lbm_step: the tile shape is tuned on the call's own steps
        int avail = left - (rec ? 1 : 0);
        const int before = avail;
        std::vector<Group*> one{&g};
        if (int r = online_tune(c, one, avail)) return r;
        left -= before - avail;
    }
    return group_step(c, g, left, flags, c->stream);
}

// lbm_step_case without allocation: the case's table entry and tile map go into the device's static
// scratch array (static_step.hip), written by a kernel whose arguments carry the entry, instead of the
// table lbm_step_case allocates.  Same kernels, same tile shape (the one lbm_step_case tuned for the
// case, else the default), same results.  Nothing is allocated, freed, copied from host memory or
// synchronised; no tuning.  Calls that overlap on one device share the scratch array and race.
int lbm_step_case_noalloc(lbm_ctx* c, int id, int n_steps, int flags)
{
    if (int r = check_case(c, id)) return r;
    if (n_steps < 0) return fail(c, LBM_E_ARG, "negative step count");
    if (n_steps == 0) return LBM_OK;
    if ((n_steps & 1) && capturing(c))
        return fail(c, LBM_E_ARG, "a captured step call must advance an even number of steps");
    CaseState& k = c->cases[id];
    if (k.solid) return fail(c, LBM_E_ARG, "a case with solid sites steps with lbm_step / lbm_step_case only");
    const bool rec = (flags & LBM_STEP_RECORD_AUX) != 0;
    if (rec && !k.aux)
        return fail(c, LBM_E_STATE, "LBM_STEP_RECORD_AUX needs the aux fields, which a first lbm_step or lbm_step_case "
                                    "with LBM_STEP_RECORD_AUX allocates");
    if (!c->scratch) return fail(c, LBM_E_STATE, "no static scratch on this device");
    const Group& g = c->single[id];
    if (int r = group_form(c, g)) return r;   // a conversion kernel when the form changes: nothing allocated
    const int ly = g.ly.empty() ? group_rows(c->cases, g.members) : g.ly[0];
    CaseDev d{};
    d.cc = k.cc;
    d.pop[0] = k.pop[k.cur];
    d.pop[1] = k.pop[k.cur ^ 1];
    d.aux = k.aux;
    d.S = k.S;
    d.ly = std::max(1, std::min(ly, k.d.ny));
    d.nstrip = strips_of(k.d);
    d.nseg = (k.d.ny + d.ly - 1) / d.ly;
    d.tile0 = 0;
    const long long ntiles = (long long)d.nstrip * d.nseg;
    if (ntiles > SCRATCH_MAX_TILES)
        return fail(c, LBM_E_ARG, "case needs " + std::to_string(ntiles) + " tiles; the static scratch holds " +
                                      std::to_string(SCRATCH_MAX_TILES) + " (use lbm_step_case)");
    HIPCHK(c, launch_fill_static_table(c->scratch, d, (int)ntiles, c->stream));
    StepLaunch a{};
    a.tab = scratch_table(c->scratch);
    a.tile_case = scratch_tiles(c->scratch);
    a.depth = g.depth;
    a.compact = k.compact ? 1 : 0;
    for (int s = 0; s < n_steps; ++s) {
        a.parity = s & 1;
        a.record = (rec && s == n_steps - 1) ? 1 : 0;
        HIPCHK(c, k.elem == 8 ? launch_step_f64(k.kind, a, (int)ntiles, c->stream)
                              : launch_step_lp(k.elem, k.sym, k.kind, a, (int)ntiles, c->stream));
        k.cur ^= 1;
    }
    k.aux_valid = rec;
    return LBM_OK;
}

int lbm_step(lbm_ctx* c, int n_steps, int flags)
{
    if (!c) return LBM_E_ARG;
    if (n_steps < 0) return fail(c, LBM_E_ARG, "negative step count");
    if (n_steps == 0) return LBM_OK;
    std::vector<Group*> act;
    for (Group& g : c->family)
        if (!g.members.empty()) act.push_back(&g);
    if (act.empty()) return LBM_OK;
    // Under a caller's graph capture only kernel launches (and the fork / join of the families'
    // streams) are queued: no tuning, no graph of the library's own, no table written.  An even step
    // count returns every family to the buffers the host bookkeeping expects after each replay.
    const bool cap = capturing(c);
    if (cap && (n_steps & 1)) return fail(c, LBM_E_ARG, "a captured step call must advance an even number of steps");
    const bool rec = (flags & LBM_STEP_RECORD_AUX) != 0;
    bool untuned = false;
    for (Group* g : act)
        if (int r = group_form(c, *g)) return r;
    for (Group* g : act) {
        if (int r = group_prepare(c, *g, rec)) return r;
        if (!cap && needs_tuning(c, *g)) untuned = true;
    }
    if (act.size() > 1 && !c->fork) {
        HIPCHK(c, hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
        for (int k = 0; k < N_KINDS; ++k) {
            HIPCHK(c, hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking));
            HIPCHK(c, hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming));
        }
    }
    // Whole graphs of graph_steps steps first (the last step stays outside when it records aux),
    // then the remaining steps launch by launch.  Families without a tuned shape are tuned first on
    // the call's own steps when there are enough of them (the recording step excepted).
    int left = n_steps;
    if (untuned) {
        c->fuse_mode = -1;   // the families' shapes change: the launch form is chosen again
        int avail = left - (rec ? 1 : 0);
        const int before = avail;
        if (int r = online_tune(c, act, avail)) return r;
        left -= before - avail;
    }
    bool tuned = true;
    for (const Group* g : act) tuned = tuned && (g->tuned || !c->autotune || g->kind == KIND_GEO);
    if (!cap && c->fuse_mode < 0 && (tuned || std::getenv("ITOLBM_FUSED"))) {
        if (!c->ev[0]) {
            HIPCHK(c, hipEventCreate(&c->ev[0]));
            HIPCHK(c, hipEventCreate(&c->ev[1]));
        }
        int avail = left - (rec ? 1 : 0);
        const int before = avail;
        if (int r = fuse_tune(c, act, avail)) return r;
        left -= before - avail;
    }
    if (c->fuse_mode > 0 && !cap) fused_members(c);
    const std::vector<Group*> run = run_groups(c, act);
    for (Group* g : run)   // tables written before tuning steps of another launch form are stale
        if (int r = group_prepare(c, *g, rec)) return r;
    const int G = cap ? 0 : c->graph_steps;
    if (G > 0 && left - (rec ? 1 : 0) >= G) {
        const int reps = (left - (rec ? 1 : 0)) / G;
        if (int r = ensure_graph(c, run, G)) return r;
        for (int i = 0; i < reps; ++i) HIPCHK(c, hipGraphLaunch(c->graph.exec, c->stream));
        for (Group* g : run) advance(c, *g, reps * G);
        left -= reps * G;
    }
    if (left > 0) return queue_steps(c, run, left, flags);
    return LBM_OK;
}

int lbm_set_graph(lbm_ctx* c, int steps_per_graph)
{
    if (!c) return LBM_E_ARG;
    if (steps_per_graph < 0 || (steps_per_graph & 1)) return fail(c, LBM_E_ARG, "steps per graph must be even and >= 0");
    c->graph_steps = steps_per_graph;
    drop_graph(c);
    return LBM_OK;
}

int lbm_sync(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (capturing(c)) return refuse_capture(c, "lbm_sync");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_set_stream(lbm_ctx* c, void* s)
{
    if (!c) return LBM_E_ARG;
    if (capturing(c)) return refuse_capture(c, "lbm_set_stream");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return LBM_OK;
}

int lbm_stream_copy(void* dst, const void* src, unsigned long long bytes, void* s)
{
    if (!dst || !src) return LBM_E_ARG;
    const hipError_t e = launch_stream_copy(dst, src, (size_t)bytes, (hipStream_t)s);
    if (e == hipErrorInvalidValue) return LBM_E_ARG;
    return e == hipSuccess ? LBM_OK : LBM_E_HIP;
}

int lbm_tile_shape(const lbm_ctx* c, int id, int* rows, int* depth, int* tuned)
{
    if (!c || id < 0 || id >= (int)c->cases.size()) return LBM_E_ARG;
    const Group& g = c->family[c->cases[id].kind];
    if (rows) *rows = g.ly.empty() ? group_rows(c->cases, g.members) : g.ly[0];
    if (depth) *depth = g.depth;
    if (tuned) *tuned = g.tuned ? 1 : 0;
    return LBM_OK;
}

int lbm_step_mode(const lbm_ctx* c, int* mode)
{
    if (!c || !mode) return LBM_E_ARG;
    *mode = c->fuse_mode;
    return LBM_OK;
}

int lbm_stored_bytes(const lbm_ctx* c, int id, int* bytes_per_update)
{
    if (!c || id < 0 || id >= (int)c->cases.size() || !bytes_per_update) return LBM_E_ARG;
    const CaseState& k = c->cases[id];
    const int per_site = k.elem == 8 ? (k.compact ? Q + 5 : 2 * Q) * 8 : 2 * k.nslot * k.elem;
    *bytes_per_update = 2 * per_site;   // read once, written once
    return LBM_OK;
}

int lbm_set_active(lbm_ctx* c, int id, int active)
{
    if (int r = check_case(c, id)) return r;
    CaseState& k = c->cases[id];
    if (k.active == (active != 0)) return LBM_OK;
    k.active = active != 0;
    Group& g = c->family[k.kind];
    g.members.clear();
    for (int j = 0; j < (int)c->cases.size(); ++j)
        if (c->cases[j].active && c->cases[j].kind == k.kind) g.members.push_back(j);
    g.ly.clear();   // row segments are re-chosen for the new member set
    g.written = false;
    g.tuned = false;
    c->fuse_mode = -1;
    return LBM_OK;
}

int lbm_set_geometry(lbm_ctx* c, int id, const int* meio, double wett_R)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_set_geometry");
    CaseState& k = c->cases[id];
    if (k.elem != 8) return fail(c, LBM_E_ARG, "solid sites need the FP64 form (LBM_PREC_F64)");
    if (!std::isfinite(wett_R)) return fail(c, LBM_E_ARG, "non-finite wettability");
    std::vector<unsigned char> m;
    long long nsolid = 0;
    if (meio) {
        m.assign((size_t)k.n, 0);
        for (long long p = 0; p < k.n; ++p) {
            if (meio[p] < 0) return fail(c, LBM_E_ARG, "negative entry in the geometry (0 = solid, > 0 = fluid)");
            m[(size_t)p] = meio[p] == 0;
            nsolid += m[(size_t)p];
        }
        if (nsolid == k.n) return fail(c, LBM_E_ARG, "the geometry has no fluid site");
    }
    // the post-stream state is kept across the change: downloaded with the old geometry, uploaded
    // with the new one (at a site that turns solid it is dropped; one that turns fluid starts at 0)
    std::vector<double> fR((size_t)k.n * Q), fB((size_t)k.n * Q);
    if (int r = lbm_download_state(c, id, LBM_STATE_POST_STREAM, fR.data(), fB.data())) return r;
    const int kind = nsolid ? KIND_GEO : step_kind(k.d.op);
    if (nsolid) {
        if (!k.solid && dev_alloc(c, (void**)&k.solid, (size_t)k.n)) return LBM_E_NOMEM;
        if (!k.geo_f && dev_alloc(c, (void**)&k.geo_f, sizeof(double) * 10 * (size_t)k.S)) return LBM_E_NOMEM;
        HIPCHK(c, hipMemcpyAsync(k.solid, m.data(), (size_t)k.n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        k.solid_h = std::move(m);
    } else {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (k.solid) {
            (void)hipFree(k.solid);
            c->bytes -= k.n;
            k.solid = nullptr;
        }
        if (k.geo_f) {
            (void)hipFree(k.geo_f);
            c->bytes -= (long long)sizeof(double) * 10 * k.S;
            k.geo_f = nullptr;
        }
        k.solid_h.clear();
    }
    k.wett = wett_R;
    if (kind != k.kind) {
        // the case changes kernel family: both families' members and tables are redone
        const int old = k.kind;
        k.kind = kind;
        k.compact = false;   // uploaded below in the full form; a later step chooses the form
        for (int kd : {old, kind}) {
            Group& g = c->family[kd];
            g.members.clear();
            for (int j = 0; j < (int)c->cases.size(); ++j)
                if (c->cases[j].active && c->cases[j].kind == kd) g.members.push_back(j);
            g.ly.clear();
            g.written = false;
            g.tuned = false;
        }
        Group& one = c->single[id];
        one.kind = kind;
        one.ly.clear();
        one.written = false;
        one.tuned = false;
    }
    drop_graph(c);   // a captured graph holds the old wettability and family launches
    c->fuse_mode = -1;
    return lbm_upload_state(c, id, fR.data(), fB.data());
}

long long lbm_read_geometry(const char* path, int pts_in, int pts_out, int* nx, int* ny, int* nz, int* meio,
                            long long cap)
{
    if (!path || pts_in < 0 || pts_out < 0) return LBM_E_ARG;
    std::ifstream in(path);
    if (!in) return LBM_E_ARG;
    std::string line;
    for (int i = 0; i < 4; ++i)
        if (!std::getline(in, line)) return LBM_E_ARG;
    if (!std::getline(in, line)) return LBM_E_ARG;
    std::istringstream dims(line);
    std::string word;
    long long X = 0, Y = 0, Z = 0;
    if (!(dims >> word >> X >> Y >> Z) || X < 1 || Y < 1 || Z < 1) return LBM_E_ARG;
    for (int i = 0; i < 5; ++i)
        if (!std::getline(in, line)) return LBM_E_ARG;
    const long long NX = X + pts_in + pts_out;
    if (NX * Y * Z > (1LL << 31)) return LBM_E_ARG;
    if (nx) *nx = (int)NX;
    if (ny) *ny = (int)Y;
    if (nz) *nz = (int)Z;
    const long long total = NX * Y * Z;
    if (!meio) {   // size query: the number of fluid points still needs the whole file
        std::vector<int> tmp((size_t)total);
        return lbm_read_geometry(path, pts_in, pts_out, nullptr, nullptr, nullptr, tmp.data(), total);
    }
    if (cap < total) return LBM_E_ARG;
    long long fluid = 0;
    for (long long z = 0; z < Z; ++z)
        for (long long y = 0; y < Y; ++y)
            for (long long x = 0; x < NX; ++x) {
                const long long pos = x + y * NX + z * NX * Y;
                int v = 0;
                if (x >= pts_in && x < NX - pts_out) {
                    double d = 0.0;   // the reference reads an int; files written as floats ("1.0") read the same
                    if (!(in >> d)) return LBM_E_ARG;
                    v = d != 0.0;
                } else {
                    v = (x != 0 && x != NX - 1);   // added layers: fluid except the outermost columns
                }
                meio[pos] = v ? (int)++fluid : 0;
            }
    return fluid;
}

int lbm_write_vtk(const char* path, int field, int nx, int ny, const double* fR, const double* fB, const double* force,
                  const int* meio, int flags)
{
    if (!path || nx < 1 || ny < 1 || !fR || (field == LBM_VTK_VELOCITY && !fB)) return LBM_E_ARG;
    if (field != LBM_VTK_DENSITY && field != LBM_VTK_VELOCITY) return LBM_E_ARG;
    if (flags & ~(LBM_VTK_BINARY | LBM_VTK_DOUBLE)) return LBM_E_ARG;
    const bool binary = (flags & LBM_VTK_BINARY) != 0, dbl = (flags & LBM_VTK_DOUBLE) != 0;
    std::ofstream out(path, std::ios::binary);
    if (!out) return LBM_E_ARG;
    const long long n = (long long)nx * ny;
    const bool vel = field == LBM_VTK_VELOCITY;
    out << "# vtk DataFile Version 2.0\n" << (vel ? "Velocidade" : "Densidade") << '\n'
        << (binary ? "BINARY" : "ASCII") << "\nDATASET STRUCTURED_POINTS\n";
    out << "DIMENSIONS " << nx << ' ' << ny << ' ' << 1 << "\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\n";
    out << "POINT_DATA " << n << '\n';
    const char* type = (binary && !dbl) ? "float" : "double";
    if (vel) out << "VECTORS velocidade " << type << '\n';
    else out << "SCALARS densidade " << type << "\nLOOKUP_TABLE default\n";
    // the value(s) of site s: density of fR (rec_density), or the colour-weighted velocity with Guo's
    // half-force correction (rec_velocity); 0 at a solid site
    auto site_values = [&](long long s, double v[3]) {
        if (meio && meio[s] == 0) {
            v[0] = v[1] = v[2] = 0.0;
            return;
        }
        if (!vel) {
            double rho = 0.0;
            for (int i = 0; i < Q; ++i) rho += fR[s * Q + i];
            v[0] = rho;
            return;
        }
        double vR[3], vB[3], rR, rB;
        host_moments(fR + s * Q, vR, rR);
        host_moments(fB + s * Q, vB, rB);
        const double cR = rR / (rR + rB), cB = 1. - cR;
        for (int a = 0; a < 3; ++a) v[a] = cR * vR[a] + cB * vB[a];
        if (force) {
            const double rho = rR + rB;
            if (rho > 0.0)
                for (int a = 0; a < 3; ++a) v[a] = v[a] - 0.5 * force[3 * s + a] / rho;
        }
    };
    const int nv = vel ? 3 : 1;
    if (!binary) {
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) {
                double v[3];
                site_values((long long)y * nx + x, v);
                for (int a = 0; a < nv; ++a) out << v[a] << ' ';
            }
            out << '\n';
        }
    } else {
        // legacy binary VTK: raw big-endian values after the header line, then a newline
        // (format rules, LBM_functions.cpp:1515-1570)
        std::vector<char> buf;
        buf.reserve((size_t)n * nv * (dbl ? 8 : 4));
        auto push = [&](const void* p, int bytes) {
            const unsigned char* b = (const unsigned char*)p;
            for (int i = bytes - 1; i >= 0; --i) buf.push_back((char)b[i]);   // the host is little-endian
        };
        for (long long s = 0; s < n; ++s) {
            double v[3];
            site_values(s, v);
            for (int a = 0; a < nv; ++a) {
                if (dbl) {
                    push(&v[a], 8);
                } else {
                    const float f = (float)v[a];
                    push(&f, 4);
                }
            }
        }
        out.write(buf.data(), (std::streamsize)buf.size());
        out << '\n';
    }
    return out.good() ? LBM_OK : LBM_E_ARG;
}

int lbm_set_reference(lbm_ctx* c, int id, double rho_red, double rho_blue)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_set_reference");
    if (!(std::isfinite(rho_red) && std::isfinite(rho_blue) && rho_red > 0.0 && rho_blue > 0.0))
        return fail(c, LBM_E_ARG, "reference densities must be finite and positive");
    CaseState& k = c->cases[id];
    if (k.elem == 8) return LBM_OK;   // FP64 stores the populations themselves
    const double mid = 0.5 * (rho_red + rho_blue) - 1.0, half = 0.5 * (rho_red - rho_blue);
    HIPCHK(c, launch_lp_recentre(k.elem, k.sym, k.pop[k.cur], (long long)k.nslot * k.cc.px * k.d.ny, k.cc.px, k.cc.ref_mid, k.cc.ref_half, mid, half,
                                 c->stream));
    k.cc.ref_mid = mid;
    k.cc.ref_half = half;
    c->family[k.kind].written = false;
    c->single[id].written = false;
    return LBM_OK;
}

int lbm_download_aux(lbm_ctx* c, int id, double* gm, double* F)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_download_aux");
    CaseState& k = c->cases[id];
    if (!k.aux_valid) return fail(c, LBM_E_STATE, "no recorded step (use LBM_STEP_RECORD_AUX)");
    std::vector<double> h(6 * (size_t)k.S);
    HIPCHK(c, hipMemcpyAsync(h.data(), k.aux, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (long long p = 0; p < k.n; ++p)
        for (int j = 0; j < 3; ++j) {
            if (gm) gm[3 * p + j] = h[(size_t)j * k.S + p];
            if (F) F[3 * p + j] = h[(size_t)(3 + j) * k.S + p];
        }
    return LBM_OK;
}

int lbm_diag_drop(lbm_ctx* c, int id, double v0, int step, double x_ini, double y_ini, lbm_drop_diag* o)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_diag_drop");
    if (!o) return fail(c, LBM_E_ARG, "null output");
    CaseState& k = c->cases[id];
    if (k.solid) return fail(c, LBM_E_STATE, "the drop diagnostics measure an all-fluid box (Imm_OPS.cpp:123-127)");
    const bool forced = k.d.op >= OP_F1;
    if (forced && !k.aux_valid) return fail(c, LBM_E_STATE, "force operators need a step with LBM_STEP_RECORD_AUX");
    if (int r = ensure_diag(c, k)) return r;
    std::memset(o, 0, sizeof(*o));
    const int nx = k.d.nx, ny = k.d.ny;
    const double* fR = nullptr;
    if (int r = state_f64(c, k, &fR)) return r;
    const double* fB = fR + Q * k.S;
    const double* F = forced ? k.aux + 3 * k.S : nullptr;

    // ---- pass 1
    DropPass1Args a1{};
    a1.fR = fR; a1.fB = fB; a1.S = k.S; a1.nx = nx; a1.ny = ny; a1.F = F;
    a1.cosx = k.tab; a1.sinx = k.tab + nx; a1.cosy = k.tab + 2 * nx; a1.siny = k.tab + 2 * nx + ny;
    a1.gal = k.gal; a1.dsum = k.dsum; a1.aphi = k.ordered; a1.partial = k.partial;
    HIPCHK(c, launch_drop_pass1(a1, k.nblk, c->stream));
    // the sums that place the drop (area, circular centroid) in the reference's site order
    {
        OrderedArgs ao{};
        ao.aphi = k.ordered; ao.n = k.n; ao.nx = nx;
        ao.cosx = a1.cosx; ao.sinx = a1.sinx; ao.cosy = a1.cosy; ao.siny = a1.siny;
        ao.out = k.ordered + k.S;
        HIPCHK(c, launch_drop_ordered(ao, c->stream));
    }
    std::vector<double> r1;
    if (int r = reduce_partials<P1_NP>(c, k, r1, P1_NS, P1_NX, P1_NN)) return r;
    {
        double h[5];
        HIPCHK(c, hipMemcpyAsync(h, k.ordered + k.S, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        r1[P1_AREA] = h[0];
        r1[P1_CXC] = h[1]; r1[P1_CXS] = h[2];
        r1[P1_CYC] = h[3]; r1[P1_CYS] = h[4];
    }
    const double n_fluid = r1[P1_NFLUID];
    const double tx = 2.0 * PI / (double)nx, ty = 2.0 * PI / (double)ny;
    double x0 = std::atan2(r1[P1_CXS], r1[P1_CXC]) / tx;
    if (x0 < 0.0) x0 += nx;
    double y0 = std::atan2(r1[P1_CYS], r1[P1_CYC]) / ty;
    if (y0 < 0.0) y0 += ny;
    const double area = r1[P1_AREA];
    const double raio = (area > 0.0) ? std::sqrt(area / (PI * (double)1)) : 0.0;
    o->min_f = r1[P1_NS + P1_NX + 0];
    o->min_rho = r1[P1_NS + P1_NX + 1];
    o->min_phi = r1[P1_NS + P1_NX + 2];
    o->max_phi = r1[P1_NS + 1];
    o->negative_total = (long long)r1[P1_NEGT];
    o->negative_color = (long long)r1[P1_NEGC];
    o->axis_ratio = 1;
    const double bad = r1[P1_BAD], mass = r1[P1_MASS], mass_r = r1[P1_MASSR];
    o->n_ruim = (int)bad;
    if (bad > 0 || !(mass > 0) || !(mass_r > 0)) {
        o->ok = 0;
        return LBM_OK;
    }
    o->mean_x = r1[P1_QX] / mass; o->mean_y = r1[P1_QY] / mass; o->mean_z = r1[P1_QZ] / mass;
    o->drop_x = r1[P1_QRX] / mass_r; o->drop_y = r1[P1_QRY] / mass_r;

    // ---- pass 2
    const double r_in_max = 0.5 * raio;
    double r_out_min = 2.0 * raio;
    if (r_out_min < raio + 4.0 * 3.45) r_out_min = raio + 4.0 * 3.45;
    const double r_caixa = 0.45 * (double)((nx < ny) ? nx : ny);
    if (r_out_min > r_caixa) r_out_min = r_caixa;
    int y_lin = (int)(y0 + 0.5);
    if (y_lin >= ny) y_lin -= ny;
    const int y_line = ((int)std::floor(y0 + 0.5)) % ny;
    const int corner_x = ((int)std::floor(x0 + 0.5) + nx / 2) % nx;
    const int corner_y = ((int)std::floor(y0 + 0.5) + ny / 2) % ny;
    DropPass2Args a2{};
    a2.fR = fR; a2.fB = fB; a2.S = k.S; a2.nx = nx; a2.ny = ny; a2.gal = k.gal; a2.dsum = k.dsum;
    a2.x0 = x0; a2.y0 = y0; a2.gx0 = x0; a2.gy0 = y0; a2.y_lin = y_lin; a2.y_line = y_line;
    a2.r_in_max = r_in_max; a2.r_out_min = r_out_min;
    a2.v0 = v0; a2.mean_x = o->mean_x; a2.mean_y = o->mean_y; a2.mean_z = o->mean_z;
    a2.pc = corner_x + (long long)nx * corner_y;
    a2.n_casca = k.n_casca; a2.hist = k.hist; a2.histn = k.histn; a2.partial = k.partial;
    HIPCHK(c, hipMemsetAsync(k.hist, 0, sizeof(double) * k.n_casca, c->stream));
    HIPCHK(c, hipMemsetAsync(k.histn, 0, sizeof(unsigned long long) * k.n_casca, c->stream));
    HIPCHK(c, launch_drop_pass2(a2, k.nblk, c->stream));
    std::vector<double> r2;
    if (int r = reduce_partials<P2_NP>(c, k, r2, P2_NS, P2_NX, 0)) return r;

    // calc_drop_radius closing (Other_functions.cpp:253-273)
    o->x0 = x0; o->y0 = y0; o->area = area; o->raio = raio;
    o->raio_grad = (r2[P2_SOMW] > 0.0) ? r2[P2_SOMWR] / r2[P2_SOMW] : 0.0;
    o->massa_R = r1[P1_MASSA_R]; o->massa_B = r1[P1_MASSA_B];
    o->qx = r1[P1_QXD]; o->qy = r1[P1_QYD]; o->qz = r1[P1_QZD];
    // calc_laplace_tension closing (Other_functions.cpp:530-551)
    const double n_in = r2[P2_NIN], n_out = r2[P2_NOUT], n_ctr = r2[P2_NCTR];
    const double p_in = (n_in > 0) ? r2[P2_PIN] / n_in : 0.0;
    const double p_out = (n_out > 0) ? r2[P2_POUT] / n_out : 0.0;
    const double var_in = (n_in > 0) ? r2[P2_PIN2] / n_in - p_in * p_in : 0.0;
    const double var_out = (n_out > 0) ? r2[P2_POUT2] / n_out - p_out * p_out : 0.0;
    o->p_in = p_in; o->sd_in = (var_in > 0.0) ? std::sqrt(var_in) : 0.0;
    o->p_out = p_out; o->sd_out = (var_out > 0.0) ? std::sqrt(var_out) : 0.0;
    o->n_in = (int)n_in; o->n_out = (int)n_out;
    o->delta_p = p_in - p_out;
    o->sigma = 1.0 * raio * o->delta_p;
    o->p_centro = (n_ctr > 0) ? r2[P2_PCTR] / n_ctr : 0.0;
    // calc_galilean closing (Galilean_diagnostics.cpp:88-104)
    const double nn = (double)k.n;
    o->max_ref = r2[P2_NS + 0]; o->max_mean = r2[P2_NS + 1]; o->max_lab = r2[P2_NS + 2];
    o->rms_ref = std::sqrt(r2[P2_SREF] / nn);
    o->rms_mean = std::sqrt(r2[P2_SMEAN] / nn);
    o->rms_lab = std::sqrt(r2[P2_SLAB] / nn);
    const double weight = r2[P2_WEIGHT], sx = r2[P2_SX], sy = r2[P2_SY];
    const double xx = r2[P2_SXX] / weight - (sx / weight) * (sx / weight);
    const double yy = r2[P2_SYY] / weight - (sy / weight) * (sy / weight);
    const double xy = r2[P2_SXY] / weight - (sx / weight) * (sy / weight), trace = xx + yy;
    const double disc = std::sqrt((xx - yy) * (xx - yy) + 4 * xy * xy);
    o->axis_ratio = (trace > disc) ? std::sqrt((trace + disc) / (trace - disc)) : INFINITY;
    o->center_dx = gi_image(x0 - (x_ini + v0 * step), nx);
    o->center_dy = gi_image(y0 - y_ini, ny);
    o->x_unwrapped = x_ini + v0 * step + o->center_dx;
    const double lw = r2[P2_LW], rw = r2[P2_RW];
    o->raio_linha = (lw > 0 && rw > 0) ? 0.5 * (r2[P2_RX] / rw - r2[P2_LX] / lw) : 0;
    o->p_canto = r2[P2_PCANTO];
    o->sigma_linha = o->raio_linha * (o->p_centro - o->p_canto);
    o->u_max = o->max_ref;
    o->u_rms = o->rms_ref;
    o->ok = std::isfinite(o->sigma) && std::isfinite(o->max_ref) && std::isfinite(o->axis_ratio);
    (void)n_fluid;
    if (!o->ok) return LBM_OK;

    // ---- calc_perfil_logistico closing (Other_functions.cpp:381-445), from the shell histogram
    {
        std::vector<double> hs(k.n_casca);
        std::vector<unsigned long long> hn(k.n_casca);
        HIPCHK(c, hipMemcpyAsync(hs.data(), k.hist, sizeof(double) * k.n_casca, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(hn.data(), k.histn, sizeof(unsigned long long) * k.n_casca, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double eps = 1.0e-30, d_casca = 0.5;
        double S0 = 0.0, Sr = 0.0, Sl = 0.0, Srr = 0.0, Srl = 0.0;
        int n_usadas = 0;
        for (int kk = 0; kk < k.n_casca; kk++) {
            const double sn = (double)hn[kk];
            if (sn < 1.0) continue;
            const double phi = hs[kk] / sn;
            if (phi <= 0.02 || phi >= 0.98) continue;
            const double r = ((double)kk + 0.5) * d_casca;
            const double l = std::log(phi / (1.0 - phi));
            S0 += 1.0; Sr += r; Sl += l; Srr += r * r; Srl += r * l;
            n_usadas++;
        }
        double K = 0.0, r_meio = 0.0, res_max = 0.0;
        if (n_usadas >= 3) {
            const double den = S0 * Srr - Sr * Sr;
            if (std::fabs(den) > eps) {
                const double a = (S0 * Srl - Sr * Sl) / den;
                const double b = (Sl - a * Sr) / S0;
                K = -a;
                if (std::fabs(a) > eps) r_meio = -b / a;
                for (int kk = 0; kk < k.n_casca; kk++) {
                    const double sn = (double)hn[kk];
                    if (sn < 1.0) continue;
                    const double phi = hs[kk] / sn;
                    if (phi <= 0.02 || phi >= 0.98) continue;
                    const double r = ((double)kk + 0.5) * d_casca;
                    const double dd = std::log(phi / (1.0 - phi)) - (a * r + b);
                    if (std::fabs(dd) > res_max) res_max = std::fabs(dd);
                }
            }
        }
        o->K_perfil = K;
        o->r_perfil = r_meio;
        o->res_perfil = res_max;
        o->esp_perfil = (K > eps) ? 2.0 * std::log(9.0) / K : 0.0;
        o->n_perfil = n_usadas;
    }

    // ---- modos_interface (Operadores_tensao.cpp:382-449)
    {
        double* ct = k.tab + 2 * nx + 2 * ny;
        ModesArgs am{};
        am.dsum = k.dsum; am.S = k.S; am.nx = nx; am.ny = ny;
        am.x0 = x0; am.y0 = y0; am.R_est = raio;
        am.ct = ct; am.st = ct + N_ANG; am.r = ct + 2 * N_ANG;
        HIPCHK(c, launch_drop_modes(am, c->stream));
        std::vector<double> rr(N_ANG);
        HIPCHK(c, hipMemcpyAsync(rr.data(), am.r, sizeof(double) * N_ANG, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double c2 = 0, s2 = 0, c4 = 0, s4 = 0, c8 = 0, s8 = 0, soma = 0, rmin = 1e30, rmax = 0;
        for (int kk = 0; kk < N_ANG; kk++) {
            const double th = 2.0 * M_PI * kk / N_ANG;
            const double r = rr[kk];
            soma += r;
            c2 += r * std::cos(2 * th); s2 += r * std::sin(2 * th);
            c4 += r * std::cos(4 * th); s4 += r * std::sin(4 * th);
            c8 += r * std::cos(8 * th); s8 += r * std::sin(8 * th);
            rmin = (r < rmin) ? r : rmin;
            rmax = (rmax < r) ? r : rmax;
        }
        o->R0 = soma / N_ANG;
        o->a2 = 2.0 * std::sqrt(c2 * c2 + s2 * s2) / N_ANG / o->R0;
        o->a4 = 2.0 * std::sqrt(c4 * c4 + s4 * s4) / N_ANG / o->R0;
        o->a8 = 2.0 * std::sqrt(c8 * c8 + s8 * s8) / N_ANG / o->R0;
        o->rmin = rmin;
        o->rmax = rmax;
    }
    return LBM_OK;
}

int lbm_diag_incl(lbm_ctx* c, int id, int p, int q, lbm_incl_diag* o, double* profile)
{
    if (int r = check_case(c, id)) return r;
    if (capturing(c)) return refuse_capture(c, "lbm_diag_incl");
    if (!o) return fail(c, LBM_E_ARG, "null output");
    if (p == 0 && q == 0) return fail(c, LBM_E_ARG, "p = q = 0");
    CaseState& k = c->cases[id];
    if (k.solid) return fail(c, LBM_E_STATE, "the band diagnostics measure an all-fluid box (Imm_INCL.cpp:241-246)");
    if (!k.aux_valid) return fail(c, LBM_E_STATE, "needs a step with LBM_STEP_RECORD_AUX");
    if (int r = ensure_diag(c, k)) return r;
    if (profile && !k.prof) {
        if (int r = dev_alloc(c, (void**)&k.prof, sizeof(double) * 5 * (size_t)k.S)) return r;
    }
    std::memset(o, 0, sizeof(*o));
    const int nx = k.d.nx, ny = k.d.ny;
    const long long M = gcdll((long long)p * nx, (long long)q * ny) / gcdll(p, q);
    const double m = std::sqrt((double)(p * p + q * q));
    const double* f64 = nullptr;
    if (int r = state_f64(c, k, &f64)) return r;
    InclArgs a{};
    a.fR = f64; a.fB = f64 + Q * k.S; a.S = k.S; a.nx = nx; a.ny = ny;
    a.F = (k.d.op >= OP_F1) ? k.aux + 3 * k.S : nullptr;
    a.gm = k.aux;
    a.p = p; a.q = q; a.M = M; a.m = m;
    a.nxn = p / m; a.nyn = q / m; a.txn = -a.nyn; a.tyn = a.nxn;
    a.fit_pass = 0;
    a.prof = profile ? k.prof : nullptr;
    a.partial = k.partial;
    HIPCHK(c, launch_incl_pass(a, k.nblk, c->stream));
    std::vector<double> r1;
    if (int r = reduce_partials<I1_NP>(c, k, r1, I1_NS, I1_NX, I1_NN)) return r;
    const double comprimento = 2.0 * ((double)nx * (double)ny / (double)M) * m;
    o->n_ruim = (int)r1[I1_NRUIM];
    o->neg_total = (long long)r1[I1_NEGT];
    o->neg_cor = (long long)r1[I1_NEGC];
    o->min_f = r1[I1_NS + I1_NX];
    o->SFx = r1[I1_SFX]; o->SFy = r1[I1_SFY];
    o->massa_R = r1[I1_MR]; o->massa_B = r1[I1_MB]; o->qx = r1[I1_QX]; o->qy = r1[I1_QY];
    o->u_max = r1[I1_NS + 0]; o->un_max = r1[I1_NS + 1]; o->ut_max = r1[I1_NS + 2];
    o->sigma_pop = r1[I1_SIGMA] / comprimento;
    o->int_grad = r1[I1_IGRAD] / comprimento;
    o->u_rms = std::sqrt(r1[I1_U2] / ((double)nx * (double)ny));
    const int nfit = (int)r1[I1_NFIT];
    o->n_ajuste = nfit;
    if (nfit > 3) {
        const double Sx = r1[I1_SX], Sy = r1[I1_SY], Sxx = r1[I1_SXX], Sxy = r1[I1_SXY];
        const double K = (nfit * Sxy - Sx * Sy) / (nfit * Sxx - Sx * Sx);
        const double aa = (Sy - K * Sx) / nfit;
        o->K = K;
        o->d0 = -aa / K;
        a.fit_pass = 1; a.fa = aa; a.fK = K; a.prof = nullptr;
        HIPCHK(c, launch_incl_pass(a, k.nblk, c->stream));
        std::vector<double> r2;
        if (int r = reduce_partials<I1_NP>(c, k, r2, I1_NS, I1_NX, I1_NN)) return r;
        o->res = r2[I1_NS + 3];
    }
    o->ok = (o->n_ruim == 0) && std::isfinite(o->u_max);
    if (profile) {
        std::vector<double> h(5 * (size_t)k.S);
        HIPCHK(c, hipMemcpyAsync(h.data(), k.prof, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (long long s = 0; s < k.n; ++s)
            for (int j = 0; j < 5; ++j) profile[5 * s + j] = h[(size_t)j * k.S + s];
    }
    return LBM_OK;
}

}  // extern "C"
