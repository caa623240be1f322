// FP64 diagnostics on the device: per-site passes over the post-collision populations with
// deterministic block reductions (per-thread sequential accumulation in site order, then a fixed
// wavefront-shuffle / LDS tree).  The host adds the block partials in block order and applies the
// reference's closing formulas (capi.hip).
//
// Reference: calc_galilean (Galilean_diagnostics.cpp:24-105), calc_drop_radius /
// calc_laplace_tension / calc_perfil_logistico (Other_functions.cpp:60-554), modos_interface
// (Operadores_tensao.cpp:380-449), mede_inclinada (Imm_INCL.cpp:85-189).
#include <hip/hip_runtime.h>

#include <cmath>

#include "diag_kernels.h"
#include "lattice.h"

namespace itolbm {

namespace {

constexpr double CS2 = 1. / 3.;

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }   // std::min
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }   // std::max

// Reduces NS sums, NX maxima and NN minima over the block (256 threads) and writes them to
// out[0 .. NS+NX+NN).  Fixed shape, so the result is reproducible run to run.
template <int NS, int NX, int NN>
__device__ void block_reduce(double* s, double* mx, double* mn, double* out)
{
    constexpr int NT = NS + NX + NN;
    __shared__ double sh[4][NT > 0 ? NT : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NS; ++k) s[k] = s[k] + __shfl_xor(s[k], off, 64);
#pragma unroll
        for (int k = 0; k < NX; ++k) mx[k] = dmax(mx[k], __shfl_xor(mx[k], off, 64));
#pragma unroll
        for (int k = 0; k < NN; ++k) mn[k] = dmin(mn[k], __shfl_xor(mn[k], off, 64));
    }
    if (lane == 0) {
        for (int k = 0; k < NS; ++k) sh[wave][k] = s[k];
        for (int k = 0; k < NX; ++k) sh[wave][NS + k] = mx[k];
        for (int k = 0; k < NN; ++k) sh[wave][NS + NX + k] = mn[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NS; ++k) out[k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
        for (int k = 0; k < NX; ++k)
            out[NS + k] = dmax(dmax(dmax(sh[0][NS + k], sh[1][NS + k]), sh[2][NS + k]), sh[3][NS + k]);
        for (int k = 0; k < NN; ++k) {
            const int j = NS + NX + k;
            out[j] = dmin(dmin(dmin(sh[0][j], sh[1][j]), sh[2][j]), sh[3][j]);
        }
    }
}

__device__ __forceinline__ void load_site(const double* __restrict__ pR, const double* __restrict__ pB,
                                          long long S, long long site, double fR[Q], double fB[Q])
{
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        fR[i] = pR[(long long)i * S + site];
        fB[i] = pB[(long long)i * S + site];
    }
}

// calcula (LBM_functions.cpp:1136-1169)
__device__ __forceinline__ void calcula(const double f[Q], double& vx, double& vy, double& vz, double& rho)
{
    double r = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        r = r + f[i];
        if (CX[i] > 0) mx = mx + f[i];
        if (CX[i] < 0) mx = mx - f[i];
        if (CY[i] > 0) my = my + f[i];
        if (CY[i] < 0) my = my - f[i];
        if (CZ[i] > 0) mz = mz + f[i];
        if (CZ[i] < 0) mz = mz - f[i];
    }
    if (r != 0.0) {
        const double inv = 1.0 / r;
        vx = mx * inv; vy = my * inv; vz = mz * inv;
    } else {
        vx = 0.; vy = 0.; vz = 0.;
    }
    rho = r;
}

__device__ __forceinline__ double image(double d, int n)
{
    if (d > 0.5 * n) d -= n;
    if (d < -0.5 * n) d += n;
    return d;
}

__device__ __forceinline__ double gi_image(double dx, double length) { return dx - length * floor(dx / length + 0.5); }

// ------------------------------------------------------------------------------ drop, pass 1
__global__ void __launch_bounds__(256) drop_pass1(DropPass1Args a)
{
    double s[P1_NS], mx[P1_NX], mn[P1_NN];
    for (int k = 0; k < P1_NS; ++k) s[k] = 0.0;
    mx[0] = 0.0;  mx[1] = 0.0;                  // u_max (drop radius), max_phi
    mn[0] = INFINITY; mn[1] = INFINITY; mn[2] = 1.0;   // min_f, min_rho, min_phi
    const long long n = (long long)a.nx * a.ny;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        double fR[Q], fB[Q];
        load_site(a.fR, a.fB, a.S, p, fR, fB);
        const int y = (int)(p / a.nx), x = (int)(p - (long long)y * a.nx);
        // calc_drop_radius, pass 1 (Other_functions.cpp:89-178)
        double vxR, vyR, vzR, rho_R, vxB, vyB, vzB, rho_B;
        calcula(fR, vxR, vyR, vzR, rho_R);
        calcula(fB, vxB, vyB, vzB, rho_B);
        s[P1_NFLUID] += 1.0;
        a.aphi[p] = 0.0;
        if (!(rho_R == rho_R) || !(rho_B == rho_B) || !(vxR == vxR) || !(vxB == vxB)) {
            s[P1_NRUIM] += 1.0;
        } else {
            s[P1_MASSA_R] = s[P1_MASSA_R] + rho_R;
            s[P1_MASSA_B] = s[P1_MASSA_B] + rho_B;
            s[P1_QXD] = s[P1_QXD] + rho_R * vxR + rho_B * vxB;
            s[P1_QYD] = s[P1_QYD] + rho_R * vyR + rho_B * vyB;
            s[P1_QZD] = s[P1_QZD] + rho_R * vzR + rho_B * vzB;
            const double rho = rho_R + rho_B;
            if (rho > 0.0) {
                double ux = (rho_R * vxR + rho_B * vxB) / rho;
                double uy = (rho_R * vyR + rho_B * vyB) / rho;
                double uz = (rho_R * vzR + rho_B * vzB) / rho;
                if (a.F) {
                    const double F0 = a.F[p], F1 = a.F[a.S + p], F2 = a.F[2 * a.S + p];
                    ux = ux - 0.5 * F0 / rho; uy = uy - 0.5 * F1 / rho; uz = uz - 0.5 * F2 / rho;
                    s[P1_QXD] = s[P1_QXD] - 0.5 * F0;
                    s[P1_QYD] = s[P1_QYD] - 0.5 * F1;
                    s[P1_QZD] = s[P1_QZD] - 0.5 * F2;
                }
                const double u2 = ux * ux + uy * uy + uz * uz;
                s[P1_U2] = s[P1_U2] + u2;
                const double u = sqrt(u2);
                if (u > mx[0]) mx[0] = u;
                const double phi = rho_R / rho;
                a.aphi[p] = phi;
                s[P1_AREA] = s[P1_AREA] + phi;
                s[P1_CXC] = s[P1_CXC] + phi * a.cosx[x];
                s[P1_CXS] = s[P1_CXS] + phi * a.sinx[x];
                s[P1_CYC] = s[P1_CYC] + phi * a.cosy[y];
                s[P1_CYS] = s[P1_CYS] + phi * a.siny[y];
            }
        }
        // calc_galilean, pass 1 (Galilean_diagnostics.cpp:31-60)
        double rr = 0, rb = 0, mxx = 0, myy = 0, mzz = 0;
        double bad = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const double fr = fR[i], fb = fB[i];
            const double f = fr + fb;
            if (!isfinite(fr) || !isfinite(fb)) bad += 1.0;
            mn[0] = dmin(mn[0], f);
            if (f < -1.e-12) s[P1_NEGT] += 1.0;
            if (fr < -1.e-12) s[P1_NEGC] += 1.0;
            if (fb < -1.e-12) s[P1_NEGC] += 1.0;
            rr += fr; rb += fb;
            if (CX[i] > 0) mxx += f;
            if (CX[i] < 0) mxx += -f;
            if (CY[i] > 0) myy += f;
            if (CY[i] < 0) myy += -f;
            if (CZ[i] > 0) mzz += f;
            if (CZ[i] < 0) mzz += -f;
        }
        const double rho = rr + rb;
        a.dsum[p] = rr - rb;
        a.dsum[a.S + p] = rr + rb;
        bool good = true;
        if (!isfinite(rho) || rho <= 0 || !isfinite(mxx) || !isfinite(myy) || !isfinite(mzz)) {
            bad += 1.0; good = false;
        } else if (a.F) {
            const double F0 = a.F[p], F1 = a.F[a.S + p], F2 = a.F[2 * a.S + p];
            if (!isfinite(F0) || !isfinite(F1) || !isfinite(F2)) {
                bad += 1.0; good = false;
            } else {
                mxx -= 0.5 * F0; myy -= 0.5 * F1; mzz -= 0.5 * F2;
            }
        }
        s[P1_BAD] += bad;
        double ux = 0, uy = 0, uz = 0, ph = 0;
        if (good) {
            mn[1] = dmin(mn[1], rho);
            ux = mxx / rho; uy = myy / rho; uz = mzz / rho; ph = rr / rho;
            mn[2] = dmin(mn[2], ph);
            mx[1] = dmax(mx[1], ph);
            s[P1_MASS] += rho; s[P1_MASSR] += rr;
            s[P1_QX] += mxx; s[P1_QY] += myy; s[P1_QZ] += mzz;
            s[P1_QRX] += rr * ux; s[P1_QRY] += rr * uy;
        }
        a.gal[p] = ux; a.gal[a.S + p] = uy; a.gal[2 * a.S + p] = uz; a.gal[3 * a.S + p] = ph;
    }
    block_reduce<P1_NS, P1_NX, P1_NN>(s, mx, mn, a.partial + (long long)blockIdx.x * P1_NP);
}

// One wavefront; lane k < 5 accumulates sum k over the sites in order p = 0 .. n-1 (x fastest), the
// order of the reference's loop, so each sum is rounded exactly as there.  A site that adds nothing
// adds +0.0 or -0.0, which leaves a running sum that starts at +0.0 unchanged.
__global__ void __launch_bounds__(64) drop_ordered(OrderedArgs a)
{
    const int k = threadIdx.x;
    if (k >= 5) return;
    const double* tab = (k == 1) ? a.cosx : (k == 2) ? a.sinx : (k == 3) ? a.cosy : a.siny;
    double acc = 0.0;
    int x = 0, y = 0;
    long long p = 0;
    constexpr int U = 32;   // loads of U sites are issued before their additions
    for (; p + U <= a.n; p += U) {
        double t[U], w[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            t[j] = a.aphi[p + j];
            w[j] = (k == 0) ? 1.0 : (k <= 2 ? tab[x] : tab[y]);
            if (++x == a.nx) { x = 0; ++y; }
        }
#pragma unroll
        for (int j = 0; j < U; ++j) acc = (k == 0) ? acc + t[j] : acc + t[j] * w[j];
    }
    for (; p < a.n; ++p) {
        const double t = a.aphi[p];
        const double w = (k == 0) ? 1.0 : (k <= 2 ? tab[x] : tab[y]);
        acc = (k == 0) ? acc + t : acc + t * w;
        if (++x == a.nx) { x = 0; ++y; }
    }
    a.out[k] = acc;
}

// ------------------------------------------------------------------------------ drop, pass 2
__global__ void __launch_bounds__(256) drop_pass2(DropPass2Args a)
{
    double s[P2_NS], mx[P2_NX], mn[1];
    for (int k = 0; k < P2_NS; ++k) s[k] = 0.0;
    for (int k = 0; k < P2_NX; ++k) mx[k] = 0.0;
    const long long n = (long long)a.nx * a.ny;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const int y = (int)(p / a.nx), x = (int)(p - (long long)y * a.nx);
        const double sumR = a.dsum[a.S + p];
        // densities of the two colours again, in the reference's order
        double rho_R = 0.0, rho_B = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            rho_R = rho_R + a.fR[(long long)i * a.S + p];
            rho_B = rho_B + a.fB[(long long)i * a.S + p];
        }
        (void)sumR;
        const bool fin = (rho_R == rho_R) && (rho_B == rho_B);
        const double dx = image((double)x - a.x0, a.nx);
        const double dy = image((double)y - a.y0, a.ny);
        const double dz = 0.0;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (fin) {
            // calc_drop_radius, pass 2 (Other_functions.cpp:213-249)
            const double w = rho_R * rho_B;
            s[P2_SOMW] = s[P2_SOMW] + w;
            s[P2_SOMWR] = s[P2_SOMWR] + w * r;
            if (y == a.y_lin) {
                if (dx < 0.0) { s[P2_SXE] = s[P2_SXE] + w * (double)x; s[P2_SWE] = s[P2_SWE] + w; }
                else { s[P2_SXD] = s[P2_SXD] + w * (double)x; s[P2_SWD] = s[P2_SWD] + w; }
            }
            // calc_laplace_tension (Other_functions.cpp:490-526)
            const double pr = CS2 * (rho_R + rho_B);
            if (r < a.r_in_max) { s[P2_PIN] = s[P2_PIN] + pr; s[P2_PIN2] = s[P2_PIN2] + pr * pr; s[P2_NIN] += 1.0; }
            else if (r >= a.r_out_min) { s[P2_POUT] = s[P2_POUT] + pr; s[P2_POUT2] = s[P2_POUT2] + pr * pr; s[P2_NOUT] += 1.0; }
            if (r < 1.5) { s[P2_PCTR] = s[P2_PCTR] + pr; s[P2_NCTR] += 1.0; }
            if (p == 0) s[P2_PCNT] = s[P2_PCNT] + pr;
        }
        // calc_perfil_logistico shells (Other_functions.cpp:341-377)
        {
            const double rho = rho_R + rho_B;
            if (!(rho <= 1.0e-30)) {
                const int k = (int)(r / 0.5);
                if (k >= 0 && k < a.n_casca) {
                    atomicAdd(a.hist + k, rho_R / rho);
                    atomicAdd(a.histn + k, 1ull);
                }
            }
        }
        // calc_galilean, pass 2 (Galilean_diagnostics.cpp:68-87)
        {
            const double ux = a.gal[p], uy = a.gal[a.S + p], uz = a.gal[2 * a.S + p], phi = a.gal[3 * a.S + p];
            const double A = (ux - a.v0) * (ux - a.v0) + uy * uy + uz * uz;
            const double B = (ux - a.mean_x) * (ux - a.mean_x) + (uy - a.mean_y) * (uy - a.mean_y) +
                             (uz - a.mean_z) * (uz - a.mean_z);
            const double Cc = ux * ux + uy * uy + uz * uz;
            s[P2_SREF] += A; s[P2_SMEAN] += B; s[P2_SLAB] += Cc;
            mx[0] = dmax(mx[0], sqrt(A));
            mx[1] = dmax(mx[1], sqrt(B));
            mx[2] = dmax(mx[2], sqrt(Cc));
            const double gx = gi_image(x - a.gx0, a.nx), gy = gi_image(y - a.gy0, a.ny);
            s[P2_WEIGHT] += phi; s[P2_SX] += phi * gx; s[P2_SY] += phi * gy;
            s[P2_SXX] += phi * gx * gx; s[P2_SYY] += phi * gy * gy; s[P2_SXY] += phi * gx * gy;
            if (y == a.y_line) {
                const double w = rho_R * rho_B;
                if (gx < 0) { s[P2_LW] += w; s[P2_LX] += w * gx; }
                else { s[P2_RW] += w; s[P2_RX] += w * gx; }
            }
            if (p == a.pc) s[P2_PCANTO] = rho_R * CS2 + rho_B * CS2;
        }
    }
    block_reduce<P2_NS, P2_NX, 0>(s, mx, mn, a.partial + (long long)blockIdx.x * P2_NP);
}

// ------------------------------------------------------------------------------ Fourier modes
// modos_interface (Operadores_tensao.cpp:382-441): one thread per ray, 60 bisection steps on the
// bilinear interpolation of rho^N; dsum holds rR - rB and rR + rB per site.
__global__ void __launch_bounds__(256) drop_modes(ModesArgs a)
{
    const int k = threadIdx.x;
    if (k >= 256) return;
    const int nx = a.nx, ny = a.ny;
    auto fase_em = [&](double x, double y) {
        x = x - nx * floor(x / nx);
        y = y - ny * floor(y / ny);
        const int i0 = (int)floor(x), j0 = (int)floor(y);
        const double tx = x - i0, ty = y - j0;
        double v = 0.0;
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                const int ii = (i0 + di) % nx, jj = (j0 + dj) % ny;
                const long long p = ii + (long long)jj * nx;
                const double wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty);
                v += wgt * a.dsum[p] / a.dsum[a.S + p];
            }
        return v;
    };
    const double ct = a.ct[k], st = a.st[k];
    double lo = 0.3 * a.R_est, hi = 1.7 * a.R_est;
    const double fa = fase_em(a.x0 + lo * ct, a.y0 + lo * st);
    for (int it = 0; it < 60; it++) {
        const double mid = 0.5 * (lo + hi);
        const double fm = fase_em(a.x0 + mid * ct, a.y0 + mid * st);
        if ((fm > 0) == (fa > 0)) lo = mid;
        else hi = mid;
    }
    a.r[k] = 0.5 * (lo + hi);
}

// ------------------------------------------------------------------------------ inclined interface
__device__ __forceinline__ double dist_central(int x, int y, int p, int q, long long M, double m)
{
    long long s = ((long long)p * x + (long long)q * y) % M;
    if (s < 0) s += M;
    double d = ((double)s - 0.5 * M);
    if (d >= 0.5 * M) d -= M;
    if (d < -0.5 * M) d += M;
    return -d / m;
}

// mede_inclinada, site pass (Imm_INCL.cpp:103-170)
__global__ void __launch_bounds__(256) incl_pass1(InclArgs a)
{
    double s[I1_NS], mx[I1_NX], mn[I1_NN];
    for (int k = 0; k < I1_NS; ++k) s[k] = 0.0;
    for (int k = 0; k < I1_NX; ++k) mx[k] = 0.0;
    mn[0] = 1e300;
    const long long n = (long long)a.nx * a.ny;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const int y = (int)(p / a.nx), x = (int)(p - (long long)y * a.nx);
        double fR[Q], fB[Q];
        load_site(a.fR, a.fB, a.S, p, fR, fB);
        double rR = 0, rB = 0, jx = 0, jy = 0, Pxx = 0, Pyy = 0, Pxy = 0, Pzz = 0;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const double f = fR[i] + fB[i];
            if (!isfinite(fR[i]) || !isfinite(fB[i])) s[I1_NRUIM] += 1.0;
            if (f < -1e-12) s[I1_NEGT] += 1.0;
            if (fR[i] < -1e-12) s[I1_NEGC] += 1.0;
            if (fB[i] < -1e-12) s[I1_NEGC] += 1.0;
            mn[0] = dmin(mn[0], f);
            rR += fR[i]; rB += fB[i];
            if (CX[i] > 0) jx += f;
            if (CX[i] < 0) jx += -f;
            if (CY[i] > 0) jy += f;
            if (CY[i] < 0) jy += -f;
            if (CX[i] != 0) Pxx += f;
            if (CY[i] != 0) Pyy += f;
            if (CX[i] * CY[i] > 0) Pxy += f;
            if (CX[i] * CY[i] < 0) Pxy += -f;
            if (CZ[i] != 0) Pzz += f;
        }
        const double rho = rR + rB;
        if (a.prof) {
            for (int k = 0; k < 5; ++k) a.prof[k * a.S + p] = NAN;
        }
        if (!(rho > 0)) { s[I1_NRUIM] += 1.0; continue; }
        if (a.F) {
            const double F0 = a.F[p], F1 = a.F[a.S + p];
            jx -= 0.5 * F0; jy -= 0.5 * F1;
            s[I1_SFX] += F0; s[I1_SFY] += F1;
        }
        const double ux = jx / rho, uy = jy / rho;
        s[I1_MR] += rR; s[I1_MB] += rB; s[I1_QX] += jx; s[I1_QY] += jy;
        const double u = hypot(ux, uy), un = fabs(ux * a.nxn + uy * a.nyn), ut = fabs(ux * a.txn + uy * a.tyn);
        mx[0] = dmax(mx[0], u); mx[1] = dmax(mx[1], un); mx[2] = dmax(mx[2], ut);
        s[I1_U2] += u * u;
        const double Pnn = a.nxn * a.nxn * Pxx + 2 * a.nxn * a.nyn * Pxy + a.nyn * a.nyn * Pyy;
        const double Ptt = a.txn * a.txn * Pxx + 2 * a.txn * a.tyn * Pxy + a.tyn * a.tyn * Pyy;
        s[I1_SIGMA] += Pnn - 0.5 * (Ptt + Pzz);
        const double g0 = a.gm[p], g1 = a.gm[a.S + p], g2 = a.gm[2 * a.S + p];
        s[I1_IGRAD] += sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        const double d = dist_central(x, y, a.p, a.q, a.M, a.m);
        const double cR = rR / rho;
        if (a.prof) {
            a.prof[p] = d; a.prof[a.S + p] = cR; a.prof[2 * a.S + p] = rho;
            a.prof[3 * a.S + p] = ux; a.prof[4 * a.S + p] = uy;
        }
        if (cR > 0.02 && cR < 0.98 && fabs(d) < 8.0) {
            const double l = log(cR / (1.0 - cR));
            s[I1_SX] += d; s[I1_SY] += l; s[I1_SXX] += d * d; s[I1_SXY] += d * l; s[I1_NFIT] += 1.0;
            if (a.fit_pass) mx[3] = dmax(mx[3], fabs(l - (a.fa + a.fK * d)));
        }
    }
    block_reduce<I1_NS, I1_NX, I1_NN>(s, mx, mn, a.partial + (long long)blockIdx.x * I1_NP);
}

}  // namespace

hipError_t launch_drop_pass1(const DropPass1Args& a, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(drop_pass1, dim3(nblk), dim3(256), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_drop_ordered(const OrderedArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(drop_ordered, dim3(1), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_drop_pass2(const DropPass2Args& a, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(drop_pass2, dim3(nblk), dim3(256), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_drop_modes(const ModesArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(drop_modes, dim3(1), dim3(256), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_incl_pass(const InclArgs& a, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(incl_pass1, dim3(nblk), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace itolbm
