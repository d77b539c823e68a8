// Spectral (S) operator: 8th-order staggered periodic stencil, its medium
// set-up, the plane-wave incident field and the scattered-field right-hand
// side.
//
// Operator (nsrno/wave_operator.py:105-109, applied there with six FFTs per
// call; here as the equivalent real-space stencil, whose Fourier symbol is
// the reference's D = 2i sum_m a_m sin((m+1/2) q) e^{iq/2} / h):
//   gx_j = ax_j (D+ u)_j,  (D+ u)_j = sum_m a_m (u_{j+m+1} - u_{j-m}) / h
//   (D- g)_j = sum_m a_m (g_{j+m} - g_{j-m-1}) / h       (D- = -(D+)^T)
//   L u = [D-x gx + D-z gz] / k0^2 + kappa_S u,  all indices periodic.
// The plan stores A = k0^2 conj(L): real face coefficients, kappa_c =
// conj(kappa_S); the adjoint apply (conj of the mass term) is k0^2 L.
//
// Kernel layout: a workgroup owns a TZ x 64 output tile and loops over its
// group of right-hand sides.  Per RHS it stages u on the tile plus a 7-node
// periodic halo (x halo on the tile rows, z halo on the tile columns; the
// corners are never needed) in LDS, forms the x- and z-face fluxes of the
// tile (64+7 and TZ+7 faces) in LDS, then the divergence.  Face and node
// coefficients are loaded once per workgroup into registers and reused for
// every right-hand side.  Algorithmic bytes per node per RHS:
// 8 (read u) + 8 (write) + (4 + 4 + 8)/R (ax, az, kappa).
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include <mutex>

#include "helm_internal.h"

namespace helm {

constexpr double FD0 = 1225.0 / 1024.0, FD1 = -245.0 / 3072.0, FD2 = 49.0 / 5120.0, FD3 = -5.0 / 7168.0;
constexpr int S8_TX = 64;
constexpr int S8_HALO = 7;

// HELM_STAG8_TILED=1: the S operator's FP32 applies through the LDS-tile
// k_stag8 instead of the streaming kernel (the FP64 forms always take
// k_stag8: 178 us per 11-field true residual on the 1056x1216 box against
// 217 us for the streaming kernel in FP64)
static bool stag8_tiled() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_STAG8_TILED");
        return v && std::atoi(v) == 1;
    }();
    return t;
}

__host__ __device__ __forceinline__ int wrap_index(int v, int m) {
    if (v < 0 || v >= m) {
        v %= m;
        if (v < 0) v += m;
    }
    return v;
}

template <typename R>
__host__ __device__ __forceinline__ R sponge_at(int i, int j, int H, int W, int cells, R amp) {
    const int iz = min(i, H - 1 - i), ix = min(j, W - 1 - j);
    const int d = cells - min(iz, ix);
    if (d <= 0 || cells <= 0) return R(0);
    const R t = R(d) / R(cells);
    return amp * t * t;
}

template <typename C, typename R>
__device__ __forceinline__ C dstag(R c0, R c1, R c2, R c3, C a0, C b0, C a1, C b1, C a2, C b2, C a3, C b3) {
    // c0 (a0 - b0) + c1 (a1 - b1) + c2 (a2 - b2) + c3 (a3 - b3)
    C d;
    d.x = c0 * (a0.x - b0.x) + c1 * (a1.x - b1.x) + c2 * (a2.x - b2.x) + c3 * (a3.x - b3.x);
    d.y = c0 * (a0.y - b0.y) + c1 * (a1.y - b1.y) + c2 * (a2.y - b2.y) + c3 * (a3.y - b3.y);
    return d;
}

// MODE 0: out = A u;  1: out = b - A u;  2: out = -(A - A0) u with A0 the
// background operator (faces 1, kappa_c 1 + i sponge).  conj: conjugate the
// mass term (A^H; the face terms are real).
template <typename R, typename C, int TZ, int MODE>
__global__ __launch_bounds__(256) void k_stag8(int H, int W, int64_t n, int nrhs, int rb, const R* __restrict__ ax,
                                               const R* __restrict__ az, const C* __restrict__ kap, R k0sq, R c0,
                                               R c1, R c2, R c3, int sp_cells, R sp_amp, int conj,
                                               const C* __restrict__ u, C* __restrict__ out,
                                               const C* __restrict__ b) {
    constexpr int TX = S8_TX, HL = S8_HALO;
    constexpr int SU = TX + 2 * HL;  // staged row length
    constexpr int NFX = TX + 7;      // x faces per tile row: -4 .. TX+2
    constexpr int NFZ = TZ + 7;      // z face rows: -4 .. TZ+2
    constexpr int RPT = TZ / 4;      // output rows per thread
    constexpr int FXT = (TZ * NFX + 255) / 256;
    constexpr int FZT = (NFZ * TX + 255) / 256;
    __shared__ C su[(TZ + 2 * HL) * SU];
    __shared__ C sgx[TZ * NFX];
    __shared__ C sgz[NFZ * TX];

    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;

    // face coefficients and kappa of the tile: every load issued before the
    // first is used (indices wrapped or clamped into the arrays, values
    // outside the tile zeroed after), one memory latency for the set-up
    R cfx[FXT], cfz[FZT];
#pragma unroll
    for (int q = 0; q < FXT; ++q) {
        const int e = tid + q * 256;
        const int ec = e < TZ * NFX ? e : 0;
        const int r = ec / NFX, f = ec - r * NFX - 4;
        cfx[q] = ax[(int64_t)wrap_index(z0 + r, H) * W + wrap_index(x0 + f, W)];
    }
#pragma unroll
    for (int q = 0; q < FZT; ++q) {
        const int e = tid + q * 256;
        const int ec = e < NFZ * TX ? e : 0;
        const int fr = ec / TX - 4, c = ec % TX;
        cfz[q] = az[(int64_t)wrap_index(z0 + fr, H) * W + wrap_index(x0 + c, W)];
    }
    C kc[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = z0 + ty + 4 * k, j = x0 + tx;
        kc[k] = kap[(int64_t)(i < H ? i : 0) * W + (j < W ? j : 0)];
    }
#pragma unroll
    for (int q = 0; q < FXT; ++q) {
        R v = tid + q * 256 < TZ * NFX ? cfx[q] : R(0);
        if (MODE == 2 && tid + q * 256 < TZ * NFX) v -= R(1);
        cfx[q] = v;
    }
#pragma unroll
    for (int q = 0; q < FZT; ++q) {
        R v = tid + q * 256 < NFZ * TX ? cfz[q] : R(0);
        if (MODE == 2 && tid + q * 256 < NFZ * TX) v -= R(1);
        cfz[q] = v;
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = z0 + ty + 4 * k, j = x0 + tx;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (i < H && j < W) {
            v = kc[k];
            if (MODE == 2) {
                v.x -= R(1);
                v.y -= sponge_at<R>(i, j, H, W, sp_cells, sp_amp);
            }
            if (conj) v.y = -v.y;
            v.x *= k0sq;
            v.y *= k0sq;
        }
        kc[k] = v;
    }

    // staging map of this thread (the same for every RHS): rows rr = ty + 4 k
    // of the halo-7 tile, columns cc = tx and tx + 64 (the corners are never read)
    constexpr int NR = (TZ + 2 * HL + 3) / 4;
    constexpr bool PF_TILE = sizeof(R) == 4;  // the register tile of the prefetch (FP32)
    int64_t rowoff[NR];
    int colg[2];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int rr = ty + 4 * k;
        rowoff[k] = rr < TZ + 2 * HL ? (int64_t)wrap_index(z0 + rr - HL, H) * W : -1;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int cc = tx + 64 * c;
        colg[c] = cc < SU ? wrap_index(x0 + cc - HL, W) : -1;
    }
    auto staged = [&](int k, int c) __attribute__((always_inline)) {
        const int rr = ty + 4 * k, cc = tx + 64 * c;
        const bool zh = rr < HL || rr >= TZ + HL, xh = cc < HL || cc >= TX + HL;
        return rowoff[k] >= 0 && colg[c] >= 0 && !(zh && xh);
    };
    // the next RHS's tile is loaded into registers while the current one is
    // computed; only LDS barriers separate the phases (a full barrier would
    // wait for those loads)
    C pre[PF_TILE ? NR : 1][2];
    auto load = [&](int r) __attribute__((always_inline)) {
        const C* ur = u + (int64_t)r * n;
#pragma unroll
        for (int k = 0; k < NR; ++k)
#pragma unroll
            for (int c = 0; c < 2; ++c)
                if (staged(k, c)) pre[k][c] = ur[rowoff[k] + colg[c]];
    };
    // (FP32 only: in FP64 the register tile does not fit beside the face
    // coefficients without spilling, so each RHS is staged straight into LDS
    // after the previous one is done)
    constexpr bool PF = sizeof(R) == 4;
    const int r0 = blockIdx.z * rb, r1 = min(nrhs, (int)(blockIdx.z + 1) * rb);
    if constexpr (PF) {
        if (r0 < r1) load(r0);
    }
    // MODE 1: b at this thread's output nodes, loaded ahead of the next
    // right-hand side's tile so that waiting for it leaves that tile in flight
    C bv[RPT];
    auto load_b = [&](int r) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int gi = z0 + ty + 4 * k, gj = x0 + tx;
            bv[k] = b[(int64_t)r * n + (int64_t)(gi < H ? gi : 0) * W + (gj < W ? gj : 0)];
        }
    };
    for (int r = r0; r < r1; ++r) {
        lds_barrier();  // the previous right-hand side's readers are done
        if constexpr (PF) {
#pragma unroll
            for (int k = 0; k < NR; ++k)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    if (staged(k, c)) su[(ty + 4 * k) * SU + tx + 64 * c] = pre[k][c];
            lds_barrier();
            if (MODE == 1) load_b(r);
            if (r + 1 < r1) load(r + 1);
        } else {
            // the whole tile is loaded before any of it is stored (addresses
            // clamped for the nodes outside the staged region)
            const C* ur = u + (int64_t)r * n;
            C tile[NR][2];
#pragma unroll
            for (int k = 0; k < NR; ++k)
#pragma unroll
                for (int c = 0; c < 2; ++c) tile[k][c] = ur[staged(k, c) ? rowoff[k] + colg[c] : 0];
            if (MODE == 1) load_b(r);
            // elements outside the staged region go to slot 0, a corner of the
            // tile that is never read: a store behind a branch would let the
            // compiler sink its load to it and wait for the loads one by one
#pragma unroll
            for (int k = 0; k < NR; ++k)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int keep = -(int)staged(k, c);  // all ones or zero: a mask, not a branch
                    su[((ty + 4 * k) * SU + tx + 64 * c) & keep] = tile[k][c];
                }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < FXT; ++q) {
            const int e = tid + q * 256;
            if (e < TZ * NFX) {
                const int rw = e / NFX, f = e - rw * NFX - 4;
                const C* row = su + (rw + HL) * SU + HL;
                C d = dstag<C, R>(c0, c1, c2, c3, row[f + 1], row[f], row[f + 2], row[f - 1], row[f + 3], row[f - 2],
                                  row[f + 4], row[f - 3]);
                d.x *= cfx[q];
                d.y *= cfx[q];
                sgx[e] = d;
            }
        }
#pragma unroll
        for (int q = 0; q < FZT; ++q) {
            const int e = tid + q * 256;
            if (e < NFZ * TX) {
                const int fr = e / TX - 4, c = e % TX;
                const C* col = su + (fr + HL) * SU + HL + c;  // col[k * SU] = u at tile row fr + k
                C d = dstag<C, R>(c0, c1, c2, c3, col[SU], col[0], col[2 * SU], col[-SU], col[3 * SU], col[-2 * SU],
                                  col[4 * SU], col[-3 * SU]);
                d.x *= cfz[q];
                d.y *= cfz[q];
                sgz[e] = d;
            }
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = ty + 4 * k, c = tx;
            const int gi = z0 + i, gj = x0 + c;
            if (gi < H && gj < W) {
                const C* gxr = sgx + i * NFX + 4 + c;   // gxr[m] = face c + m
                const C* gzc = sgz + (i + 4) * TX + c;  // gzc[m * TX] = face row i + m
                C acc = dstag<C, R>(c0, c1, c2, c3, gxr[0], gxr[-1], gxr[1], gxr[-2], gxr[2], gxr[-3], gxr[3],
                                    gxr[-4]);
                const C dz = dstag<C, R>(c0, c1, c2, c3, gzc[0], gzc[-TX], gzc[TX], gzc[-2 * TX], gzc[2 * TX],
                                         gzc[-3 * TX], gzc[3 * TX], gzc[-4 * TX]);
                acc.x += dz.x;
                acc.y += dz.y;
                cfma(acc, kc[k], su[(i + HL) * SU + HL + c]);
                const int64_t o = (int64_t)r * n + (int64_t)gi * W + gj;
                if (MODE == 1) acc = csub(bv[k], acc);
                if (MODE == 2) {
                    acc.x = -acc.x;
                    acc.y = -acc.y;
                }
                out[o] = acc;
            }
        }
    }
}

template <int TZ>
static int stag8_rhs_per_group(const Level& L, int nrhs) {
    const int64_t tiles = (int64_t)((L.nx + S8_TX - 1) / S8_TX) * ((L.nz + TZ - 1) / TZ);
    const int64_t groups = std::max<int64_t>(1, std::min<int64_t>(nrhs, (2048 + tiles - 1) / tiles));
    return (int)((nrhs + groups - 1) / groups);
}

int32_t launch_stag8_32(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                        const float2* b, hipStream_t s) {
    constexpr int TZ = 16;
    HELM_CHECK(L.ax32 && L.az32, HELM_ERR_STATE, "8-tap operator: coefficients not set");
    // the streaming kernel with unit split-PML factors is this operator: fluxes
    // ax D+x u, az D+z u, their differences, plus k0^2 kappa_c u (conjugated
    // mass in the adjoint)
    if (!stag8_tiled() && L.cmass32 && L.ctab32) return launch_cond8_32(L, u, out, nrhs, adjoint, b, s);
    const int rb = stag8_rhs_per_group<TZ>(L, nrhs);
    const dim3 block(64, 4), grid((L.nx + S8_TX - 1) / S8_TX, (L.nz + TZ - 1) / TZ, (nrhs + rb - 1) / rb);
    const float ih = (float)(1.0 / L.h_m);
    const float c0 = (float)FD0 * ih, c1 = (float)FD1 * ih, c2 = (float)FD2 * ih, c3 = (float)FD3 * ih;
    if (b)
        hipLaunchKernelGGL((k_stag8<float, float2, TZ, 1>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.ax32,
                           L.az32, L.kap32, (float)k0sq, c0, c1, c2, c3, L.sp_cells, (float)L.sp_amp, (int)adjoint,
                           u, out, b);
    else
        hipLaunchKernelGGL((k_stag8<float, float2, TZ, 0>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.ax32,
                           L.az32, L.kap32, (float)k0sq, c0, c1, c2, c3, L.sp_cells, (float)L.sp_amp, (int)adjoint,
                           u, out, (const float2*)nullptr);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

int32_t launch_stag8_64(const Level& L, double k0sq, const double2* u, double2* out, int nrhs, bool adjoint,
                        const double2* b, int mode, hipStream_t s) {
    constexpr int TZ = 8;
    HELM_CHECK(L.ax64 && L.az64, HELM_ERR_STATE, "8-tap operator: coefficients not set");
    const int rb = stag8_rhs_per_group<TZ>(L, nrhs);
    const dim3 block(64, 4), grid((L.nx + S8_TX - 1) / S8_TX, (L.nz + TZ - 1) / TZ, (nrhs + rb - 1) / rb);
    const double ih = 1.0 / L.h_m;
    const double c0 = FD0 * ih, c1 = FD1 * ih, c2 = FD2 * ih, c3 = FD3 * ih;
#define HELM_S8_LAUNCH(M, BP)                                                                                    \
    hipLaunchKernelGGL((k_stag8<double, double2, TZ, M>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.ax64, \
                       L.az64, L.kap64, k0sq, c0, c1, c2, c3, L.sp_cells, L.sp_amp, (int)adjoint, u, out, BP)
    if (mode == 1) {
        HELM_CHECK(b, HELM_ERR_ARG, "8-tap residual needs b");
        HELM_S8_LAUNCH(1, b);
    } else if (mode == 2) {
        HELM_S8_LAUNCH(2, (const double2*)nullptr);
    } else {
        HELM_S8_LAUNCH(0, (const double2*)nullptr);
    }
#undef HELM_S8_LAUNCH
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

// ---- medium (wave_operator.py:46-70) -------------------------------------
struct SpGeom {
    int H, W, h, w, pz, px;
    double dx, z0, rho0, c0, omega, alpha_np, alpha_scale, sp_amp, relax_dt;
    int sp_cells, relaxation;
    double epa0, epa1, tau0, tau1;
};

__device__ __forceinline__ double sp_a(const float* __restrict__ rho, const SpGeom& g, int bi, int bj) {
    bi = wrap_index(bi, g.H);
    bj = wrap_index(bj, g.W);
    const int oi = min(max(bi - g.pz, 0), g.h - 1), oj = min(max(bj - g.px, 0), g.w - 1);
    return g.rho0 / (double)rho[(int64_t)oi * g.w + oj];
}

__device__ __forceinline__ double harmonic(double a, double b) { return 2.0 * a * b / (a + b); }

__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

__global__ void k_sp_medium(SpGeom g, const float* __restrict__ c, const float* __restrict__ rho, double* ax64,
                            double* az64, float* ax32, float* az32, double2* kap64, float2* kap32, double* eta64,
                            float* eta32) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= g.W || i >= g.H) return;
    const int oi = min(max(i - g.pz, 0), g.h - 1), oj = min(max(j - g.px, 0), g.w - 1);
    const double a = sp_a(rho, g, i, j);
    const double cp = (double)c[(int64_t)oi * g.w + oj];
    const double bb = a * (g.c0 / cp) * (g.c0 / cp);
    const double axe = harmonic(a, sp_a(rho, g, i, j + 1)), axw = harmonic(a, sp_a(rho, g, i, j - 1));
    const double azs = harmonic(a, sp_a(rho, g, i + 1, j)), azn = harmonic(a, sp_a(rho, g, i - 1, j));
    const double z = (double)(i - g.pz) * g.dx + g.z0;
    double tissue = 0.0;
    if (z >= -1e-6 * g.dx) {  // z >= 0 up to rounding (the row at z = 0 is tissue)
        const double lc = log(cp / g.c0) / 0.03, la = log(a) / 0.05;
        tissue = 1.0 - exp(-lc * lc - la * la);
    }
    const double sp = sponge_at<double>(i, j, g.H, g.W, g.sp_cells, g.sp_amp);
    double2 ks;
    if (g.relaxation) {
        const double e0 = g.alpha_scale * tissue * cp / g.c0 * g.epa0;
        const double e1 = g.alpha_scale * tissue * cp / g.c0 * g.epa1;
        const double s1 = 1.0 + 0.5 * (e0 + e1);
        const double amp = s1 * s1;
        double sn, cs;
        sincos(g.omega * g.relax_dt, &sn, &cs);
        const double2 step = make_double2(cs, sn);
        const double2 sm1 = make_double2(cs - 1.0, sn);
        // h = sum_k e_k / (amp (1 + tau_k (step - 1)/dt))
        double2 hs = make_double2(0.0, 0.0);
        const double ek[2] = {e0, e1}, tk[2] = {g.tau0, g.tau1};
        for (int q = 0; q < 2; ++q) {
            const double2 den = make_double2(amp * (1.0 + tk[q] * sm1.x / g.relax_dt), amp * tk[q] * sm1.y / g.relax_dt);
            hs = cadd(hs, cdiv(make_double2(ek[q], 0.0), den));
        }
        // modulus = amp (1 - h) / (1 - (1 - 1/step) h)
        const double2 inv_step = cdiv(make_double2(1.0, 0.0), step);
        const double2 one_m = make_double2(1.0 - inv_step.x, -inv_step.y);
        const double2 num = make_double2(amp * (1.0 - hs.x), -amp * hs.y);
        const double2 den = csub(make_double2(1.0, 0.0), cmul(one_m, hs));
        const double2 modulus = cdiv(num, den);
        ks = cdiv(make_double2(bb, 0.0), modulus);
        ks.y -= sp;
    } else {
        const double t = g.alpha_np * tissue * cp / g.omega;
        ks = make_double2(bb * (1.0 - t * t), -2.0 * bb * t - sp);  // b (1 - i t)^2 - i sponge
    }
    const int64_t o = (int64_t)i * g.W + j;
    ax64[o] = axe;
    az64[o] = azs;
    ax32[o] = (float)axe;
    az32[o] = (float)azs;
    const double2 kt = cconj(ks);  // stored in this library's convention
    kap64[o] = kt;
    kap32[o] = make_float2((float)kt.x, (float)kt.y);
    const double e = 0.25 * (axe + axw + azs + azn);
    eta64[o] = e;
    eta32[o] = (float)e;
}

// ---- incident field (wave_operator.py:111-127) -----------------------------
// sp_buf layout: [0, W) twiddles e^{2 pi i m / W}; [W, 2W) boundary trace;
// [2W, 3W) k_z; [3W, 4W) boundary spectrum.
__device__ __forceinline__ double stag_sum(double q) {
    return FD0 * sin(0.5 * q) + FD1 * sin(1.5 * q) + FD2 * sin(2.5 * q) + FD3 * sin(3.5 * q);
}
__device__ __forceinline__ double stag_dsum(double q) {
    return FD0 * 0.5 * cos(0.5 * q) + FD1 * 1.5 * cos(1.5 * q) + FD2 * 2.5 * cos(2.5 * q) + FD3 * 3.5 * cos(3.5 * q);
}

__global__ void k_inc_prepare(int W, int px, double dx, double x0, double k0, double sin_theta, double aperture,
                              double2* buf) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= W) return;
    double sn, cs;
    sincospi(2.0 * m / W, &sn, &cs);
    buf[m] = make_double2(cs, sn);
    const double x = (double)(m - px) * dx + x0;
    double ap = 0.0;
    if (fabs(x) <= aperture) ap = 0.5 * (1.0 + cospi(x / aperture));
    sincos(-k0 * x * sin_theta, &sn, &cs);
    buf[W + m] = make_double2(ap * cs, ap * sn);
    // discrete dispersion: solve 2 S(q) = sqrt(k0^2 - sx) h for the propagating branch
    const int kk = m < (W + 1) / 2 ? m : m - W;  // fftfreq order
    const double qx = 2.0 * M_PI * (double)kk / W;
    const double dxs = 2.0 * stag_sum(qx) / dx;
    const double sx = dxs * dxs;
    const double desired = sqrt(fmax(k0 * k0 - sx, 0.0)) * dx;
    double q = fmin(fmax(desired, 0.0), M_PI);
    for (int it = 0; it < 6; ++it) {
        const double value = 2.0 * stag_sum(q), deriv = 2.0 * stag_dsum(q);
        q = fmin(fmax(q - (value - desired) / fmax(deriv, 0.05), 0.0), M_PI);
    }
    buf[2 * W + m] = make_double2(q / dx, -sqrt(fmax(sx - k0 * k0, 0.0)));
}

// spectrum[k] = sum_j trace[j] e^{-2 pi i j k / W}
__global__ void k_inc_spectrum(int W, double2* buf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= W) return;
    const double2* tw = buf;
    const double2* tr = buf + W;
    double2 acc = make_double2(0.0, 0.0);
    int idx = 0;
    for (int j = 0; j < W; ++j) {
        cfma(acc, tr[j], cconj(tw[idx]));
        idx += k;
        if (idx >= W) idx -= W;
    }
    buf[3 * W + k] = acc;
}

// field[i][x] = (1/W) sum_k spectrum[k] e^{-i |z_i| kz_k} e^{2 pi i k x / W}; one workgroup per row
constexpr int INC_CHUNK = 1024;
constexpr int INC_MAXX = 16;  // x values per thread: W <= 4096
__global__ __launch_bounds__(256) void k_inc_field(int H, int W, int pz, double dx, double z0,
                                                   const double2* __restrict__ buf, double2* __restrict__ field) {
    __shared__ double2 row[INC_CHUNK];
    extern __shared__ double2 tw[];  // the W twiddles, gathered by every output from LDS
    const int i = blockIdx.x;
    const double az = fabs((double)(i - pz) * dx + z0);
    for (int k = threadIdx.x; k < W; k += blockDim.x) tw[k] = buf[k];
    const double2* kz = buf + 2 * W;
    const double2* spec = buf + 3 * W;
    double2 acc[INC_MAXX];
#pragma unroll
    for (int q = 0; q < INC_MAXX; ++q) acc[q] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < W; k0 += INC_CHUNK) {
        const int kn = min(INC_CHUNK, W - k0);
        __syncthreads();
        for (int k = threadIdx.x; k < kn; k += blockDim.x) {
            const double2 kk = kz[k0 + k];
            // e^{-i |z| kz} = e^{|z| Im kz} e^{-i |z| Re kz}
            const double mag = exp(az * kk.y);
            double sn, cs;
            sincos(-az * kk.x, &sn, &cs);
            row[k] = cmul(spec[k0 + k], make_double2(mag * cs, mag * sn));
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < INC_MAXX; ++q) {
            const int x = threadIdx.x + q * 256;
            if (x < W) {
                int idx = (int)(((int64_t)k0 * x) % W);
                double2 a = acc[q];
                for (int k = 0; k < kn; ++k) {
                    cfma(a, row[k], tw[idx]);
                    idx += x;
                    if (idx >= W) idx -= W;
                }
                acc[q] = a;
            }
        }
    }
    const double s = 1.0 / W;
#pragma unroll
    for (int q = 0; q < INC_MAXX; ++q) {
        const int x = threadIdx.x + q * 256;
        if (x < W) field[(int64_t)i * W + x] = cscale(acc[q], s);
    }
}

// Scratch of helm_spectral_incident_noalloc: the incident-field work buffer
// (twiddles, boundary trace, k_z, boundary spectrum: 4 W complex doubles) for
// boxes up to 256 * INC_MAXX = 4096 nodes wide, 256 KiB per device, held
// until the process ends.  Host code must not pass `scratch` itself to a
// kernel (on the host it names a stand-in, not device memory): the device
// address of each GPU's copy comes from hipGetSymbolAddress, once per device
// with that device current, and is kept in scratch_ptr[device].
constexpr int INC_SCRATCH = 4 * 256 * INC_MAXX;
static __device__ __align__(256) double2 scratch[INC_SCRATCH];
constexpr int MAX_DEVICES = 64;
static double2* scratch_ptr[MAX_DEVICES] = {};
static std::mutex scratch_mutex;

static int32_t scratch_address(int device, double2** out) {
    HELM_CHECK(device >= 0 && device < MAX_DEVICES, HELM_ERR_ARG, "no-alloc scratch: device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device]) {
        int current = 0;
        HELM_HIP(hipGetDevice(&current));
        if (current != device) HELM_HIP(hipSetDevice(device));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (current != device) HELM_HIP(hipSetDevice(current));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("hipGetSymbolAddress (incident scratch): ") + hipGetErrorString(e));
            return HELM_ERR_HIP;
        }
        scratch_ptr[device] = (double2*)ptr;
    }
    *out = scratch_ptr[device];
    return HELM_OK;
}

template <typename T>
static int32_t sp_alloc(T** ptr, size_t count) {
    if (*ptr) return HELM_OK;
    hipError_t e = hipMalloc((void**)ptr, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        *ptr = nullptr;
        set_error(std::string("hipMalloc (spectral operator): ") + hipGetErrorString(e));
        return HELM_ERR_ALLOC;
    }
    return HELM_OK;
}

// the three incident-field kernels on the work buffer buf (4 W complex doubles)
static int32_t launch_incident(helm_plan* p, double sin_theta, double aperture_m, double2* buf, double2* field,
                               hipStream_t s) {
    const Level& L = p->levels[0];
    const int W = L.nx, H = L.nz;
    const helm_spectral_medium& m = p->spectral;
    hipLaunchKernelGGL(k_inc_prepare, dim3((W + 255) / 256), dim3(256), 0, s, W, m.px, L.h_m, m.x0_m, p->k0,
                       sin_theta, aperture_m, buf);
    hipLaunchKernelGGL(k_inc_spectrum, dim3((W + 255) / 256), dim3(256), 0, s, W, buf);
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)k_inc_field, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
    (void)attr;
    hipLaunchKernelGGL(k_inc_field, dim3(H), dim3(256), (size_t)W * sizeof(double2), s, H, W, m.pz, L.h_m, m.z0_m,
                       buf, field);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

}  // namespace helm

namespace helm {
// The S operator's FP32 applies through the conditioned streaming kernel:
// mass k0^2 kappa_c (rounded as k_stag8 forms it: float(k0^2) * kappa_c in
// FP32) and its conjugate (the adjoint's mass; the faces are real)
__global__ void k_sp_mass(int64_t n, double k0sq, const float2* __restrict__ kap32, float2* __restrict__ m32,
                          float2* __restrict__ mh32) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float2 k32 = kap32[e];
    const float s = (float)k0sq;
    m32[e] = make_float2(k32.x * s, k32.y * s);
    mh32[e] = make_float2(k32.x * s, -k32.y * s);
}
}  // namespace helm

using namespace helm;

extern "C" {

int32_t helm_set_medium_spectral(helm_plan* p, const float* c, const float* rho, const helm_spectral_medium* m,
                                 void* stream) {
    HELM_CHECK(p && c && rho && m, HELM_ERR_ARG, "helm_set_medium_spectral: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_STAG8, HELM_ERR_STATE,
               "helm_set_medium_spectral: the plan was not created with op_kind HELM_OP_STAG8");
    Level& L = p->levels[0];
    HELM_CHECK(m->h >= 1 && m->w >= 1 && m->pz >= 0 && m->px >= 0 && m->pz + m->h <= L.nz && m->px + m->w <= L.nx,
               HELM_ERR_ARG, "helm_set_medium_spectral: the medium does not fit the box");
    HELM_CHECK(std::isfinite(m->rho0) && m->rho0 > 0, HELM_ERR_ARG, "rho0 must be finite and positive");
    HELM_CHECK(std::isfinite(m->alpha_scale) && m->alpha_scale >= 0, HELM_ERR_ARG,
               "alpha_scale must be finite and nonnegative");
    HELM_CHECK(m->sponge_cells >= 0 && std::isfinite(m->sponge_amplitude) && m->sponge_amplitude >= 0, HELM_ERR_ARG,
               "invalid sponge");
    HELM_CHECK(!m->relaxation || (std::isfinite(m->relax_dt_s) && m->relax_dt_s > 0), HELM_ERR_ARG,
               "relaxation needs relax_dt_s > 0");
    hipStream_t s = (hipStream_t)stream;
    HELM_TRY(sp_alloc(&L.ax64, L.n));
    HELM_TRY(sp_alloc(&L.az64, L.n));
    HELM_TRY(sp_alloc(&L.ax32, L.n));
    HELM_TRY(sp_alloc(&L.az32, L.n));
    SpGeom g{};
    g.H = L.nz;
    g.W = L.nx;
    g.h = m->h;
    g.w = m->w;
    g.pz = m->pz;
    g.px = m->px;
    g.dx = L.h_m;
    g.z0 = m->z0_m;
    g.rho0 = m->rho0;
    g.c0 = p->desc.c0_m_s;
    g.omega = p->omega;
    g.alpha_scale = m->alpha_scale;
    g.alpha_np = m->alpha_scale * 100.0 * std::log(10.0) / 20.0 * std::pow(p->desc.frequency_mhz, 1.1);
    g.sp_cells = m->sponge_cells;
    g.sp_amp = m->sponge_amplitude;
    g.relaxation = m->relaxation;
    g.relax_dt = m->relax_dt_s;
    g.epa0 = m->eta_per_unit_alpha[0];
    g.epa1 = m->eta_per_unit_alpha[1];
    g.tau0 = m->taus_s[0];
    g.tau1 = m->taus_s[1];
    hipLaunchKernelGGL(k_sp_medium, dim3((L.nx + 255) / 256, L.nz), dim3(256), 0, s, g, c, rho, L.ax64, L.az64,
                       L.ax32, L.az32, L.kap64, L.kap32, L.eta64, L.eta32);
    HELM_HIP(hipGetLastError());
    L.op = HELM_OP_STAG8;
    L.sp_cells = m->sponge_cells;
    L.sp_amp = m->sponge_amplitude;
    // k0 of wave_operator.py:55 (the relaxation form matches the 2.5 ns time stepper)
    p->k0 = m->relaxation ? 2.0 * std::sin(p->omega * m->relax_dt_s / 2.0) / m->relax_dt_s / p->desc.c0_m_s
                          : p->omega / p->desc.c0_m_s;
    // the streaming kernel's FP32 operands: mass k0^2 kappa_c, unit per-column and
    // per-row factors (sv = 1, g - 1 = 0, cA = 1, cB = 0: no PML), the whole
    // box as the native box (periodic rows and columns, as here)
    HELM_TRY(sp_alloc(&L.cmass32, L.n));
    HELM_TRY(sp_alloc(&L.cmassh32, L.n));
    HELM_TRY(sp_alloc(&L.ctab32, (size_t)4 * (L.nx + L.nz)));
    hipLaunchKernelGGL(k_sp_mass, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, L.n, p->k0 * p->k0, L.kap32,
                       L.cmass32, L.cmassh32);
    HELM_HIP(hipGetLastError());
    {
        std::vector<float2> t32((size_t)4 * (L.nx + L.nz));
        auto fill = [&](int n, size_t base) {
            for (int k = 0; k < n; ++k)
                for (int t = 0; t < 4; ++t) t32[base + (size_t)t * n + k] = make_float2((t == 0 || t == 2) ? 1.f : 0.f, 0.f);
        };
        fill(L.nx, 0);
        fill(L.nz, (size_t)4 * L.nx);
        HELM_TRY(upload_h2d(L.ctab32, t32.data(), t32.size() * sizeof(float2)));
    }
    L.cbox[0] = 0;
    L.cbox[1] = 0;
    L.cbox[2] = L.nz;
    L.cbox[3] = L.nx;
    p->spectral = *m;
    p->medium_set = true;
    if (p->mg.levels > 0) HELM_TRY(mg_build(p, s));
    return HELM_OK;
}

int32_t helm_spectral_incident(helm_plan* p, double sin_theta, double aperture_m, void* field, void* stream) {
    HELM_CHECK(p && field, HELM_ERR_ARG, "helm_spectral_incident: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_STAG8 && p->medium_set, HELM_ERR_STATE,
               "helm_spectral_incident: call helm_set_medium_spectral first");
    HELM_CHECK(std::isfinite(sin_theta) && std::isfinite(aperture_m) && aperture_m > 0, HELM_ERR_ARG,
               "helm_spectral_incident: invalid angle or aperture");
    const Level& L = p->levels[0];
    const int W = L.nx, H = L.nz;
    HELM_CHECK(W <= 256 * INC_MAXX, HELM_ERR_ARG, "helm_spectral_incident: box wider than 4096 nodes");
    hipStream_t s = (hipStream_t)stream;
    HELM_TRY(sp_alloc(&p->sp_buf, (size_t)4 * W));
    return launch_incident(p, sin_theta, aperture_m, p->sp_buf, (double2*)field, s);
}

int32_t helm_spectral_incident_noalloc(helm_plan* p, double sin_theta, double aperture_m, void* field,
                                       void* stream) {
    HELM_CHECK(p && field, HELM_ERR_ARG, "helm_spectral_incident_noalloc: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_STAG8 && p->medium_set, HELM_ERR_STATE,
               "helm_spectral_incident_noalloc: call helm_set_medium_spectral first");
    HELM_CHECK(std::isfinite(sin_theta) && std::isfinite(aperture_m) && aperture_m > 0, HELM_ERR_ARG,
               "helm_spectral_incident_noalloc: invalid angle or aperture");
    const int W = p->levels[0].nx;
    HELM_CHECK(W >= 1 && 4 * W <= INC_SCRATCH, HELM_ERR_ARG,
               "helm_spectral_incident_noalloc: box wider than 4096 nodes (the static scratch holds 4 x 4096 "
               "complex doubles)");
    double2* buf = nullptr;
    HELM_TRY(scratch_address(p->device, &buf));
    return launch_incident(p, sin_theta, aperture_m, buf, (double2*)field, (hipStream_t)stream);
}

int32_t helm_spectral_rhs(helm_plan* p, const void* inc, void* rhs, int32_t nrhs, void* stream) {
    HELM_CHECK(p && inc && rhs, HELM_ERR_ARG, "helm_spectral_rhs: null argument");
    HELM_CHECK(inc != rhs, HELM_ERR_ARG, "helm_spectral_rhs: inc and rhs must not alias");
    HELM_CHECK(p->desc.op_kind == HELM_OP_STAG8 && p->medium_set, HELM_ERR_STATE,
               "helm_spectral_rhs: call helm_set_medium_spectral first");
    HELM_CHECK(nrhs >= 1 && (int64_t)nrhs * p->levels[0].n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_spectral_rhs: nrhs out of range");
    return launch_stag8_64(p->levels[0], p->k0 * p->k0, (const double2*)inc, (double2*)rhs, nrhs, true, nullptr, 2,
                           (hipStream_t)stream);
}

}  // extern "C"
