// Conditioned split-PML operator (op_kind HELM_OP_COND8): the masked,
// projected frequency operator of the dense path, its analytic adjoint, its
// medium set-up and its source term.
//
// Operator (nsrno/conditioned_wave.py:52-80 applied through
// nsrno/dense_wave.py:66-68; the reference applies it with eleven FFTs; here
// it is the equivalent real-space stencil, whose Fourier symbol is the
// reference's D = 2i sum_m a_m sin((m+1/2) q) e^{iq/2} / h, on a box whose
// padding (>= 16 nodes) keeps the periodic wrap out of the native domain):
//   p_m = M p (M = 1 inside the native medium, 0 outside)
//   Fx  = ax sv_x (D+x p_m)                          (faces j+1/2)
//   (A p) = M [ cA_x D-x(Fx) + cB_x D-x((g_x - 1) Fx) + (same in z) + mass p_m ]
//         + (1 - M) p
// with, per column j (rows likewise), from the split-PML factors of the
// reference (D = e^{i omega dt}, scalar = (D-1)/(D k0^2)):
//   sv = (D-1)/(D-av), g = 1 + (1-av)/D, left = m/(D-ar), right = 1 + (1-m)/(D-ar),
//   cA = scalar (left + right), cB = scalar right        (m = [ar != 1])
// which is scalar [left D-(F) + right D-(g F)] rearranged.  The adjoint
// (dense_wave.py:70-82):
//   (A^H y) = M [ D-x( conj(ax sv_x) [D+x(conj(cA_x) y_m) + conj(g_x - 1) D+x(conj(cB_x) y_m)] )
//               + (same in z) + conj(mass) y_m ] + (1 - M) y.
// Away from the PML g = 1 exactly, so tiles whose faces all have g = 1 skip
// the second divergence, and tiles whose columns (rows) share one cA skip the
// per-node scaling of the adjoint's differences.
//
// Kernel layout as the spectral 8-tap kernel: a workgroup owns a TZ x 64
// output tile, stages the masked field with a 7-node halo in LDS, forms the
// face fluxes in LDS, then the divergences; coefficients are loaded once per
// workgroup and reused for its group of right-hand sides.  Algorithmic bytes
// per node per RHS: 8 (read) + 8 (write) for FP32, plus (4 + 4 + 8)/R for
// ax, az, mass.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <type_traits>
#include <utility>

#include "helm_internal.h"

namespace helm {

constexpr double CD0 = 1225.0 / 1024.0, CD1 = -245.0 / 3072.0, CD2 = 49.0 / 5120.0, CD3 = -5.0 / 7168.0;
constexpr int C8_TX = 64;
constexpr int C8_HL = 7;

__device__ __forceinline__ int cwrap(int v, int m) {
    if (v < 0 || v >= m) {
        v %= m;
        if (v < 0) v += m;
    }
    return v;
}

template <typename C, typename R>
__device__ __forceinline__ C c8diff(R c0, R c1, R c2, R c3, C a0, C b0, C a1, C b1, C a2, C b2, C a3, C b3) {
    C d;
    d.x = c0 * (a0.x - b0.x) + c1 * (a1.x - b1.x) + c2 * (a2.x - b2.x) + c3 * (a3.x - b3.x);
    d.y = c0 * (a0.y - b0.y) + c1 * (a1.y - b1.y) + c2 * (a2.y - b2.y) + c3 * (a3.y - b3.y);
    return d;
}

template <typename C>
__device__ __forceinline__ C cm(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
// conj(a) * b
template <typename C>
__device__ __forceinline__ C cjm(C a, C b) {
    C r;
    r.x = a.x * b.x + a.y * b.y;
    r.y = a.x * b.y - a.y * b.x;
    return r;
}
template <typename C>
__device__ __forceinline__ bool ceq(C a, C b) {
    return a.x == b.x && a.y == b.y;
}

// Table layout (per axis, length n = W for columns, H for rows):
// t[0 n + k] = sv, t[1 n + k] = g - 1, t[2 n + k] = cA, t[3 n + k] = cB.
// MODE 0: out = A u (or A^H u with ADJ); 1: out = b - A u.
template <typename R, typename C, int TZ, int MODE, bool ADJ>
__global__ __launch_bounds__(256) void k_cond8(int H, int W, int64_t n, int nrhs, int rb, const R* __restrict__ ax,
                                               const R* __restrict__ az, const C* __restrict__ mass,
                                               const C* __restrict__ tx, const C* __restrict__ tz, int bz0, int bx0,
                                               int bh, int bw, R c0, R c1, R c2, R c3, const C* __restrict__ u,
                                               C* __restrict__ out, const C* __restrict__ b) {
    constexpr int TX = C8_TX, HL = C8_HL;
    constexpr int SU = TX + 2 * HL;
    constexpr int NFX = TX + 7;  // x faces -4 .. TX+2 of each tile row
    constexpr int NFZ = TZ + 7;  // z face rows -4 .. TZ+2
    constexpr int NZT = TZ + 2 * HL;
    constexpr int RPT = TZ / 4;
    constexpr int FXT = (TZ * NFX + 255) / 256;
    constexpr int FZT = (NFZ * TX + 255) / 256;
    __shared__ C su[NZT * SU];
    __shared__ C sgx[TZ * NFX];
    __shared__ C sgz[NFZ * TX];
    __shared__ C xsv[SU], xg[SU], xa[SU], xb[SU];      // columns x0-7 .. x0+TX+6
    __shared__ C zsv[NZT], zg[NZT], za[NZT], zb[NZT];  // rows z0-7 .. z0+TZ+6

    const int tx_ = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx_;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;

    for (int k = tid; k < SU; k += 256) {
        const int j = cwrap(x0 + k - HL, W);
        xsv[k] = tx[j];
        xg[k] = tx[W + j];
        xa[k] = tx[2 * W + j];
        xb[k] = tx[3 * W + j];
    }
    for (int k = tid; k < NZT; k += 256) {
        const int i = cwrap(z0 + k - HL, H);
        zsv[k] = tz[i];
        zg[k] = tz[H + i];
        za[k] = tz[2 * H + i];
        zb[k] = tz[3 * H + i];
    }
    __syncthreads();
    // second divergence needed: some face of the tile window has g != 1
    bool px_ = false, pz_ = false, nux = false, nuz = false;
    if (tid < NFX) px_ = xg[tid - 4 + HL].x != R(0) || xg[tid - 4 + HL].y != R(0);
    if (tid < NFZ) pz_ = zg[tid - 4 + HL].x != R(0) || zg[tid - 4 + HL].y != R(0);
    if (ADJ) {  // node scalings of the window not all equal
        if (tid < SU) nux = !ceq(xa[tid], xa[HL]);
        if (tid < NZT) nuz = !ceq(za[tid], za[HL]);
    }
    const bool pmlx = __syncthreads_or(px_);
    const bool pmlz = __syncthreads_or(pz_);
    const bool varx = ADJ ? (__syncthreads_or(nux) || pmlx) : false;
    const bool varz = ADJ ? (__syncthreads_or(nuz) || pmlz) : false;

    // face coefficients ax sv (x faces of the tile rows) and az sv (z faces of the tile columns)
    C cfx[FXT], cfz[FZT];
#pragma unroll
    for (int q = 0; q < FXT; ++q) {
        const int e = tid + q * 256;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (e < TZ * NFX) {
            const int r = e / NFX, f = e - r * NFX - 4;
            const R a = ax[(int64_t)cwrap(z0 + r, H) * W + cwrap(x0 + f, W)];
            v = xsv[f + HL];
            v.x *= a;
            v.y *= a;
        }
        cfx[q] = v;
    }
#pragma unroll
    for (int q = 0; q < FZT; ++q) {
        const int e = tid + q * 256;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (e < NFZ * TX) {
            const int fr = e / TX - 4, c = e % TX;
            const R a = az[(int64_t)cwrap(z0 + fr, H) * W + cwrap(x0 + c, W)];
            v = zsv[fr + HL];
            v.x *= a;
            v.y *= a;
        }
        cfz[q] = v;
    }
    C mk[RPT];
    bool inside[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = z0 + ty + 4 * k, j = x0 + tx_;
        inside[k] = i < H && j < W && i >= bz0 && i < bz0 + bh && j >= bx0 && j < bx0 + bw;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (inside[k]) v = mass[(int64_t)i * W + j];
        if (ADJ) v.y = -v.y;
        mk[k] = v;
    }

    const int r1 = min(nrhs, (int)(blockIdx.z + 1) * rb);
    for (int r = blockIdx.z * rb; r < r1; ++r) {
        const C* ur = u + (int64_t)r * n;
        __syncthreads();
        for (int rr = ty; rr < NZT; rr += 4) {
            const bool zh = rr < HL || rr >= TZ + HL;
            const int gz = cwrap(z0 + rr - HL, H);
            const bool zin = gz >= bz0 && gz < bz0 + bh;
            for (int cc = tx_; cc < SU; cc += 64) {
                const bool xh = cc < HL || cc >= TX + HL;
                if (zh && xh) continue;
                const int gx = cwrap(x0 + cc - HL, W);
                C v;
                v.x = R(0);
                v.y = R(0);
                if (zin && gx >= bx0 && gx < bx0 + bw) v = ur[(int64_t)gz * W + gx];
                su[rr * SU + cc] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < FXT; ++q) {
            const int e = tid + q * 256;
            if (e < TZ * NFX) {
                const int rw = e / NFX, f = e - rw * NFX - 4;
                const C* row = su + (rw + HL) * SU + HL;
                C flux;
                if (!ADJ) {
                    flux = cm(cfx[q], c8diff<C, R>(c0, c1, c2, c3, row[f + 1], row[f], row[f + 2], row[f - 1],
                                                   row[f + 3], row[f - 2], row[f + 4], row[f - 3]));
                } else if (!varx) {
                    const C d = c8diff<C, R>(c0, c1, c2, c3, row[f + 1], row[f], row[f + 2], row[f - 1], row[f + 3],
                                             row[f - 2], row[f + 4], row[f - 3]);
                    flux = cjm(cfx[q], cjm(xa[HL], d));
                } else {
                    const C* sa = xa + HL;
                    const C d = c8diff<C, R>(c0, c1, c2, c3, cjm(sa[f + 1], row[f + 1]), cjm(sa[f], row[f]),
                                             cjm(sa[f + 2], row[f + 2]), cjm(sa[f - 1], row[f - 1]),
                                             cjm(sa[f + 3], row[f + 3]), cjm(sa[f - 2], row[f - 2]),
                                             cjm(sa[f + 4], row[f + 4]), cjm(sa[f - 3], row[f - 3]));
                    C t = d;
                    if (pmlx) {
                        const C* sb = xb + HL;
                        const C d2 = c8diff<C, R>(c0, c1, c2, c3, cjm(sb[f + 1], row[f + 1]), cjm(sb[f], row[f]),
                                                  cjm(sb[f + 2], row[f + 2]), cjm(sb[f - 1], row[f - 1]),
                                                  cjm(sb[f + 3], row[f + 3]), cjm(sb[f - 2], row[f - 2]),
                                                  cjm(sb[f + 4], row[f + 4]), cjm(sb[f - 3], row[f - 3]));
                        const C gm = xg[f + HL];
                        t = cadd(t, cjm(gm, d2));
                    }
                    flux = cjm(cfx[q], t);
                }
                sgx[e] = flux;
            }
        }
#pragma unroll
        for (int q = 0; q < FZT; ++q) {
            const int e = tid + q * 256;
            if (e < NFZ * TX) {
                const int fr = e / TX - 4, c = e % TX;
                const C* col = su + (fr + HL) * SU + HL + c;  // col[k * SU] = field at tile row fr + k
                C flux;
                if (!ADJ) {
                    flux = cm(cfz[q], c8diff<C, R>(c0, c1, c2, c3, col[SU], col[0], col[2 * SU], col[-SU],
                                                   col[3 * SU], col[-2 * SU], col[4 * SU], col[-3 * SU]));
                } else if (!varz) {
                    const C d = c8diff<C, R>(c0, c1, c2, c3, col[SU], col[0], col[2 * SU], col[-SU], col[3 * SU],
                                             col[-2 * SU], col[4 * SU], col[-3 * SU]);
                    flux = cjm(cfz[q], cjm(za[HL], d));
                } else {
                    const C* sa = za + HL + fr;  // sa[k] = row fr + k
                    const C d = c8diff<C, R>(c0, c1, c2, c3, cjm(sa[1], col[SU]), cjm(sa[0], col[0]),
                                             cjm(sa[2], col[2 * SU]), cjm(sa[-1], col[-SU]),
                                             cjm(sa[3], col[3 * SU]), cjm(sa[-2], col[-2 * SU]),
                                             cjm(sa[4], col[4 * SU]), cjm(sa[-3], col[-3 * SU]));
                    C t = d;
                    if (pmlz) {
                        const C* sb = zb + HL + fr;
                        const C d2 = c8diff<C, R>(c0, c1, c2, c3, cjm(sb[1], col[SU]), cjm(sb[0], col[0]),
                                                  cjm(sb[2], col[2 * SU]), cjm(sb[-1], col[-SU]),
                                                  cjm(sb[3], col[3 * SU]), cjm(sb[-2], col[-2 * SU]),
                                                  cjm(sb[4], col[4 * SU]), cjm(sb[-3], col[-3 * SU]));
                        t = cadd(t, cjm(zg[fr + HL], d2));
                    }
                    flux = cjm(cfz[q], t);
                }
                sgz[e] = flux;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = ty + 4 * k, c = tx_;
            const int gi = z0 + i, gj = x0 + c;
            if (gi < H && gj < W) {
                const int64_t o = (int64_t)r * n + (int64_t)gi * W + gj;
                C acc;
                if (inside[k]) {
                    const C* gxr = sgx + i * NFX + 4 + c;   // gxr[m] = face c + m
                    const C* gzc = sgz + (i + 4) * TX + c;  // gzc[m * TX] = face row i + m
                    const C dx = c8diff<C, R>(c0, c1, c2, c3, gxr[0], gxr[-1], gxr[1], gxr[-2], gxr[2], gxr[-3],
                                              gxr[3], gxr[-4]);
                    const C dz = c8diff<C, R>(c0, c1, c2, c3, gzc[0], gzc[-TX], gzc[TX], gzc[-2 * TX], gzc[2 * TX],
                                              gzc[-3 * TX], gzc[3 * TX], gzc[-4 * TX]);
                    if (!ADJ) {
                        acc = cm(xa[c + HL], dx);
                        cfma(acc, za[i + HL], dz);
                        if (pmlx) {
                            const C* gg = xg + HL + c;  // gg[m] = g - 1 at face c + m
                            const C d2 = c8diff<C, R>(c0, c1, c2, c3, cm(gg[0], gxr[0]), cm(gg[-1], gxr[-1]),
                                                      cm(gg[1], gxr[1]), cm(gg[-2], gxr[-2]), cm(gg[2], gxr[2]),
                                                      cm(gg[-3], gxr[-3]), cm(gg[3], gxr[3]), cm(gg[-4], gxr[-4]));
                            cfma(acc, xb[c + HL], d2);
                        }
                        if (pmlz) {
                            const C* gg = zg + HL + i;  // gg[m] = g - 1 at face row i + m
                            const C d2 = c8diff<C, R>(c0, c1, c2, c3, cm(gg[0], gzc[0]), cm(gg[-1], gzc[-TX]),
                                                      cm(gg[1], gzc[TX]), cm(gg[-2], gzc[-2 * TX]),
                                                      cm(gg[2], gzc[2 * TX]), cm(gg[-3], gzc[-3 * TX]),
                                                      cm(gg[3], gzc[3 * TX]), cm(gg[-4], gzc[-4 * TX]));
                            cfma(acc, zb[i + HL], d2);
                        }
                    } else {
                        acc = cadd(dx, dz);
                    }
                    cfma(acc, mk[k], su[(i + HL) * SU + HL + c]);
                } else {
                    acc = u[o];
                }
                if (MODE == 1) acc = csub(b[o], acc);
                out[o] = acc;
            }
        }
    }
}

// ---- streaming form ------------------------------------------------------
// One wavefront per workgroup owns a strip of 64 columns and a run of tzr
// output rows for RB right-hand sides, marching down the rows.  The z
// differences use two register rings: the field rows t .. t+7 and the z-face
// fluxes t-4 .. t+3 of output row t, each row in slot (row - i0) mod 8.  The
// row loop is unrolled eight times, so every slot is a compile-time constant
// and nothing moves between registers from row to row.  The x differences go
// through LDS once per block of NB output rows: the block's field rows (all in
// the ring at its first row) and their 7-column halos are staged together and
// its NB x 71 x-face fluxes formed in NB + 1 passes of the wavefront (NB passes
// of 64 faces, one of the NB x 7 halo faces) where one row at a time takes two.
// Field rows, z-face coefficients, mass and b are loaded PD rows ahead, a
// block's halo columns and x-face coefficients one block ahead.  Forward: the
// staged field is M u; adjoint: M y, scaled per column (x) and per row (z) by
// conj(cA), conj(cB) as the differences are formed.
// a row index at most a few periods outside [0, m) into it, with no division
// (the streaming kernel wraps several rows per step)
__device__ __forceinline__ int rwrap(int v, int m) {
    while (v < 0) v += m;
    while (v >= m) v -= m;
    return v;
}
template <int K>
using CK = std::integral_constant<int, K>;
template <typename F, int... Ks>
__device__ __forceinline__ void each_phase(std::integer_sequence<int, Ks...>, F&& f) {
    (f(CK<Ks>{}), ...);
}

template <typename R, typename C, int RB, int NB, int MODE, bool ADJ>
__global__ __launch_bounds__(64, (RB == 1 && sizeof(R) == 4) ? 4 : 1) void k_cond8s(int H, int W, int64_t n, int nrhs, int tzr,
                                               const R* __restrict__ ax, const R* __restrict__ az,
                                               const C* __restrict__ mass, const C* __restrict__ tx,
                                               const C* __restrict__ tz, int bz0, int bx0, int bh, int bw,
                                               R c0, R c1, R c2, R c3,
                                               const C* __restrict__ u, C* __restrict__ out,
                                               const C* __restrict__ b) {
    static_assert(NB == 4 || NB == 8, "blocks of 4 or 8 rows");
    constexpr int NS = 78, NF = 71;  // staged nodes x0-7 .. x0+70, x faces x0-4 .. x0+66 (+1/2) of a row
    constexpr int PD = 2;            // rows loaded ahead
    constexpr int NHE = (14 * NB + 63) / 64;  // halo nodes per lane and block
    __shared__ C st[RB][NB * NS];
    __shared__ C sbt[ADJ ? RB : 1][ADJ ? NB * NS : 1];
    __shared__ C sf[RB][NB * NF];
    __shared__ C sgx[NF];  // g - 1 of the strip's x faces (forward PML term)
    const int l = threadIdx.x;
    const int x0 = blockIdx.x * 64, j = x0 + l;
    const int i0 = blockIdx.y * tzr;
    if (i0 >= H) return;
    const int i1 = min(H, i0 + tzr);
    const int r0 = blockIdx.z * RB;
    const C zero = C{R(0), R(0)};
    const bool jok = j < W;
    const int jc = cwrap(j, W);
    const bool xin = jc >= bx0 && jc < bx0 + bw;
    auto zin = [&](int r) { return r >= bz0 && r < bz0 + bh; };
    // row `row` of RHS r0+q (clamped into range) from base pointer f; rows are
    // uniform, so a load is a scalar base plus the lane's column offset
    auto rowp = [&](const C* f, int q, int row) -> const C* {
        return f + (int64_t)(r0 + q < nrhs ? r0 + q : nrhs - 1) * n + (int64_t)row * W;
    };

    // x faces: lane l forms face l of each block row (face column x0-4+l) and,
    // in the halo pass (lanes < 7 NB), face 64 + l%7 of block row l/7
    const int fa = cwrap(x0 - 4 + l, W);
    const int hr = l < 7 * NB ? l / 7 : 0, hf = 64 + (l < 7 * NB ? l % 7 : 0);
    const int fh = cwrap(x0 - 4 + hf, W);
    const C svA = tx[fa], gA = tx[W + fa], svH = tx[fh], gH = tx[W + fh];
    const C cAj = tx[2 * W + jc], cBj = tx[3 * W + jc];
    sgx[l] = gA;
    if (l < 7) sgx[64 + l] = gH;
    const bool pmlx = __any((gA.x != R(0) || gA.y != R(0)) || (l < 7 * NB && (gH.x != R(0) || gH.y != R(0))));
    // halo nodes: entry e = l + 64 h of 14 NB (block row e/14; columns x0-7.. x0-1, x0+64 .. x0+70)
    int hrow[NHE], hcol[NHE], hpos[NHE];
    bool hok[NHE], hin[NHE];
    C hA[NHE], hB[NHE];
#pragma unroll
    for (int h = 0; h < NHE; ++h) {
        const int e = l + 64 * h, ec = e < 14 * NB ? e : 0;
        const int k = ec % 14;
        hok[h] = e < 14 * NB;
        hrow[h] = ec / 14;
        hcol[h] = cwrap(k < 7 ? x0 - 7 + k : x0 + 57 + k, W);
        hpos[h] = hrow[h] * NS + (k < 7 ? k : k + 64);
        hin[h] = hcol[h] >= bx0 && hcol[h] < bx0 + bw;
        hA[h] = ADJ ? tx[2 * W + hcol[h]] : zero;
        hB[h] = ADJ ? tx[3 * W + hcol[h]] : zero;
    }

    // block loads (rows tb .. tb+NB-1): x-face coefficients and halo nodes
    R AXV[NB], AXH;
    C HV[RB][NHE];
    int hoff[NHE];  // halo node h of a block: row hrow[h] and column hcol[h] as an offset from the block's first row
#pragma unroll
    for (int h = 0; h < NHE; ++h) hoff[h] = hrow[h] * W + hcol[h];
    const int axh_off = hr * W + fh;
    auto pf_block = [&](auto WRAPc, int tb) {
        if constexpr (decltype(WRAPc)::value) {
#pragma unroll
            for (int p = 0; p < NB; ++p) AXV[p] = (ax + (int64_t)rwrap(tb + p, H) * W)[fa];
            AXH = ax[(int64_t)rwrap(tb + hr, H) * W + fh];
#pragma unroll
            for (int h = 0; h < NHE; ++h)
#pragma unroll
                for (int q = 0; q < RB; ++q)
                    HV[q][h] = rowp(u, q, 0)[(int64_t)rwrap(tb + hrow[h], H) * W + hcol[h]];
        } else {
            const R* axb = ax + (int64_t)tb * W;
#pragma unroll
            for (int p = 0; p < NB; ++p) AXV[p] = axb[(int64_t)p * W + fa];
            AXH = axb[axh_off];
#pragma unroll
            for (int h = 0; h < NHE; ++h)
#pragma unroll
                for (int q = 0; q < RB; ++q) HV[q][h] = rowp(u, q, tb)[hoff[h]];
        }
    };
    // value of a loaded node in the staged field M u (zero outside the box and past the last RHS)
    auto masked = [&](int q, C v, int row, bool colin) -> C {
        return (r0 + q < nrhs && colin && zin(row)) ? v : zero;
    };

    C U[RB][8], G[RB][8];  // rings: field row r in slot (r - i0) & 7, z face f (+1/2) in slot (f - i0) & 7
    int gring = 0;         // bit s: the z face in ring slot s has g != 1 (forward PML term)
    C PF[RB][PD], PM[PD], PB[RB][PD];
    R PA[PD];
    // per-lane load and store addresses of step t: field row t+7+PD, az row
    // t+3+PD, mass and b of row t+PD, output row t (column jc)
    const C* pU[RB];
    const C* pB[RB];
    C* pO[RB];
    const R* pAz;
    const C* pM;
    auto setp = [&](int t) {
        const int ru = rwrap(t + 7 + PD, H), rm = rwrap(t + PD, H), ro = rwrap(t, H);
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            pU[q] = rowp(u, q, ru) + jc;
            pB[q] = MODE == 1 ? rowp(b, q, rm) + jc : nullptr;
            pO[q] = out + (rowp(u, q, ro) - u) + jc;
        }
        pAz = az + (int64_t)rwrap(t + 3 + PD, H) * W + jc;
        pM = mass + (int64_t)rm * W + jc;
    };

    // one row step: t = tb + K (K = (t - i0) & 7); OUT: t is an output row of
    // the run; WRAP: rows are wrapped and addresses formed per step (runs at
    // the box's ends), else they advance one row per step
    auto step = [&](auto Kc, auto OUTc, auto WRAPc, int t, int tb) {
        constexpr int K = decltype(Kc)::value;
        constexpr bool OUT = decltype(OUTc)::value;
        constexpr bool WRAP = decltype(WRAPc)::value;
        constexpr int ks = K % PD;
        auto wr = [&](int v) { return WRAP ? rwrap(v, H) : v; };
        // field row t+7 into the ring; z-face coefficient of face t+3, mass and b of row t
        {
            const int row = wr(t + 7);
#pragma unroll
            for (int q = 0; q < RB; ++q) U[q][(K + 7) & 7] = masked(q, PF[q][ks], row, xin);
        }
        const R a = PA[ks];
        const C mk = PM[ks];
        C bo[RB];
#pragma unroll
        for (int q = 0; q < RB; ++q) bo[q] = PB[q][ks];
        // loads PD rows ahead
        if constexpr (WRAP) setp(t);
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            PF[q][ks] = *pU[q];
            if constexpr (MODE == 1) PB[q][ks] = *pB[q];
            else PB[q][ks] = zero;
        }
        PA[ks] = *pAz;
        PM[ks] = *pM;
        // uniform per-row coefficients: face t+3 and row t (in runs without
        // wrap at constant offsets from the group's first row)
        const C* tzf = tz + tb;          // tzf[k] = sv of face tb + k
        const C* tzg = tz + H + tb;      // g - 1
        const C* tza = tz + 2 * H + tb;  // cA of row tb + k
        const C* tzb = tz + 3 * H + tb;  // cB
        // z face t+3 from rows t .. t+7 (slots K .. K+7)
        {
            const int fz = wr(t + 3);
            const C sv = WRAP ? tz[fz] : tzf[K + 3], gm = WRAP ? tz[H + fz] : tzg[K + 3];
            C cz[8], czb[8];
            bool pz = false;
            if constexpr (!ADJ) {
                constexpr int bit = 1 << ((K + 3) & 7);
                gring = (gring & ~bit) | ((gm.x != R(0) || gm.y != R(0)) ? bit : 0);
            } else {
                pz = gm.x != R(0) || gm.y != R(0);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int row = wr(t + k);
                    cz[k] = WRAP ? tz[2 * H + row] : tza[K + k];
                    czb[k] = pz ? (WRAP ? tz[3 * H + row] : tzb[K + k]) : zero;
                }
            }
#pragma unroll
            for (int q = 0; q < RB; ++q) {
#define UR(k) U[q][(K + (k)) & 7]
                C d;
                if constexpr (!ADJ) {
                    d = c8diff<C, R>(c0, c1, c2, c3, UR(4), UR(3), UR(5), UR(2), UR(6), UR(1), UR(7), UR(0));
                    d = cscale(cm(sv, d), a);
                } else {
                    d = c8diff<C, R>(c0, c1, c2, c3, cjm(cz[4], UR(4)), cjm(cz[3], UR(3)), cjm(cz[5], UR(5)),
                                     cjm(cz[2], UR(2)), cjm(cz[6], UR(6)), cjm(cz[1], UR(1)), cjm(cz[7], UR(7)),
                                     cjm(cz[0], UR(0)));
                    if (pz) {
                        const C d2 = c8diff<C, R>(c0, c1, c2, c3, cjm(czb[4], UR(4)), cjm(czb[3], UR(3)),
                                                  cjm(czb[5], UR(5)), cjm(czb[2], UR(2)), cjm(czb[6], UR(6)),
                                                  cjm(czb[1], UR(1)), cjm(czb[7], UR(7)), cjm(czb[0], UR(0)));
                        d = cadd(d, cjm(gm, d2));
                    }
                    d = cscale(cjm(sv, d), a);  // conj(az sv), az real
                }
#undef UR
                G[q][(K + 3) & 7] = d;
            }
        }
        if constexpr (OUT) {
            if constexpr (K % NB == 0) {
                // x fluxes of block rows t .. t+NB-1 (ring slots K .. K+NB-1)
                lds_barrier();  // the previous block's readers are done
#pragma unroll
                for (int q = 0; q < RB; ++q) {
#pragma unroll
                    for (int p = 0; p < NB; ++p) {
                        const C v = U[q][(K + p) & 7];
                        if constexpr (!ADJ) {
                            st[q][p * NS + 7 + l] = v;
                        } else {
                            st[q][p * NS + 7 + l] = cjm(cAj, v);
                            if (pmlx) sbt[q][p * NS + 7 + l] = cjm(cBj, v);
                        }
                    }
#pragma unroll
                    for (int h = 0; h < NHE; ++h) {
                        if (!hok[h]) continue;
                        const C v = masked(q, HV[q][h], wr(t + hrow[h]), hin[h]);
                        if constexpr (!ADJ) {
                            st[q][hpos[h]] = v;
                        } else {
                            st[q][hpos[h]] = cjm(hA[h], v);
                            if (pmlx) sbt[q][hpos[h]] = cjm(hB[h], v);
                        }
                    }
                }
                lds_barrier();
                auto xflux = [&](int q, int base, C sv, C g, R axf) -> C {
                    const C* s = st[q] + base;  // s[m] = node of face - m .. face + 1 + m
                    C d = c8diff<C, R>(c0, c1, c2, c3, s[1], s[0], s[2], s[-1], s[3], s[-2], s[4], s[-3]);
                    if constexpr (ADJ) {
                        if (pmlx) {
                            const C* sB = sbt[q] + base;
                            const C d2 = c8diff<C, R>(c0, c1, c2, c3, sB[1], sB[0], sB[2], sB[-1], sB[3], sB[-2],
                                                      sB[4], sB[-3]);
                            d = cadd(d, cjm(g, d2));
                        }
                    }
                    return ADJ ? cscale(cjm(sv, d), axf) : cscale(cm(sv, d), axf);
                };
#pragma unroll
                for (int q = 0; q < RB; ++q) {
#pragma unroll
                    for (int p = 0; p < NB; ++p) sf[q][p * NF + l] = xflux(q, p * NS + l + 3, svA, gA, AXV[p]);
                    if (l < 7 * NB) sf[q][hr * NF + hf] = xflux(q, hr * NS + hf + 3, svH, gH, AXH);
                }
                pf_block(WRAPc, t + NB);  // the next block's coefficients and halo nodes
                lds_barrier();
            }
            // output row t
            const int ic = wr(t);
            const bool inside = jok && xin && zin(ic);
            const C cAz = WRAP ? tz[2 * H + ic] : tza[K];
            const bool pzo = !ADJ && gring != 0;  // some face t-4 .. t+3 has g != 1
            C mkc = mk;
            if (ADJ) mkc.y = -mkc.y;
            constexpr int pb = K % NB;
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int r = r0 + q;
                if (t >= i1 || !jok || r >= nrhs) continue;
                C acc;
                if (inside) {
#define GR(k) G[q][(K + 4 + (k)) & 7]
                    const C* fx = sf[q] + pb * NF + l + 4;  // fx[m] = face x0 + l + m (+1/2)
                    const C dx = c8diff<C, R>(c0, c1, c2, c3, fx[0], fx[-1], fx[1], fx[-2], fx[2], fx[-3], fx[3],
                                              fx[-4]);
                    const C dz = c8diff<C, R>(c0, c1, c2, c3, GR(4), GR(3), GR(5), GR(2), GR(6), GR(1), GR(7), GR(0));
                    if constexpr (!ADJ) {
                        acc = cm(cAj, dx);
                        cfma(acc, cAz, dz);
                        if (pmlx) {
                            const C* gg = sgx + l + 4;  // gg[m] = g - 1 at face x0 + l + m
                            const C d2 = c8diff<C, R>(c0, c1, c2, c3, cm(gg[0], fx[0]), cm(gg[-1], fx[-1]),
                                                      cm(gg[1], fx[1]), cm(gg[-2], fx[-2]), cm(gg[2], fx[2]),
                                                      cm(gg[-3], fx[-3]), cm(gg[3], fx[3]), cm(gg[-4], fx[-4]));
                            cfma(acc, cBj, d2);
                        }
                        if (pzo) {
                            C gz[8];  // g - 1 of faces t-4 .. t+3
#pragma unroll
                            for (int k = 0; k < 8; ++k) gz[k] = WRAP ? tz[H + rwrap(t - 4 + k, H)] : tzg[K - 4 + k];
                            const C d2 = c8diff<C, R>(c0, c1, c2, c3, cm(gz[4], GR(4)), cm(gz[3], GR(3)),
                                                      cm(gz[5], GR(5)), cm(gz[2], GR(2)), cm(gz[6], GR(6)),
                                                      cm(gz[1], GR(1)), cm(gz[7], GR(7)), cm(gz[0], GR(0)));
                            cfma(acc, WRAP ? tz[3 * H + ic] : tzb[K], d2);
                        }
                    } else {
                        acc = cadd(dx, dz);
                    }
#undef GR
                    cfma(acc, mkc, U[q][K]);
                } else {
                    acc = u[pO[q] - out];  // identity rows outside the medium
                }
                if (MODE == 1) acc = csub(bo[q], acc);
                *pO[q] = acc;
            }
        }
        if constexpr (!WRAP) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                pU[q] += W;
                if constexpr (MODE == 1) pB[q] += W;
                pO[q] += W;
            }
            pAz += W;
            pM += W;
        }
    };
    // a run: rows i0-7 .. i0-1 (phases 1 .. 7) form z faces i0-4 .. i0+2,
    // then the output rows in groups of eight (past the run's end: computed,
    // not stored)
    auto run = [&](auto WRAPc) {
        pf_block(WRAPc, i0);
#pragma unroll
        for (int k = 1; k < 8; ++k) {  // rows i0-7 .. i0-1
            const int row = rwrap(i0 - 8 + k, H);
#pragma unroll
            for (int q = 0; q < RB; ++q) U[q][k] = masked(q, rowp(u, q, row)[jc], row, xin);
        }
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            // row i0 + d (consumed at step i0+d-7, slot (d + 1) % PD); z face i0-4+d (step i0-7+d, slot (d + 1) % PD)
            const int s = (d + 1) % PD;
            const int row = rwrap(i0 + d, H);
#pragma unroll
            for (int q = 0; q < RB; ++q) PF[q][s] = rowp(u, q, row)[jc];
            PA[s] = (az + (int64_t)rwrap(i0 - 4 + d, H) * W)[jc];
            PM[s] = zero;
#pragma unroll
            for (int q = 0; q < RB; ++q) PB[q][s] = zero;
        }
        if constexpr (!decltype(WRAPc)::value) setp(i0 - 7);
        each_phase(std::integer_sequence<int, 1, 2, 3, 4, 5, 6, 7>{}, [&](auto Kc) {
            step(Kc, std::false_type{}, WRAPc, i0 - 8 + decltype(Kc)::value, i0 - 8);
        });
        for (int tb = i0; tb < i1; tb += 8)
            each_phase(std::integer_sequence<int, 0, 1, 2, 3, 4, 5, 6, 7>{}, [&](auto Kc) {
                step(Kc, std::true_type{}, WRAPc, tb + decltype(Kc)::value, tb);
            });
    };
    // every row the run touches (i0-8 .. its last group's row + 15 and the
    // block loads beyond) inside [0, H): no wrap
    const int last = i0 + ((i1 - i0 + 7) & ~7) - 8;
    if (i0 >= 8 && last + 8 + NB + NB + PD + 8 <= H) run(std::false_type{});
    else run(std::true_type{});
}

static int cond_stream_tz() {
    static const int t = [] {
        const char* v = std::getenv("HELM_COND_TZ");
        const int x = v ? std::atoi(v) : 0;
        return x >= 8 && x <= 4096 ? x : 32;
    }();
    return t;
}
static int cond_stream_rb() {
    static const int t = [] {
        const char* v = std::getenv("HELM_COND_RB");
        const int x = v ? std::atoi(v) : 0;
        return (x == 1 || x == 2 || x == 4) ? x : 1;
    }();
    return t;
}

// rows per x-flux block of the streaming kernel (HELM_COND_NB: 4, the
// default, or 8)
static int cond_stream_nb(int) {
    static const int t = [] {
        const char* v = std::getenv("HELM_COND_NB");
        const int x = v ? std::atoi(v) : 0;
        return x == 8 ? 8 : 4;
    }();
    return t;
}

template <int RB, int NB>
static void launch_cond8s_nb(const Level& L, const float2* mass, const float2* u, float2* out, int nrhs, bool adjoint,
                             const float2* b, hipStream_t s) {
    const int tzr = L.st_tz ? L.st_tz : cond_stream_tz();
    const dim3 block(64), grid((L.nx + 63) / 64, (L.nz + tzr - 1) / tzr, (nrhs + RB - 1) / RB);
    const float ih = (float)(1.0 / L.h_m);
    const float c0 = (float)CD0 * ih, c1 = (float)CD1 * ih, c2 = (float)CD2 * ih, c3 = (float)CD3 * ih;
    const float2* tx = L.ctab32;
    const float2* tz = L.ctab32 + 4 * L.nx;
#define HELM_C8S(M, A, BP)                                                                                       \
    hipLaunchKernelGGL((k_cond8s<float, float2, RB, NB, M, A>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, tzr,  \
                       L.ax32, L.az32, mass, tx, tz, L.cbox[0], L.cbox[1], L.cbox[2], L.cbox[3], c0, c1, c2, c3,   \
                       u, out, BP)
    if (b) {
        if (adjoint) HELM_C8S(1, true, b);
        else HELM_C8S(1, false, b);
    } else {
        if (adjoint) HELM_C8S(0, true, (const float2*)nullptr);
        else HELM_C8S(0, false, (const float2*)nullptr);
    }
#undef HELM_C8S
}
template <int RB>
static void launch_cond8s(const Level& L, const float2* mass, const float2* u, float2* out, int nrhs, bool adjoint,
                          const float2* b, hipStream_t s) {
    if (cond_stream_nb(RB) == 4) launch_cond8s_nb<RB, 4>(L, mass, u, out, nrhs, adjoint, b, s);
    else launch_cond8s_nb<RB, 8>(L, mass, u, out, nrhs, adjoint, b, s);
}

// FP64 (true residuals of the refinement): the streaming kernel with one RHS
// per wavefront and 32-row runs (HELM_COND64_TILED=1: the LDS-tile kernel)
static bool cond64_tiled() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_COND64_TILED");
        return v && std::atoi(v) == 1;
    }();
    return t;
}
static void launch_cond8s64(const Level& L, const double2* mass, const double2* u, double2* out, int nrhs,
                            bool adjoint, const double2* b, hipStream_t s) {
    constexpr int TZR = 32;
    const dim3 block(64), grid((L.nx + 63) / 64, (L.nz + TZR - 1) / TZR, nrhs);
    const double ih = 1.0 / L.h_m;
    const double c0 = CD0 * ih, c1 = CD1 * ih, c2 = CD2 * ih, c3 = CD3 * ih;
    const double2* tx = L.ctab64;
    const double2* tz = L.ctab64 + 4 * L.nx;
#define HELM_C8S64(M, A, BP)                                                                                     \
    hipLaunchKernelGGL((k_cond8s<double, double2, 1, 8, M, A>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, TZR,  \
                       L.ax64, L.az64, mass, tx, tz, L.cbox[0], L.cbox[1], L.cbox[2], L.cbox[3], c0, c1, c2, c3,   \
                       u, out, BP)
    if (b) {
        if (adjoint) HELM_C8S64(1, true, b);
        else HELM_C8S64(1, false, b);
    } else {
        if (adjoint) HELM_C8S64(0, true, (const double2*)nullptr);
        else HELM_C8S64(0, false, (const double2*)nullptr);
    }
#undef HELM_C8S64
}

template <int TZ>
static int cond8_rhs_per_group(const Level& L, int nrhs) {
    const int64_t tiles = (int64_t)((L.nx + C8_TX - 1) / C8_TX) * ((L.nz + TZ - 1) / TZ);
    const int64_t groups = std::max<int64_t>(1, std::min<int64_t>(nrhs, (2048 + tiles - 1) / tiles));
    return (int)((nrhs + groups - 1) / groups);
}

int32_t launch_cond8_32(const Level& L, const float2* u, float2* out, int nrhs, bool adjoint, const float2* b,
                        hipStream_t s) {
    constexpr int TZ = 16;
    HELM_CHECK(L.ax32 && L.az32 && L.cmass32 && L.ctab32, HELM_ERR_STATE, "conditioned operator: medium not set");
    if (cond_stream_tz() > 0 && !std::getenv("HELM_COND_TILED")) {
        // a level with a conjugate mass takes its adjoint through the forward kernel
        const bool hm = adjoint && L.cmassh32;
        const float2* mass = hm ? L.cmassh32 : L.cmass32;
        const bool adj = adjoint && !hm;
        switch (L.st_rb ? L.st_rb : cond_stream_rb()) {
            case 1: launch_cond8s<1>(L, mass, u, out, nrhs, adj, b, s); break;
            case 4: launch_cond8s<4>(L, mass, u, out, nrhs, adj, b, s); break;
            default: launch_cond8s<2>(L, mass, u, out, nrhs, adj, b, s); break;
        }
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    const int rb = cond8_rhs_per_group<TZ>(L, nrhs);
    const dim3 block(64, 4), grid((L.nx + C8_TX - 1) / C8_TX, (L.nz + TZ - 1) / TZ, (nrhs + rb - 1) / rb);
    const float ih = (float)(1.0 / L.h_m);
    const float c0 = (float)CD0 * ih, c1 = (float)CD1 * ih, c2 = (float)CD2 * ih, c3 = (float)CD3 * ih;
    const float2* tx = L.ctab32;
    const float2* tz = L.ctab32 + 4 * L.nx;
#define HELM_C8_32(M, A, BP)                                                                                   \
    hipLaunchKernelGGL((k_cond8<float, float2, TZ, M, A>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.ax32, \
                       L.az32, L.cmass32, tx, tz, L.cbox[0], L.cbox[1], L.cbox[2], L.cbox[3], c0, c1, c2, c3, u,  \
                       out, BP)
    if (b) {
        if (adjoint) HELM_C8_32(1, true, b);
        else HELM_C8_32(1, false, b);
    } else {
        if (adjoint) HELM_C8_32(0, true, (const float2*)nullptr);
        else HELM_C8_32(0, false, (const float2*)nullptr);
    }
#undef HELM_C8_32
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

int32_t launch_cond8_64(const Level& L, const double2* u, double2* out, int nrhs, bool adjoint, const double2* b,
                        hipStream_t s) {
    constexpr int TZ = 8;
    HELM_CHECK(L.ax64 && L.az64 && L.cmass64 && L.ctab64, HELM_ERR_STATE, "conditioned operator: medium not set");
    if (!cond64_tiled()) {
        launch_cond8s64(L, L.cmass64, u, out, nrhs, adjoint, b, s);
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    const int rb = cond8_rhs_per_group<TZ>(L, nrhs);
    const dim3 block(64, 4), grid((L.nx + C8_TX - 1) / C8_TX, (L.nz + TZ - 1) / TZ, (nrhs + rb - 1) / rb);
    const double ih = 1.0 / L.h_m;
    const double c0 = CD0 * ih, c1 = CD1 * ih, c2 = CD2 * ih, c3 = CD3 * ih;
    const double2* tx = L.ctab64;
    const double2* tz = L.ctab64 + 4 * L.nx;
#define HELM_C8_64(M, A, BP)                                                                                     \
    hipLaunchKernelGGL((k_cond8<double, double2, TZ, M, A>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.ax64, \
                       L.az64, L.cmass64, tx, tz, L.cbox[0], L.cbox[1], L.cbox[2], L.cbox[3], c0, c1, c2, c3, u,    \
                       out, BP)
    if (b) {
        if (adjoint) HELM_C8_64(1, true, b);
        else HELM_C8_64(1, false, b);
    } else {
        if (adjoint) HELM_C8_64(0, true, (const double2*)nullptr);
        else HELM_C8_64(0, false, (const double2*)nullptr);
    }
#undef HELM_C8_64
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

// ---- host-side factors (conditioned_wave.py:33-41, 58-69) ---------------
using cd = std::complex<double>;

struct CondAxis {
    int size, offset, pml;
    double dt, dx, c0;
    // damping of FrequencyGeometry (conditioned_wave.py:34-37) at box position t (node index)
    double damping(double t, double stagger) const {
        const double pos = t - offset;
        const double left = pos < pml ? std::pow(1.0 - (pos + stagger) / pml, 4) : 0.0;
        const double right = pos >= size - pml ? std::pow(1.0 - (size - 1 - pos - stagger) / pml, 4) : 0.0;
        return (left + right) * 2.0 * c0 / dx;
    }
    double ar(double t) const { return 1.0 - dt * damping(t, 0.0); }
    double av(double t) const { return 1.0 - dt * damping(t, 0.5); }  // face t + 1/2
};

struct CondFactors {
    cd sv, gm1, cA, cB, left, right;
};

static CondFactors cond_factors(const CondAxis& a, double t, cd D, double k0) {
    const cd q = 1.0 / D, scalar = (D - 1.0) / D / (k0 * k0);
    const double ar = a.ar(t), av = a.av(t);
    const double m = ar != 1.0 ? 1.0 : 0.0;
    CondFactors f;
    f.left = m / (D - ar);
    f.right = 1.0 + (1.0 - m) / (D - ar);
    f.sv = (D - 1.0) / (D - av);
    f.gm1 = (1.0 - av) * q;
    f.cA = scalar * (f.left + f.right);
    f.cB = scalar * f.right;
    return f;
}

// Multigrid stretch of the 5-point approximation at box position t (fine
// node units): node factor s with 1/s = k0^2 scalar (left + right g_mean),
// g_mean the mean gain of the node's two faces, and face factor 1/s_face = sv;
// returned in this library's convention (conjugated).
static double2 cond_node_stretch(const CondAxis& a, double t, cd D, double k0) {
    const CondFactors f = cond_factors(a, t, D, k0);
    const cd gl = 1.0 + (1.0 - a.av(t - 1.0)) / D, gr = 1.0 + f.gm1;
    const cd inv = (D - 1.0) / D * (f.left + f.right * 0.5 * (gl + gr));
    const cd s = std::conj(1.0 / inv);
    return make_double2(s.real(), s.imag());
}
static double2 cond_face_istretch(const CondAxis& a, double tface, cd D) {
    // face at box position tface = t + 1/2 of node t
    const cd sv = (D - 1.0) / (D - a.av(tface - 0.5));
    const cd c = std::conj(sv);
    return make_double2(c.real(), c.imag());
}

static CondAxis cond_axis(const helm_plan* p, bool xaxis) {
    const helm_conditioned_medium& m = p->cond;
    CondAxis a;
    a.size = xaxis ? m.w : m.h;
    a.offset = xaxis ? m.px : m.pz;
    a.pml = m.pml;
    a.dt = m.dt_s;
    a.dx = p->levels[0].h_m;
    a.c0 = p->desc.c0_m_s;
    return a;
}

int32_t cond_level_profiles(const helm_plan* p, int level, std::vector<double2>& sx, std::vector<double2>& sz,
                            std::vector<double2>& isxe, std::vector<double2>& isze) {
    const cd D = std::exp(cd(0.0, 2.0 * M_PI * p->desc.frequency_mhz * 1e6 * p->cond.dt_s));
    const double k0 = p->k0, scale = (double)(1 << level);
    const CondAxis ax = cond_axis(p, true), az = cond_axis(p, false);
    for (size_t j = 0; j < sx.size(); ++j) sx[j] = cond_node_stretch(ax, j * scale, D, k0);
    for (size_t i = 0; i < sz.size(); ++i) sz[i] = cond_node_stretch(az, i * scale, D, k0);
    for (size_t k = 0; k < isxe.size(); ++k) isxe[k] = cond_face_istretch(ax, (k - 0.5) * scale, D);
    for (size_t k = 0; k < isze.size(); ++k) isze[k] = cond_face_istretch(az, (k - 0.5) * scale, D);
    return HELM_OK;
}

// ---- medium (conditioned_wave.py:45-66) ----------------------------------
struct CondGeom {
    int H, W, h, w, pz, px, pml;
    double rho0, c0, dt, epa0, epa1, tau0, tau1;
    double2 D;
};

__device__ __forceinline__ int cond_src(int box, int off, int size, int pml) {
    // box index -> native index of the cropped-and-replicated coefficients
    return min(max(box - off, pml), size - 1 - pml);
}

__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

__global__ void k_cond_medium(CondGeom g, const double* __restrict__ c, const double* __restrict__ rho,
                              const double* __restrict__ alpha, double* ax64, double* az64, float* ax32, float* az32,
                              double2* mass64, float2* mass32, double* eta64, float* eta32, double2* kap64,
                              float2* kap32) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= g.W) return;
    auto at = [&](const double* f, int bi, int bj) {
        bi = cwrap(bi, g.H);
        bj = cwrap(bj, g.W);
        return f[(int64_t)cond_src(bi, g.pz, g.h, g.pml) * g.w + cond_src(bj, g.px, g.w, g.pml)];
    };
    const bool in = i >= g.pz && i < g.pz + g.h && j >= g.px && j < g.px + g.w;
    const double a = g.rho0 / at(rho, i, j);
    const double ae = g.rho0 / at(rho, i, j + 1), as = g.rho0 / at(rho, i + 1, j);
    const double axv = in ? 2.0 * a * ae / (a + ae) : 0.0;
    const double azv = in ? 2.0 * a * as / (a + as) : 0.0;
    const double ci = at(c, i, j), al = alpha ? at(alpha, i, j) : 0.0;
    const double e0 = al * ci / g.c0 * g.epa0, e1 = al * ci / g.c0 * g.epa1;
    const double amp = (1.0 + (e0 + e1) / 2.0) * (1.0 + (e0 + e1) / 2.0);
    const double2 dm1 = make_double2(g.D.x - 1.0, g.D.y);
    const double2 one = make_double2(1.0, 0.0);
    // h = sum_k e_k / (amp (1 + tau_k (D - 1)/dt))
    double2 hs = make_double2(0.0, 0.0);
    {
        const double2 d0 = make_double2(amp * (1.0 + g.tau0 * dm1.x / g.dt), amp * (g.tau0 * dm1.y / g.dt));
        const double2 d1 = make_double2(amp * (1.0 + g.tau1 * dm1.x / g.dt), amp * (g.tau1 * dm1.y / g.dt));
        hs = cadd(zdiv(make_double2(e0, 0.0), d0), zdiv(make_double2(e1, 0.0), d1));
    }
    // modulus = amp (1 - h) / (1 - (1 - 1/D) h)
    const double2 qd = zdiv(one, g.D);
    const double2 omq = make_double2(1.0 - qd.x, -qd.y);
    const double2 num = cscale(csub(one, hs), amp);
    const double2 den = csub(one, cmul(omq, hs));
    const double2 modulus = zdiv(num, den);
    const double r = g.c0 / ci;
    const double2 mk = zdiv(make_double2(a * (r * r), 0.0), modulus);
    const double2 ms = in ? mk : make_double2(0.0, 0.0);
    const int64_t o = (int64_t)i * g.W + j;
    ax64[o] = axv;
    az64[o] = azv;
    ax32[o] = (float)axv;
    az32[o] = (float)azv;
    mass64[o] = ms;
    mass32[o] = make_float2((float)ms.x, (float)ms.y);
    // the multigrid's 5-point approximation, in this library's convention: eta = a, kappa = conj(mass)
    eta64[o] = a;
    eta32[o] = (float)a;
    kap64[o] = make_double2(mk.x, -mk.y);
    kap32[o] = make_float2((float)mk.x, (float)-mk.y);
}

// dense_wave.source_rhs (dense_wave.py:59-62): line [nrhs][w] on box row `face`
__global__ void k_cond_source(CondGeom g, int face, int zi, int xi, double dx, const double* __restrict__ c,
                              const double* __restrict__ rho, const double2* __restrict__ line, double2* rhs,
                              int nrhs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, r = blockIdx.z;
    if (j >= g.W || r >= nrhs) return;
    double2 v = make_double2(0.0, 0.0);
    if (i == face && j >= g.px && j < g.px + g.w) {
        const double sc = c[(int64_t)cond_src(zi + g.pz, g.pz, g.h, g.pml) * g.w + cond_src(xi + g.px, g.px, g.w, g.pml)];
        const double a = g.rho0 / rho[(int64_t)cond_src(i, g.pz, g.h, g.pml) * g.w + cond_src(j, g.px, g.w, g.pml)];
        const double2 l = line[(int64_t)r * g.w + (j - g.px)];
        const double2 scale = zdiv(make_double2(a * g.c0 * g.c0 * 2.0 * g.dt, 0.0),
                                   make_double2(sc * dx * (g.D.x - 1.0), sc * dx * g.D.y));
        v = cmul(l, scale);
    }
    rhs[(int64_t)r * g.H * g.W + (int64_t)i * g.W + j] = v;
}

// preconditioner wrapper: forward z = MG^{-1}(k0^2 sx sz v), adjoint
// z = conj(k0^2 sx sz) MG^{-H}(v); outside the native box z = v (the operator
// is the identity there and decoupled from the box).
__global__ void k_cond_prescale(int H, int W, int64_t n, int nrhs, int bz0, int bx0, int bh, int bw, float k0sq,
                                const float2* __restrict__ sx, const float2* __restrict__ sz,
                                const float2* __restrict__ v, float2* out, const double* __restrict__ fscale) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nrhs) return;
    const int64_t node = e % n;
    const int i = (int)(node / W), j = (int)(node % W);
    float2 r = make_float2(0.f, 0.f);
    if (i >= bz0 && i < bz0 + bh && j >= bx0 && j < bx0 + bw) {
        const float2 f = cscale(cconj(cmul(sx[j], sz[i])), k0sq);  // the profiles hold conj(s)
        r = cmul(f, v[e]);
        if (fscale) r = cscale(r, (float)fscale[e / n]);
    }
    out[e] = r;
}

__global__ void k_cond_post(int H, int W, int64_t n, int nrhs, int bz0, int bx0, int bh, int bw, float k0sq,
                            int adjoint, const float2* __restrict__ sx, const float2* __restrict__ sz,
                            const float2* __restrict__ v, float2* z, const double* __restrict__ fscale) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nrhs) return;
    const int64_t node = e % n;
    const int i = (int)(node / W), j = (int)(node % W);
    if (i >= bz0 && i < bz0 + bh && j >= bx0 && j < bx0 + bw) {
        if (adjoint) z[e] = cmul(cscale(cmul(sx[j], sz[i]), k0sq), z[e]);
    } else {
        z[e] = fscale ? cscale(v[e], (float)fscale[e / n]) : v[e];
    }
}

// forward form of k_cond_post: only the nodes outside the native box (the
// frame of bands above, below, left and right of it) are written, z = v there
__global__ void k_cond_frame(int H, int W, int64_t n, int nrhs, int bz0, int bx0, int bh, int bw,
                             const float2* __restrict__ v, float2* z, const double* __restrict__ fscale) {
    // frame nodes of one RHS: rows outside [bz0, bz0+bh) whole, then the left and right parts of the box rows
    const int64_t top = (int64_t)bz0 * W, bottom = (int64_t)(H - bz0 - bh) * W;
    const int side = W - bw;
    const int64_t per = top + bottom + (int64_t)bh * side;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * nrhs) return;
    const int r = (int)(e / per);
    int64_t k = e - (int64_t)r * per, node;
    if (k < top) {
        node = k;
    } else if (k < top + bottom) {
        node = (int64_t)(bz0 + bh) * W + (k - top);
    } else {
        k -= top + bottom;
        const int64_t row = k / side;
        const int c = (int)(k - row * side);
        node = (bz0 + row) * W + (c < bx0 ? c : c + bw);
    }
    const int64_t o = (int64_t)r * n + node;
    z[o] = fscale ? cscale(v[o], (float)fscale[r]) : v[o];
}

int32_t cond_precond(helm_plan* p, const float2* v, float2* z, int nrhs, bool adjoint, hipStream_t s,
                     const double* fscale) {
    const Level& L = p->levels[0];
    const int64_t cnt = (int64_t)nrhs * L.n;
    const dim3 grid((unsigned)((cnt + 255) / 256)), block(256);
    const float k0sq = (float)(p->k0 * p->k0);
    if (!adjoint) {
        if (p->cond_cap < cnt) {
            if (p->cond_scr) (void)hipFree(p->cond_scr);
            p->cond_scr = nullptr;
            p->cond_cap = 0;
            HELM_HIP(hipMalloc((void**)&p->cond_scr, cnt * sizeof(float2) + 256));
            p->cond_cap = cnt;
            p->generation++;
        }
        hipLaunchKernelGGL(k_cond_prescale, grid, block, 0, s, L.nz, L.nx, L.n, nrhs, L.cbox[0], L.cbox[1],
                           L.cbox[2], L.cbox[3], k0sq, L.sx32, L.sz32, v, p->cond_scr, fscale);
        HELM_HIP(hipGetLastError());
        HELM_TRY(mg_vcycle(p, p->cond_scr, z, nrhs, true, s));
    } else {
        HELM_TRY(mg_vcycle(p, v, z, nrhs, false, s, fscale));
    }
    if (adjoint) {
        hipLaunchKernelGGL(k_cond_post, grid, block, 0, s, L.nz, L.nx, L.n, nrhs, L.cbox[0], L.cbox[1], L.cbox[2],
                           L.cbox[3], k0sq, (int)adjoint, L.sx32, L.sz32, v, z, fscale);
    } else {
        const int64_t frame = ((int64_t)L.nz * L.nx - (int64_t)L.cbox[2] * L.cbox[3]) * nrhs;
        if (frame > 0)
            hipLaunchKernelGGL(k_cond_frame, dim3((unsigned)((frame + 255) / 256)), block, 0, s, L.nz, L.nx, L.n,
                               nrhs, L.cbox[0], L.cbox[1], L.cbox[2], L.cbox[3], v, z, fscale);
    }
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

// ---- material gradient (dense_wave.py:250-255) ---------------------------
// For G = -Re sum_r sum_i conj(lam_ri) M_i ((A v_r)_i - s_ri), v = M (u + pref):
//   S_ax[q]  = dG-part / d ax_q = -sum_r sv_x (D+x v_r)_q [(D+x yA_r)_q + (g_x-1)_q (D+x yB_r)_q],
//   yA = conj(lam) M cA, yB = conj(lam) M cB   (D-^T = -D+), likewise S_az,
//   S_m[q]   = sum_r conj(lam_rq) M_q v_rq   (coefficient of mass_q),
//   T[q]     = sum_r conj(lam_rq) M_q line_r K / c_src on the source row (coefficient of a_q in s).
__global__ void k_cond_sens(int H, int W, int64_t n, int nrhs, const double2* __restrict__ tx,
                            const double2* __restrict__ tz, int bz0, int bx0, int bh, int bw, double c0, double c1,
                            double c2, double c3, const double2* __restrict__ v, const double2* __restrict__ lam,
                            double2* s_ax, double2* s_az, double2* s_m) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= W) return;
    auto in = [&](int zi, int xj) { return zi >= bz0 && zi < bz0 + bh && xj >= bx0 && xj < bx0 + bw; };
    const double cf[4] = {c0, c1, c2, c3};
    double2 sax = make_double2(0, 0), saz = make_double2(0, 0), sm = make_double2(0, 0);
    const double2 svx = tx[j], gx = tx[W + j], svz = tz[i], gz = tz[H + i];
    for (int r = 0; r < nrhs; ++r) {
        const double2* vr = v + (int64_t)r * n;
        const double2* lr = lam + (int64_t)r * n;
        auto vm = [&](int zi, int xj) {
            zi = cwrap(zi, H);
            xj = cwrap(xj, W);
            return in(zi, xj) ? vr[(int64_t)zi * W + xj] : make_double2(0, 0);
        };
        // y = conj(lam) M coef at a wrapped node
        auto yl = [&](int zi, int xj, double2 coef) {
            zi = cwrap(zi, H);
            xj = cwrap(xj, W);
            if (!in(zi, xj)) return make_double2(0, 0);
            return cmul(cconj(lr[(int64_t)zi * W + xj]), coef);
        };
        double2 dvx = make_double2(0, 0), dvz = dvx, dax = dvx, dbx = dvx, daz = dvx, dbz = dvx;
        for (int m = 0; m < 4; ++m) {
            const int jp = cwrap(j + m + 1, W), jm = cwrap(j - m, W);
            const int ip = cwrap(i + m + 1, H), im = cwrap(i - m, H);
            dvx = cadd(dvx, cscale(csub(vm(i, jp), vm(i, jm)), cf[m]));
            dvz = cadd(dvz, cscale(csub(vm(ip, j), vm(im, j)), cf[m]));
            dax = cadd(dax, cscale(csub(yl(i, jp, tx[2 * W + jp]), yl(i, jm, tx[2 * W + jm])), cf[m]));
            dbx = cadd(dbx, cscale(csub(yl(i, jp, tx[3 * W + jp]), yl(i, jm, tx[3 * W + jm])), cf[m]));
            daz = cadd(daz, cscale(csub(yl(ip, j, tz[2 * H + ip]), yl(im, j, tz[2 * H + im])), cf[m]));
            dbz = cadd(dbz, cscale(csub(yl(ip, j, tz[3 * H + ip]), yl(im, j, tz[3 * H + im])), cf[m]));
        }
        sax = csub(sax, cmul(cmul(svx, dvx), cadd(dax, cmul(gx, dbx))));
        saz = csub(saz, cmul(cmul(svz, dvz), cadd(daz, cmul(gz, dbz))));
        if (in(i, j)) sm = cadd(sm, cmul(cconj(lr[(int64_t)i * W + j]), vr[(int64_t)i * W + j]));
    }
    const int64_t o = (int64_t)i * W + j;
    s_ax[o] = sax;
    s_az[o] = saz;
    s_m[o] = sm;
}

// chain rule to the native maps: box node q takes native node map(q) (crop +
// replicate), a = rho0/rho; grad_a is accumulated into grad_rho and turned into
// d/d rho by k_cond_grad_rho.
__global__ void k_cond_grad(CondGeom g, int nrhs, int face, int zsrc, int xsrc, double dx,
                            const double* __restrict__ c, const double* __restrict__ rho,
                            const double* __restrict__ alpha, const double2* __restrict__ s_ax,
                            const double2* __restrict__ s_az, const double2* __restrict__ s_m,
                            const double2* __restrict__ lam, const double2* __restrict__ line, double* grad_c,
                            double* grad_a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= g.W) return;
    auto in = [&](int zi, int xj) { return zi >= g.pz && zi < g.pz + g.h && xj >= g.px && xj < g.px + g.w; };
    auto nat = [&](int zi, int xj) {
        zi = cwrap(zi, g.H);
        xj = cwrap(xj, g.W);
        return (int64_t)cond_src(zi, g.pz, g.h, g.pml) * g.w + cond_src(xj, g.px, g.w, g.pml);
    };
    auto aat = [&](int zi, int xj) { return g.rho0 / rho[nat(zi, xj)]; };
    auto d1 = [](double a, double b) { return 2.0 * b * b / ((a + b) * (a + b)); };  // d/da 2ab/(a+b)
    auto d2 = [](double a, double b) { return 2.0 * a * a / ((a + b) * (a + b)); };  // d/db
    const int64_t q = (int64_t)i * g.W + j;
    const int jw = cwrap(j - 1, g.W), iw = cwrap(i - 1, g.H);
    const double a = aat(i, j);
    double ga = 0.0;
    // faces q+1/2 (a_q first) and q-1/2 (a_q second); the face coefficients carry M of their node
    if (in(i, j)) {
        ga -= s_ax[q].x * d1(a, aat(i, j + 1)) + s_az[q].x * d1(a, aat(i + 1, j));
    }
    if (in(i, jw)) ga -= s_ax[(int64_t)i * g.W + jw].x * d2(aat(i, jw), a);
    if (in(iw, j)) ga -= s_az[(int64_t)iw * g.W + j].x * d2(aat(iw, j), a);
    double gc = 0.0;
    if (in(i, j)) {
        // mass = a c0^2 / (c^2 K(c)), K the relaxation modulus
        const double cv = c[nat(i, j)], al = alpha ? alpha[nat(i, j)] : 0.0;
        const double e0 = al * cv / g.c0 * g.epa0, e1 = al * cv / g.c0 * g.epa1;
        const double de0 = al / g.c0 * g.epa0, de1 = al / g.c0 * g.epa1;
        const double ah = 1.0 + (e0 + e1) / 2.0, A = ah * ah, dA = ah * (de0 + de1);
        const double2 dm1 = make_double2(g.D.x - 1.0, g.D.y);
        const double2 dk0 = make_double2(1.0 + g.tau0 * dm1.x / g.dt, g.tau0 * dm1.y / g.dt);
        const double2 dk1 = make_double2(1.0 + g.tau1 * dm1.x / g.dt, g.tau1 * dm1.y / g.dt);
        const double2 i0 = zdiv(make_double2(1.0, 0.0), dk0), i1 = zdiv(make_double2(1.0, 0.0), dk1);
        const double2 hs = cadd(cscale(i0, e0 / A), cscale(i1, e1 / A));
        const double2 dh = cadd(cscale(i0, (de0 * A - e0 * dA) / (A * A)), cscale(i1, (de1 * A - e1 * dA) / (A * A)));
        const double2 qd = zdiv(make_double2(1.0, 0.0), g.D);
        const double2 one = make_double2(1.0, 0.0);
        const double2 den = csub(one, cmul(make_double2(1.0 - qd.x, -qd.y), hs));
        const double2 omh = csub(one, hs);
        const double2 K = zdiv(cscale(omh, A), den);
        const double2 dK = csub(zdiv(cscale(omh, dA), den), zdiv(cscale(cmul(qd, dh), A), cmul(den, den)));
        const double2 iK = zdiv(one, K);
        const double c2 = g.c0 * g.c0;
        const double2 dm_da = cscale(iK, c2 / (cv * cv));
        // d mass / dc = a c0^2 (-2/(c^3 K) - dK/(c^2 K^2))
        const double2 dm_dc = cscale(csub(cscale(iK, -2.0 / (cv * cv * cv)), cscale(cmul(dK, cmul(iK, iK)), 1.0 / (cv * cv))),
                                     a * c2);
        const double2 sm = s_m[q];
        ga -= sm.x * dm_da.x - sm.y * dm_da.y;  // Re(s_m dm/da)
        gc -= sm.x * dm_dc.x - sm.y * dm_dc.y;
        if (line && i == face) {
            // T_q = sum_r conj(lam) line K / c_src,  K = c0^2 2 dt / (dx (D - 1))
            const double csrc = c[(int64_t)cond_src(zsrc + g.pz, g.pz, g.h, g.pml) * g.w +
                                  cond_src(xsrc + g.px, g.px, g.w, g.pml)];
            const double2 Kc = zdiv(make_double2(c2 * 2.0 * g.dt, 0.0), make_double2(dx * dm1.x, dx * dm1.y));
            double2 t = make_double2(0, 0);
            for (int r = 0; r < nrhs; ++r)
                t = cadd(t, cmul(cconj(lam[(int64_t)r * g.H * g.W + q]), line[(int64_t)r * g.w + (j - g.px)]));
            t = cscale(cmul(t, Kc), 1.0 / csrc);
            ga += t.x;
            atomicAdd(&grad_c[(int64_t)cond_src(zsrc + g.pz, g.pz, g.h, g.pml) * g.w +
                              cond_src(xsrc + g.px, g.px, g.w, g.pml)],
                      -a * t.x / csrc);
        }
    }
    if (ga != 0.0) atomicAdd(&grad_a[nat(i, j)], ga);
    if (gc != 0.0) atomicAdd(&grad_c[nat(i, j)], gc);
}

__global__ void k_cond_grad_rho(int64_t n, double rho0, const double* __restrict__ rho, double* grad) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) grad[k] *= -rho0 / (rho[k] * rho[k]);
}

template <typename T>
static int32_t c_alloc(T** ptr, size_t count) {
    if (*ptr) return HELM_OK;
    hipError_t e = hipMalloc((void**)ptr, count * sizeof(T) + 256);
    if (e != hipSuccess) {
        *ptr = nullptr;
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return HELM_ERR_ALLOC;
    }
    return HELM_OK;
}

static CondGeom cond_geom(const helm_plan* p) {
    const helm_conditioned_medium& m = p->cond;
    const Level& L = p->levels[0];
    CondGeom g{};
    g.H = L.nz;
    g.W = L.nx;
    g.h = m.h;
    g.w = m.w;
    g.pz = m.pz;
    g.px = m.px;
    g.pml = m.pml;
    g.rho0 = m.rho0;
    g.c0 = p->desc.c0_m_s;
    g.dt = m.dt_s;
    g.epa0 = m.eta_per_unit_alpha[0];
    g.epa1 = m.eta_per_unit_alpha[1];
    g.tau0 = m.taus_s[0];
    g.tau1 = m.taus_s[1];
    const double ph = 2.0 * M_PI * p->desc.frequency_mhz * 1e6 * m.dt_s;
    g.D = make_double2(std::cos(ph), std::sin(ph));
    return g;
}

}  // namespace helm

using namespace helm;

extern "C" {

int32_t helm_set_medium_conditioned(helm_plan* p, const double* c, const double* rho, const double* alpha,
                                    const helm_conditioned_medium* m, void* stream) {
    HELM_CHECK(p && c && rho && m, HELM_ERR_ARG, "helm_set_medium_conditioned: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_COND8, HELM_ERR_STATE,
               "helm_set_medium_conditioned: the plan was not created with op_kind HELM_OP_COND8");
    Level& L = p->levels[0];
    HELM_CHECK(m->h >= 1 && m->w >= 1 && m->pz >= 0 && m->px >= 0 && m->pz + m->h <= L.nz && m->px + m->w <= L.nx,
               HELM_ERR_ARG, "helm_set_medium_conditioned: the medium does not fit the box");
    HELM_CHECK(m->pz >= 8 && m->px >= 8 && L.nz - m->pz - m->h >= 8 && L.nx - m->px - m->w >= 8, HELM_ERR_ARG,
               "helm_set_medium_conditioned: the box must pad the medium by >= 8 nodes on every side");
    HELM_CHECK(m->pml >= 1 && 2 * m->pml < m->h && 2 * m->pml < m->w, HELM_ERR_ARG,
               "helm_set_medium_conditioned: pml must be >= 1 and leave a medium core");
    HELM_CHECK(std::isfinite(m->dt_s) && m->dt_s > 0 && std::isfinite(m->rho0) && m->rho0 > 0, HELM_ERR_ARG,
               "helm_set_medium_conditioned: dt_s and rho0 must be finite and positive");
    HELM_CHECK(std::sin(M_PI * p->desc.frequency_mhz * 1e6 * m->dt_s) > 0, HELM_ERR_ARG,
               "helm_set_medium_conditioned: frequency * dt_s must lie in (0, 1)");
    hipStream_t s = (hipStream_t)stream;
    p->cond = *m;
    p->k0 = 2.0 * std::sin(M_PI * p->desc.frequency_mhz * 1e6 * m->dt_s) / (m->dt_s * p->desc.c0_m_s);
    HELM_TRY(c_alloc(&L.ax64, L.n));
    HELM_TRY(c_alloc(&L.az64, L.n));
    HELM_TRY(c_alloc(&L.ax32, L.n));
    HELM_TRY(c_alloc(&L.az32, L.n));
    HELM_TRY(c_alloc(&L.cmass64, L.n));
    HELM_TRY(c_alloc(&L.cmass32, L.n));
    HELM_TRY(c_alloc(&L.ctab64, (size_t)4 * (L.nx + L.nz)));
    HELM_TRY(c_alloc(&L.ctab32, (size_t)4 * (L.nx + L.nz)));
    const CondGeom g = cond_geom(p);
    hipLaunchKernelGGL(k_cond_medium, dim3((L.nx + 255) / 256, L.nz), dim3(256), 0, s, g, c, rho, alpha, L.ax64,
                       L.az64, L.ax32, L.az32, L.cmass64, L.cmass32, L.eta64, L.eta32, L.kap64, L.kap32);
    HELM_HIP(hipGetLastError());
    // per-column and per-row factors
    const cd D(g.D.x, g.D.y);
    std::vector<double2> t64((size_t)4 * (L.nx + L.nz));
    std::vector<float2> t32(t64.size());
    auto fill = [&](const CondAxis& a, int n, size_t base) {
        for (int k = 0; k < n; ++k) {
            const CondFactors f = cond_factors(a, k, D, p->k0);
            const cd v[4] = {f.sv, f.gm1, f.cA, f.cB};
            for (int t = 0; t < 4; ++t) t64[base + (size_t)t * n + k] = make_double2(v[t].real(), v[t].imag());
        }
    };
    fill(cond_axis(p, true), L.nx, 0);
    fill(cond_axis(p, false), L.nz, (size_t)4 * L.nx);
    for (size_t k = 0; k < t64.size(); ++k) t32[k] = make_float2((float)t64[k].x, (float)t64[k].y);
    HELM_HIP(hipMemcpyAsync(L.ctab64, t64.data(), t64.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    HELM_HIP(hipMemcpyAsync(L.ctab32, t32.data(), t32.size() * sizeof(float2), hipMemcpyHostToDevice, s));
    HELM_HIP(hipStreamSynchronize(s));  // the host tables go out of scope
    L.op = HELM_OP_COND8;
    L.cbox[0] = m->pz;
    L.cbox[1] = m->px;
    L.cbox[2] = m->h;
    L.cbox[3] = m->w;
    // multigrid stretch profiles of every level from the split-PML factors
    for (size_t l = 0; l < p->levels.size(); ++l) HELM_TRY(upload_cond_profiles(p, (int)l));
    p->medium_set = true;
    p->generation++;
    if (p->mg.levels > 0) HELM_TRY(mg_build(p, s));
    return HELM_OK;
}

int32_t helm_conditioned_source(helm_plan* p, const double* c, const double* rho, const void* line, int32_t face,
                                int32_t source_z, int32_t source_x, void* rhs, int32_t nrhs, void* stream) {
    HELM_CHECK(p && c && rho && line && rhs, HELM_ERR_ARG, "helm_conditioned_source: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_COND8 && p->medium_set, HELM_ERR_STATE,
               "helm_conditioned_source: call helm_set_medium_conditioned first");
    const Level& L = p->levels[0];
    HELM_CHECK(face >= 0 && face < L.nz, HELM_ERR_ARG, "helm_conditioned_source: face row outside the box");
    HELM_CHECK(nrhs >= 1 && nrhs <= 65535 && (int64_t)nrhs * L.n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_conditioned_source: nrhs out of range");
    const CondGeom g = cond_geom(p);
    hipLaunchKernelGGL(k_cond_source, dim3((L.nx + 255) / 256, L.nz, nrhs), dim3(256), 0, (hipStream_t)stream, g,
                       face, source_z, source_x, L.h_m, c, rho, (const double2*)line, (double2*)rhs, nrhs);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

int32_t helm_conditioned_gradient(helm_plan* p, const double* c, const double* rho, const double* alpha,
                                  const void* v, const void* lam, const void* line, int32_t nrhs, int32_t face,
                                  int32_t source_z, int32_t source_x, double* grad_c, double* grad_rho,
                                  void* stream) {
    HELM_CHECK(p && c && rho && v && lam && grad_c && grad_rho, HELM_ERR_ARG,
               "helm_conditioned_gradient: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_COND8 && p->medium_set, HELM_ERR_STATE,
               "helm_conditioned_gradient: call helm_set_medium_conditioned first");
    const Level& L = p->levels[0];
    HELM_CHECK(nrhs >= 1 && (int64_t)nrhs * L.n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_conditioned_gradient: nrhs out of range");
    HELM_CHECK(!line || (face >= 0 && face < L.nz), HELM_ERR_ARG, "helm_conditioned_gradient: face outside the box");
    hipStream_t s = (hipStream_t)stream;
    const CondGeom g = cond_geom(p);
    if (!p->cond_grad) HELM_HIP(hipMalloc((void**)&p->cond_grad, 3 * L.n * sizeof(double2) + 256));
    double2* sax = p->cond_grad;
    double2* saz = sax + L.n;
    double2* sm = saz + L.n;
    const size_t nn = (size_t)g.h * g.w;
    HELM_HIP(hipMemsetAsync(grad_c, 0, nn * sizeof(double), s));
    HELM_HIP(hipMemsetAsync(grad_rho, 0, nn * sizeof(double), s));
    const double ih = 1.0 / L.h_m;
    const dim3 grid((L.nx + 255) / 256, L.nz), block(256);
    hipLaunchKernelGGL(k_cond_sens, grid, block, 0, s, L.nz, L.nx, L.n, nrhs, (const double2*)L.ctab64,
                       (const double2*)(L.ctab64 + 4 * L.nx), L.cbox[0], L.cbox[1], L.cbox[2], L.cbox[3], CD0 * ih,
                       CD1 * ih, CD2 * ih, CD3 * ih, (const double2*)v, (const double2*)lam, sax, saz, sm);
    hipLaunchKernelGGL(k_cond_grad, grid, block, 0, s, g, nrhs, face, source_z, source_x, L.h_m, c, rho, alpha, sax,
                       saz, sm, (const double2*)lam, (const double2*)line, grad_c, grad_rho);
    hipLaunchKernelGGL(k_cond_grad_rho, dim3((unsigned)((nn + 255) / 256)), block, 0, s, (int64_t)nn, g.rho0, rho,
                       grad_rho);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

}  // extern "C"
