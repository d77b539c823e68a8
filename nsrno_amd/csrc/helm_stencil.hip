// Matrix-free flux-form Helmholtz stencil, port injection and sampling.
//
// Operator (nsrno/theranostic/wave.py:174-188, assembled there as a sparse
// matrix; applied here without forming it):
//   (A u)[i,j] = dg u[i,j] + fw u[i,j-1] + fe u[i,j+1] + fn u[i-1,j] + fs u[i+1,j]
//   fw = sz_i / sxe_j     * (eta[i,j] + eta[i,j-1]) / (2 h^2)   (ghost eta = 1)
//   fe = sz_i / sxe_{j+1} * (eta[i,j] + eta[i,j+1]) / (2 h^2)
//   fn = sx_j / sze_i     * (eta[i,j] + eta[i-1,j]) / (2 h^2)
//   fs = sx_j / sze_{i+1} * (eta[i,j] + eta[i+1,j]) / (2 h^2)
//   dg = k0^2 sz_i sx_j kappa_c[i,j] - (fw + fe + fn + fs)
// Neighbours outside the grid are the zero ghosts of wave.py:109-115.
// The adjoint A^H is conj(A) because A is complex symmetric (wave.py:12-14).
//
// Memory layout in HBM: fields [rhs][z][x] contiguous (x fastest, coalesced
// along a 64-lane wavefront), per-node eta (f32) and kappa_c (c64) read once
// per node and reused for every right-hand side in registers; 1-D complex
// profiles are tiny and stay in cache.  Algorithmic bytes per node per RHS:
// 8 (read u) + 8 (write Au) + 12/R (coefficients).
#include <algorithm>
#include <cstdlib>

#include "helm_internal.h"
#include "helm_coef.h"

namespace helm {

// out[r] = A u[r]  (MODE 0)   or   out[r] = b[r] - A u[r]  (MODE 1)
template <typename R, typename C, int MODE>
__global__ __launch_bounds__(256) void k_apply(int nz, int nx, int64_t n, int nrhs, int rb, const C* __restrict__ sx,
                                               const C* __restrict__ sz, const C* __restrict__ isxe,
                                               const C* __restrict__ isze, const R* __restrict__ eta,
                                               const C* __restrict__ kap, R k0sq, R hh, R shift, int conj,
                                               const C* __restrict__ u, C* __restrict__ out,
                                               const C* __restrict__ b) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nz || j >= nx) return;
    Coef<C> c;
    node_coef<R, C>(i, j, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, conj != 0, c);
    const int64_t idx = (int64_t)i * nx + j;
    const bool hw = j > 0, he = j < nx - 1, hn = i > 0, hs = i < nz - 1;
    const int r1 = min(nrhs, (int)(blockIdx.z + 1) * rb);
    for (int r = blockIdx.z * rb; r < r1; ++r) {
        const C* ur = u + (int64_t)r * n;
        C acc = cmul(c.dg, ur[idx]);
        if (hw) cfma(acc, c.fw, ur[idx - 1]);
        if (he) cfma(acc, c.fe, ur[idx + 1]);
        if (hn) cfma(acc, c.fn, ur[idx - nx]);
        if (hs) cfma(acc, c.fs, ur[idx + nx]);
        if (MODE == 1) acc = csub(b[(int64_t)r * n + idx], acc);
        out[(int64_t)r * n + idx] = acc;
    }
}

// Streaming form of the same operator (FP32): one wavefront per (64-column
// strip, run of tz rows, group of RB right-hand sides).  Each lane owns one
// column and walks down the run holding rows i-1, i, i+1 of every RHS of its
// group in registers, with row i+2 already in flight; the west/east
// neighbours come from the adjacent lanes (cross-lane shuffles), and only
// the two edge lanes load one extra value per row.  The node coefficients
// are formed once per node for the whole group.  Each u value is read from
// HBM once, plus (tz+2)/tz for the run's halo rows.  Each row step waits for
// one memory latency, so short runs (8 rows) with many wavefronts in flight
// do best.  eta is carried down the column in a register window and its
// west/east values come from the neighbouring lanes, so a node's stencil
// costs two loads (eta of row i+2, kappa) instead of the seven of node_coef.
// (Loading the edge lanes' halo values two rows ahead as well costs more
// registers than it saves: measured 10% slower.)
__device__ __forceinline__ float2 shfl_up1(float2 v) {
    return make_float2(__shfl_up(v.x, 1, 64), __shfl_up(v.y, 1, 64));
}
__device__ __forceinline__ float2 shfl_down1(float2 v) {
    return make_float2(__shfl_down(v.x, 1, 64), __shfl_down(v.y, 1, 64));
}

template <int RB, int MODE>
__global__ __launch_bounds__(256) void k_apply_stream(int nz, int nx, int64_t n, int nrhs, int tz,
                                                      const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                      const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                      const float* __restrict__ eta, const float2* __restrict__ kap,
                                                      float k0sq, float hh, float shift, int conj,
                                                      const float2* __restrict__ u, float2* __restrict__ out,
                                                      const float2* __restrict__ b) {
    const int lane = threadIdx.x;
    const int x0 = blockIdx.x * 64, j = x0 + lane;
    const int i0 = __builtin_amdgcn_readfirstlane((blockIdx.y * 4 + threadIdx.y) * tz);  // uniform per wavefront
    if (i0 >= nz) return;
    const int i1 = min(nz, i0 + tz);
    const int r0 = blockIdx.z * RB;
    const bool colok = j < nx;
    const int jj = colok ? j : nx - 1;  // lanes past the grid mirror the last column (results dropped)
    const bool hw = jj > 0, he = jj < nx - 1;
    const bool cj = conj != 0;
    // u and eta at (i, col), zero (eta: 1) outside the grid or past the last
    // RHS; the load is unconditional (row and RHS clamped into range) so that
    // no branch splits the loads a row step issues
    auto ld = [=](int r, int i, int col) -> float2 {
        const float2 v = u[(int64_t)(r < nrhs ? r : nrhs - 1) * n + (int64_t)min(max(i, 0), nz - 1) * nx + col];
        return keep_or_zero(v, r < nrhs && i >= 0 && i < nz);
    };
    auto le = [=](int i, int col) -> float {
        return keep_or(eta[(int64_t)min(max(i, 0), nz - 1) * nx + col], i >= 0 && i < nz, 1.f);
    };
    // per-column stretch factors of this lane (constant down the run)
    const float2 sxj = sx[jj], iw = isxe[jj], ie = isxe[jj + 1];
    float2 un[RB], uc[RB], us[RB], up[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        un[q] = ld(r0 + q, i0 - 1, jj);
        uc[q] = ld(r0 + q, i0, jj);
        us[q] = ld(r0 + q, i0 + 1, jj);
    }
    // eta window down the column: rows i-1, i, i+1
    float en = le(i0 - 1, jj), ec = le(i0, jj), es = le(i0 + 1, jj);
    // Row step i uses operands loaded one step earlier (u two rows ahead):
    // row i + 2 of u and eta, and for row i + 1 the strip's halo columns
    // (edge lanes), the edge lanes' outer eta and kappa.
    const int xw = x0 > 0 ? x0 - 1 : 0, xe = x0 + 64 < nx ? x0 + 64 : nx - 1;
    const int jwe = lane == 0 ? (hw ? jj - 1 : jj) : (he ? jj + 1 : jj);  // edge lanes' outer eta column
    float2 pw[RB], pe[RB], pk;
    float pup, peh;
    auto pf_row = [&](int i) {  // operands of row step i (i may pass the run: loaded, unused)
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            up[q] = ld(r0 + q, i + 2, jj);
            pw[q] = keep_or_zero(ld(r0 + q, i, xw), lane == 0 && x0 > 0);
            pe[q] = keep_or_zero(ld(r0 + q, i, xe), lane == 63 && x0 + 64 < nx);
        }
        pup = le(i + 2, jj);
        const int ic = min(i, nz - 1);
        pk = kap[(int64_t)ic * nx + jj];
        peh = eta[(int64_t)ic * nx + jwe];
    };
    pf_row(i0);
    for (int i = i0; i < i1; ++i) {
        float2 ew[RB], ee[RB], upc[RB];
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            ew[q] = pw[q];
            ee[q] = pe[q];
            upc[q] = up[q];
        }
        const float eup = pup, ehl = peh;
        const float2 kc = pk;
        pf_row(i + 1);
        // west/east eta from the neighbouring lanes (edge lanes use theirs)
        float eww = __shfl_up(ec, 1, 64), eee = __shfl_down(ec, 1, 64);
        if (lane == 0) eww = hw ? ehl : 1.f;
        if (lane == 63) eee = he ? ehl : 1.f;
        if (!hw) eww = 1.f;
        if (!he) eee = 1.f;
        // the node's stencil, as node_coef forms it
        const float2 szi = sz[i], in_ = isze[i], is_ = isze[i + 1];
        Coef<float2> c;
        c.fw = cscale(cmul(szi, iw), hh * (ec + eww));
        c.fe = cscale(cmul(szi, ie), hh * (ec + eee));
        c.fn = cscale(cmul(sxj, in_), hh * (ec + en));
        c.fs = cscale(cmul(sxj, is_), hh * (ec + es));
        float2 mass = cscale(cmul(cmul(szi, sxj), kc), k0sq);
        if (shift != 0.f) {
            const float2 m2 = make_float2(mass.x - shift * mass.y, mass.y + shift * mass.x);
            mass = m2;
        }
        c.dg = csub(mass, cadd(cadd(c.fw, c.fe), cadd(c.fn, c.fs)));
        if (cj) {
            c.dg = cconj(c.dg);
            c.fw = cconj(c.fw);
            c.fe = cconj(c.fe);
            c.fn = cconj(c.fn);
            c.fs = cconj(c.fs);
        }
        const bool hn = i > 0, hs = i < nz - 1;
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            float2 w = shfl_up1(uc[q]), e = shfl_down1(uc[q]);
            if (lane == 0) w = ew[q];
            if (lane == 63) e = ee[q];
            float2 acc = cmul(c.dg, uc[q]);
            if (hw) cfma(acc, c.fw, w);
            if (he) cfma(acc, c.fe, e);
            if (hn) cfma(acc, c.fn, un[q]);
            if (hs) cfma(acc, c.fs, us[q]);
            const int r = r0 + q;
            if (colok && r < nrhs) {
                const int64_t o = (int64_t)r * n + (int64_t)i * nx + j;
                if (MODE == 1) acc = csub(b[o], acc);
                out[o] = acc;
            }
            un[q] = uc[q];
            uc[q] = us[q];
            us[q] = upc[q];
        }
        en = ec;
        ec = es;
        es = eup;
    }
}

// Block form of the streaming kernel: the run length TZ is a compile-time
// constant and every value the run needs (TZ + 2 field rows of each RHS of
// the group, eta, kappa, and the strip's west/east halo columns) is loaded
// before the first node is formed, so a wavefront waits for one memory
// latency per run instead of one per row.  The halo columns are packed one
// value per lane (lane q * TZ + t holds row t of RHS q) and handed to the
// edge lane with a readlane, so they cost two registers instead of 2 RB TZ.
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <int RB, int TZ, int MODE>
__device__ __forceinline__ void k_apply_block_body(const VBlock& B, int nz, int nx, int64_t n, int nrhs,
                                                     const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                     const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                     const float* __restrict__ eta, const float2* __restrict__ kap,
                                                     float k0sq, float hh, float shift, int conj,
                                                     const float2* __restrict__ u, float2* __restrict__ out,
                                                     const float2* __restrict__ b, int xcd) {
    static_assert(RB * TZ <= 64, "halo columns are packed one value per lane");
    int bx, by, bz;
    xcd_tile(B, bx, by, bz, xcd != 0, xcd == 2);
    const int lane = threadIdx.x;
    const int x0 = bx * 64, j = x0 + lane;
    // the run's first row, uniform over the wavefront (so the row profiles below are scalar loads)
    const int i0 = __builtin_amdgcn_readfirstlane((by * 4 + threadIdx.y) * TZ);
    if (i0 >= nz) return;
    const int r0 = bz * RB;
    const bool colok = j < nx;
    const int jj = colok ? j : nx - 1;  // lanes past the grid mirror the last column (results dropped)
    const bool hw = jj > 0, he = jj < nx - 1;
    const bool wstrip = x0 > 0, estrip = x0 + 64 < nx;
    // row profiles of the run (rows past the grid clamped; their results are dropped)
    float2 SZ[TZ], IZ[TZ + 1];
#pragma unroll
    for (int t = 0; t < TZ; ++t) SZ[t] = sz[min(i0 + t, nz - 1)];
#pragma unroll
    for (int t = 0; t <= TZ; ++t) IZ[t] = isze[min(i0 + t, nz)];
    float2 U[RB][TZ + 2];
    float E[TZ + 2];
    float2 K[TZ];
#pragma unroll
    for (int t = 0; t < TZ + 2; ++t) {
        const int i = i0 - 1 + t;
        const bool rok = i >= 0 && i < nz;
#pragma unroll
        for (int q = 0; q < RB; ++q)
            U[q][t] = (rok && r0 + q < nrhs) ? u[(int64_t)(r0 + q) * n + (int64_t)i * nx + jj] : make_float2(0.f, 0.f);
        E[t] = rok ? eta[(int64_t)i * nx + jj] : 1.f;  // the operator's ghost value outside the grid
    }
#pragma unroll
    for (int t = 0; t < TZ; ++t)
        K[t] = i0 + t < nz ? kap[(int64_t)(i0 + t) * nx + jj] : make_float2(0.f, 0.f);
    // halo columns x0 - 1 and x0 + 64: lane q * TZ + t holds row i0 + t of RHS r0 + q
    float2 HW = make_float2(0.f, 0.f), HE = make_float2(0.f, 0.f);
    float EH = 1.f;  // lanes 0..TZ-1: west eta of row t; lanes 32..32+TZ-1: east eta
    {
        const int q = lane / TZ, t = lane % TZ, i = i0 + t;
        if (lane < RB * TZ && i < nz && r0 + q < nrhs) {
            const int64_t o = (int64_t)(r0 + q) * n + (int64_t)i * nx;
            if (wstrip) HW = u[o + x0 - 1];
            if (estrip) HE = u[o + x0 + 64];
        }
        if (lane < TZ && i0 + lane < nz && wstrip) EH = eta[(int64_t)(i0 + lane) * nx + x0 - 1];
        if (lane >= 32 && lane < 32 + TZ && i0 + lane - 32 < nz && estrip)
            EH = eta[(int64_t)(i0 + lane - 32) * nx + x0 + 64];
    }
    const float2 sxj = sx[jj], iw = isxe[jj], ie = isxe[jj + 1];
    const bool cj = conj != 0;
#pragma unroll
    for (int t = 0; t < TZ; ++t) {
        const int i = i0 + t;
        const float en = E[t], ec = E[t + 1], es = E[t + 2];
        float eww = __shfl_up(ec, 1, 64), eee = __shfl_down(ec, 1, 64);
        const float ehw = readlane_f(EH, t), ehe = readlane_f(EH, 32 + t);
        if (lane == 0) eww = ehw;
        if (lane == 63) eee = ehe;
        if (!hw) eww = 1.f;
        if (!he) eee = 1.f;
        const float2 szi = SZ[t], in_ = IZ[t], is_ = IZ[t + 1];
        Coef<float2> c;
        c.fw = cscale(cmul(szi, iw), hh * (ec + eww));
        c.fe = cscale(cmul(szi, ie), hh * (ec + eee));
        c.fn = cscale(cmul(sxj, in_), hh * (ec + en));
        c.fs = cscale(cmul(sxj, is_), hh * (ec + es));
        float2 mass = cscale(cmul(cmul(szi, sxj), K[t]), k0sq);
        if (shift != 0.f) {
            const float2 m2 = make_float2(mass.x - shift * mass.y, mass.y + shift * mass.x);
            mass = m2;
        }
        c.dg = csub(mass, cadd(cadd(c.fw, c.fe), cadd(c.fn, c.fs)));
        if (cj) {
            c.dg = cconj(c.dg);
            c.fw = cconj(c.fw);
            c.fe = cconj(c.fe);
            c.fn = cconj(c.fn);
            c.fs = cconj(c.fs);
        }
        const bool hn = i > 0, hs = i < nz - 1;
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const float2 uc = U[q][t + 1];
            float2 w = shfl_up1(uc), e = shfl_down1(uc);
            const float2 hwv = make_float2(readlane_f(HW.x, q * TZ + t), readlane_f(HW.y, q * TZ + t));
            const float2 hev = make_float2(readlane_f(HE.x, q * TZ + t), readlane_f(HE.y, q * TZ + t));
            if (lane == 0) w = hwv;
            if (lane == 63) e = hev;
            float2 acc = cmul(c.dg, uc);
            if (hw) cfma(acc, c.fw, w);
            if (he) cfma(acc, c.fe, e);
            if (hn) cfma(acc, c.fn, U[q][t]);
            if (hs) cfma(acc, c.fs, U[q][t + 2]);
            const int r = r0 + q;
            if (colok && r < nrhs && i < nz) {
                const int64_t o = (int64_t)r * n + (int64_t)i * nx + j;
                if (MODE == 1) acc = csub(b[o], acc);
                out[o] = acc;
            }
        }
    }
}
template <int RB, int TZ, int MODE>
__global__ __launch_bounds__(256) void k_apply_block(int nz, int nx, int64_t n, int nrhs,
                                                     const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                     const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                     const float* __restrict__ eta, const float2* __restrict__ kap,
                                                     float k0sq, float hh, float shift, int conj,
                                                     const float2* __restrict__ u, float2* __restrict__ out,
                                                     const float2* __restrict__ b, int xcd) {
    k_apply_block_body<RB, TZ, MODE>(hw_block(), nz, nx, n, nrhs, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, conj, u, out, b, xcd);
}

// Paired form of the block kernel (nx even): each lane owns two adjacent
// columns, so field and kappa rows move as 16-byte vectors, eta as 8-byte
// ones, and a wavefront covers 128 columns; the face between a lane's two
// columns is formed once, and only the outer neighbours (west of the first,
// east of the second column) come from the adjacent lanes.  The arithmetic
// per node, and its order, are those of k_apply_block.
template <int RB, int TZ>
__global__ __launch_bounds__(256) void k_apply_pair(int nz, int nx, int64_t n, int nrhs,
                                                    const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                    const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                    const float* __restrict__ eta, const float2* __restrict__ kap,
                                                    float k0sq, float hh, float shift, int conj,
                                                    const float2* __restrict__ u, float2* __restrict__ out, int xcd) {
    static_assert(RB * TZ <= 64, "halo columns are packed one value per lane");
    int bx, by, bz;
    xcd_tile(bx, by, bz, xcd != 0, xcd == 2);
    const int lane = threadIdx.x;
    const int x0 = bx * 128, c0 = x0 + 2 * lane;  // columns c0, c0 + 1
    const int i0 = __builtin_amdgcn_readfirstlane((by * 4 + threadIdx.y) * TZ);
    if (i0 >= nz) return;
    const int r0 = bz * RB;
    const bool colok = c0 < nx;  // nx is even: both columns or neither
    const int cc = colok ? c0 : nx - 2;  // lanes past the grid mirror the last pair (results dropped)
    const bool hw = cc > 0, he = cc + 1 < nx - 1;  // west of c0, east of c0 + 1
    const bool wstrip = x0 > 0, estrip = x0 + 128 < nx;
    float2 SZ[TZ], IZ[TZ + 1];
#pragma unroll
    for (int t = 0; t < TZ; ++t) SZ[t] = sz[min(i0 + t, nz - 1)];
#pragma unroll
    for (int t = 0; t <= TZ; ++t) IZ[t] = isze[min(i0 + t, nz)];
    const float4* u4 = reinterpret_cast<const float4*>(u);
    const float4* k4 = reinterpret_cast<const float4*>(kap);
    const float2* e2 = reinterpret_cast<const float2*>(eta);
    float4 U[RB][TZ + 2];
    float2 E[TZ + 2];
    float4 K[TZ];
#pragma unroll
    for (int t = 0; t < TZ + 2; ++t) {
        const int i = i0 - 1 + t;
        const bool rok = i >= 0 && i < nz;
#pragma unroll
        for (int q = 0; q < RB; ++q)
            U[q][t] = (rok && r0 + q < nrhs) ? u4[((int64_t)(r0 + q) * n + (int64_t)i * nx + cc) >> 1]
                                             : make_float4(0.f, 0.f, 0.f, 0.f);
        E[t] = rok ? e2[((int64_t)i * nx + cc) >> 1] : make_float2(1.f, 1.f);  // ghost eta 1
    }
#pragma unroll
    for (int t = 0; t < TZ; ++t)
        K[t] = i0 + t < nz ? k4[((int64_t)(i0 + t) * nx + cc) >> 1] : make_float4(0.f, 0.f, 0.f, 0.f);
    float2 HW = make_float2(0.f, 0.f), HE = make_float2(0.f, 0.f);
    float EH = 1.f;  // lanes 0..TZ-1: west eta of row t; lanes 32..32+TZ-1: east eta
    {
        const int q = lane / TZ, t = lane % TZ, i = i0 + t;
        if (lane < RB * TZ && i < nz && r0 + q < nrhs) {
            const int64_t o = (int64_t)(r0 + q) * n + (int64_t)i * nx;
            if (wstrip) HW = u[o + x0 - 1];
            if (estrip) HE = u[o + x0 + 128];
        }
        if (lane < TZ && i0 + lane < nz && wstrip) EH = eta[(int64_t)(i0 + lane) * nx + x0 - 1];
        if (lane >= 32 && lane < 32 + TZ && i0 + lane - 32 < nz && estrip)
            EH = eta[(int64_t)(i0 + lane - 32) * nx + x0 + 128];
    }
    const float2 sx0 = sx[cc], sx1 = sx[cc + 1], iw = isxe[cc], im = isxe[cc + 1], ie = isxe[cc + 2];
    const bool cj = conj != 0;
#pragma unroll
    for (int t = 0; t < TZ; ++t) {
        const int i = i0 + t;
        const float2 en = E[t], ec = E[t + 1], es = E[t + 2];
        float eww = __shfl_up(ec.y, 1, 64), eee = __shfl_down(ec.x, 1, 64);
        const float ehw = readlane_f(EH, t), ehe = readlane_f(EH, 32 + t);
        if (lane == 0) eww = ehw;
        if (lane == 63) eee = ehe;
        if (!hw) eww = 1.f;
        if (!he) eee = 1.f;
        const float2 szi = SZ[t], in_ = IZ[t], is_ = IZ[t + 1];
        // the two nodes' stencils, as node_coef forms them (the middle face once)
        const float2 fmid = cscale(cmul(szi, im), hh * (ec.x + ec.y));
        Coef<float2> a, b;
        a.fw = cscale(cmul(szi, iw), hh * (ec.x + eww));
        a.fe = fmid;
        a.fn = cscale(cmul(sx0, in_), hh * (ec.x + en.x));
        a.fs = cscale(cmul(sx0, is_), hh * (ec.x + es.x));
        b.fw = fmid;
        b.fe = cscale(cmul(szi, ie), hh * (ec.y + eee));
        b.fn = cscale(cmul(sx1, in_), hh * (ec.y + en.y));
        b.fs = cscale(cmul(sx1, is_), hh * (ec.y + es.y));
        float2 ma = cscale(cmul(cmul(szi, sx0), make_float2(K[t].x, K[t].y)), k0sq);
        float2 mb = cscale(cmul(cmul(szi, sx1), make_float2(K[t].z, K[t].w)), k0sq);
        if (shift != 0.f) {
            ma = make_float2(ma.x - shift * ma.y, ma.y + shift * ma.x);
            mb = make_float2(mb.x - shift * mb.y, mb.y + shift * mb.x);
        }
        a.dg = csub(ma, cadd(cadd(a.fw, a.fe), cadd(a.fn, a.fs)));
        b.dg = csub(mb, cadd(cadd(b.fw, b.fe), cadd(b.fn, b.fs)));
        if (cj) {
            a.dg = cconj(a.dg);
            a.fw = cconj(a.fw);
            a.fe = cconj(a.fe);
            a.fn = cconj(a.fn);
            a.fs = cconj(a.fs);
            b.dg = cconj(b.dg);
            b.fw = cconj(b.fw);
            b.fe = cconj(b.fe);
            b.fn = cconj(b.fn);
            b.fs = cconj(b.fs);
        }
        const bool hn = i > 0, hs = i < nz - 1;
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const float4 uc = U[q][t + 1], un = U[q][t], us = U[q][t + 2];
            const float2 u0 = make_float2(uc.x, uc.y), u1 = make_float2(uc.z, uc.w);
            float2 w = make_float2(__shfl_up(uc.z, 1, 64), __shfl_up(uc.w, 1, 64));
            float2 e = make_float2(__shfl_down(uc.x, 1, 64), __shfl_down(uc.y, 1, 64));
            const float2 hwv = make_float2(readlane_f(HW.x, q * TZ + t), readlane_f(HW.y, q * TZ + t));
            const float2 hev = make_float2(readlane_f(HE.x, q * TZ + t), readlane_f(HE.y, q * TZ + t));
            if (lane == 0) w = hwv;
            if (lane == 63) e = hev;
            float2 acc0 = cmul(a.dg, u0);
            if (hw) cfma(acc0, a.fw, w);
            cfma(acc0, a.fe, u1);
            if (hn) cfma(acc0, a.fn, make_float2(un.x, un.y));
            if (hs) cfma(acc0, a.fs, make_float2(us.x, us.y));
            float2 acc1 = cmul(b.dg, u1);
            cfma(acc1, b.fw, u0);
            if (he) cfma(acc1, b.fe, e);
            if (hn) cfma(acc1, b.fn, make_float2(un.z, un.w));
            if (hs) cfma(acc1, b.fs, make_float2(us.z, us.w));
            const int r = r0 + q;
            if (colok && r < nrhs && i < nz)
                reinterpret_cast<float4*>(out)[((int64_t)r * n + (int64_t)i * nx + c0) >> 1] =
                    make_float4(acc0.x, acc0.y, acc1.x, acc1.y);
        }
    }
}

// streaming-kernel shape (rows per run, RHS per group): tuning knobs read
// once from HELM_STENCIL_TZ and HELM_STENCIL_RB
static int stream_tz() {
    static const int t = [] {
        const char* v = std::getenv("HELM_STENCIL_TZ");
        const int x = v ? std::atoi(v) : 0;
        return x >= 2 && x <= 1024 ? x : 8;  // measured best on 576^2 x 11 (8 rows, 4 RHS)
    }();
    return t;
}
static int stream_rb() {
    static const int t = [] {
        const char* v = std::getenv("HELM_STENCIL_RB");
        const int x = v ? std::atoi(v) : 0;
        return (x == 1 || x == 2 || x == 4) ? x : 4;
    }();
    return t;
}

template <int RB>
static void launch_stream(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                          const float2* b, float shift, hipStream_t s) {
    const int tz = L.st_tz ? L.st_tz : stream_tz();
    const int runs = (L.nz + tz - 1) / tz;
    const dim3 block(64, 4), grid((L.nx + 63) / 64, (runs + 3) / 4, (nrhs + RB - 1) / RB);
    const float hh = (float)(0.5 / (L.h_m * L.h_m));
    if (b)
        hipLaunchKernelGGL((k_apply_stream<RB, 1>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, tz, L.sx32, L.sz32,
                           L.isxe32, L.isze32, L.eta32, L.kap32, (float)k0sq, hh, shift, (int)adjoint, u, out, b);
    else
        hipLaunchKernelGGL((k_apply_stream<RB, 0>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, tz, L.sx32, L.sz32,
                           L.isxe32, L.isze32, L.eta32, L.kap32, (float)k0sq, hh, shift, (int)adjoint, u, out,
                           (const float2*)nullptr);
}

// form before tuning: the block form (4-row runs unless HELM_STENCIL_TZ says
// otherwise); HELM_STENCIL_BLOCK=0 selects the streaming form
static bool stream_block() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_STENCIL_BLOCK");
        return !(v && std::atoi(v) == 0);
    }();
    return t;
}
static int block_tz() {
    static const int t = [] {
        const char* v = std::getenv("HELM_STENCIL_TZ");
        const int x = v ? std::atoi(v) : 0;
        return x >= 2 && x <= 16 ? x : 4;  // measured best or near it on 576^2 .. 2048^2 x 11
    }();
    return t;
}
// block form's workgroup order: HELM_STENCIL_XCD=1 gives each XCD a contiguous
// run of workgroups (neighbouring runs share their halo rows in one L2)
static int stencil_xcd() {
    static const int t = [] {
        const char* v = std::getenv("HELM_STENCIL_XCD");
        return v && std::atoi(v) == 1 ? 1 : 0;
    }();
    return t;
}

template <int RB, int TZ>
static void launch_block_tz(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                            const float2* b, float shift, hipStream_t s) {
    const int runs = (L.nz + TZ - 1) / TZ;
    const dim3 block(64, 4), grid((L.nx + 63) / 64, (runs + 3) / 4, (nrhs + RB - 1) / RB);
    const float hh = (float)(0.5 / (L.h_m * L.h_m));
    const int xcd = L.st_tz ? (int)L.st_xcd : stencil_xcd();
    if (b)
        hipLaunchKernelGGL((k_apply_block<RB, TZ, 1>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, L.sx32, L.sz32,
                           L.isxe32, L.isze32, L.eta32, L.kap32, (float)k0sq, hh, shift, (int)adjoint, u, out, b, xcd);
    else
        hipLaunchKernelGGL((k_apply_block<RB, TZ, 0>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, L.sx32, L.sz32,
                           L.isxe32, L.isze32, L.eta32, L.kap32, (float)k0sq, hh, shift, (int)adjoint, u, out,
                           (const float2*)nullptr, xcd);
}

template <int RB>
static void launch_block(const Level& L, int tz, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                         const float2* b, float shift, hipStream_t s) {
    if (tz <= 2)
        launch_block_tz<RB, 2>(L, k0sq, u, out, nrhs, adjoint, b, shift, s);
    else if (tz <= 4)
        launch_block_tz<RB, 4>(L, k0sq, u, out, nrhs, adjoint, b, shift, s);
    else if (tz <= 8 || RB * 16 > 64)  // a run's halo packs RB x TZ values into one wavefront
        launch_block_tz<RB, 8>(L, k0sq, u, out, nrhs, adjoint, b, shift, s);
    else
        launch_block_tz<RB, (RB * 16 <= 64 ? 16 : 8)>(L, k0sq, u, out, nrhs, adjoint, b, shift, s);
}

template <int RB, int TZ>
static void launch_pair_tz(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                           float shift, hipStream_t s) {
    const int runs = (L.nz + TZ - 1) / TZ;
    const dim3 block(64, 4), grid((L.nx + 127) / 128, (runs + 3) / 4, (nrhs + RB - 1) / RB);
    hipLaunchKernelGGL((k_apply_pair<RB, TZ>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, L.sx32, L.sz32, L.isxe32,
                       L.isze32, L.eta32, L.kap32, (float)k0sq, (float)(0.5 / (L.h_m * L.h_m)), shift, (int)adjoint,
                       u, out, (int)L.st_xcd);
}
template <int RB>
static void launch_pair(const Level& L, int tz, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                        float shift, hipStream_t s) {
    if (tz <= 2)
        launch_pair_tz<RB, 2>(L, k0sq, u, out, nrhs, adjoint, shift, s);
    else if (tz <= 4)
        launch_pair_tz<RB, 4>(L, k0sq, u, out, nrhs, adjoint, shift, s);
    else
        launch_pair_tz<RB, 8>(L, k0sq, u, out, nrhs, adjoint, shift, s);
}
// the paired form needs an even row length and 16-byte aligned fields and kappa
static bool pair_ok(const Level& L, const float2* u, const float2* out, const float2* b) {
    auto al = [](const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; };
    return !b && L.nx % 2 == 0 && L.nx >= 128 && al(u, 16) && al(out, 16) && al(L.kap32, 16) && al(L.eta32, 8);
}

// right-hand sides per workgroup: enough groups for >= ~2048 workgroups
static int rhs_per_group(const Level& L, int nrhs) {
    const int64_t bxy = (int64_t)((L.nx + 63) / 64) * ((L.nz + 3) / 4);
    const int64_t groups = std::max<int64_t>(1, std::min<int64_t>(nrhs, (2048 + bxy - 1) / bxy));
    return (int)((nrhs + groups - 1) / groups);
}

int32_t launch_apply32(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                       const float2* b, float shift, hipStream_t s) {
    if (L.op == HELM_OP_STAG8) {
        HELM_CHECK(shift == 0.0f, HELM_ERR_ARG, "the 8-tap operator has no shifted form");
        return launch_stag8_32(L, k0sq, u, out, nrhs, adjoint, b, s);
    }
    if (L.op == HELM_OP_COND8) {
        HELM_CHECK(shift == 0.0f, HELM_ERR_ARG, "the conditioned operator has no shifted form");
        return launch_cond8_32(L, u, out, nrhs, adjoint, b, s);
    }
    return launch_flux32(L, k0sq, u, out, nrhs, adjoint, b, shift, s);
}

int32_t launch_apply64(const Level& L, double k0sq, const double2* u, double2* out, int nrhs, bool adjoint,
                       const double2* b, hipStream_t s) {
    if (L.op == HELM_OP_STAG8) return launch_stag8_64(L, k0sq, u, out, nrhs, adjoint, b, b ? 1 : 0, s);
    if (L.op == HELM_OP_COND8) return launch_cond8_64(L, u, out, nrhs, adjoint, b, s);
    return launch_flux64(L, k0sq, u, out, nrhs, adjoint, b, s);
}

int32_t launch_flux32(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                      const float2* b, float shift, hipStream_t s) {
    if (L.nx >= 64) {  // streaming kernel (a strip needs a full wavefront of columns to pay off)
        const int rb = L.st_rb ? L.st_rb : stream_rb();
        if (L.st_tz && L.st_pair && pair_ok(L, u, out, b)) {
            switch (rb) {
                case 1: launch_pair<1>(L, L.st_tz, k0sq, u, out, nrhs, adjoint, shift, s); break;
                case 2: launch_pair<2>(L, L.st_tz, k0sq, u, out, nrhs, adjoint, shift, s); break;
                case 8: launch_pair<8>(L, L.st_tz, k0sq, u, out, nrhs, adjoint, shift, s); break;
                default: launch_pair<4>(L, L.st_tz, k0sq, u, out, nrhs, adjoint, shift, s); break;
            }
        } else if (L.st_tz ? L.st_block : stream_block()) {
            const int tz = L.st_tz ? L.st_tz : block_tz();
            switch (rb) {
                case 1: launch_block<1>(L, tz, k0sq, u, out, nrhs, adjoint, b, shift, s); break;
                case 2: launch_block<2>(L, tz, k0sq, u, out, nrhs, adjoint, b, shift, s); break;
                case 8: launch_block<8>(L, tz, k0sq, u, out, nrhs, adjoint, b, shift, s); break;
                default: launch_block<4>(L, tz, k0sq, u, out, nrhs, adjoint, b, shift, s); break;
            }
        } else {
            switch (rb) {
                case 1: launch_stream<1>(L, k0sq, u, out, nrhs, adjoint, b, shift, s); break;
                case 2: launch_stream<2>(L, k0sq, u, out, nrhs, adjoint, b, shift, s); break;
                default: launch_stream<4>(L, k0sq, u, out, nrhs, adjoint, b, shift, s); break;
            }
        }
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    const int rb = rhs_per_group(L, nrhs);
    dim3 block(64, 4), grid((L.nx + 63) / 64, (L.nz + 3) / 4, (nrhs + rb - 1) / rb);
    const float hh = (float)(0.5 / (L.h_m * L.h_m));
    if (b)
        hipLaunchKernelGGL((k_apply<float, float2, 1>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.sx32, L.sz32,
                           L.isxe32, L.isze32, L.eta32, L.kap32, (float)k0sq, hh, shift, (int)adjoint, u, out, b);
    else
        hipLaunchKernelGGL((k_apply<float, float2, 0>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.sx32, L.sz32,
                           L.isxe32, L.isze32, L.eta32, L.kap32, (float)k0sq, hh, shift, (int)adjoint, u, out,
                           (const float2*)nullptr);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

int32_t launch_flux64(const Level& L, double k0sq, const double2* u, double2* out, int nrhs, bool adjoint,
                      const double2* b, hipStream_t s) {
    const int rb = rhs_per_group(L, nrhs);
    dim3 block(64, 4), grid((L.nx + 63) / 64, (L.nz + 3) / 4, (nrhs + rb - 1) / rb);
    const double hh = 0.5 / (L.h_m * L.h_m);
    if (b)
        hipLaunchKernelGGL((k_apply<double, double2, 1>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.sx64,
                           L.sz64, L.isxe64, L.isze64, L.eta64, L.kap64, k0sq, hh, 0.0, (int)adjoint, u, out, b);
    else
        hipLaunchKernelGGL((k_apply<double, double2, 0>), grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.sx64,
                           L.sz64, L.isxe64, L.isze64, L.eta64, L.kap64, k0sq, hh, 0.0, (int)adjoint, u, out,
                           (const double2*)nullptr);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

// ---- ports ---------------------------------------------------------------
__global__ void k_fill_c128(int64_t count, double2* p) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = make_double2(0.0, 0.0);
}

// rhs[r][node_idx[q]] = -inv_h2 * sum_t w_node[t] * amp[r][port_of_tap[t]]
__global__ void k_inject(int n_nodes, int n_ports, int64_t n, int nrhs, const int* __restrict__ node_ptr,
                         const int* __restrict__ node_idx, const int* __restrict__ port_of_tap,
                         const double* __restrict__ w_node, const double2* __restrict__ amp, double2* rhs,
                         double inv_h2) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (q >= n_nodes || r >= nrhs) return;
    double2 acc = make_double2(0.0, 0.0);
    for (int t = node_ptr[q]; t < node_ptr[q + 1]; ++t) {
        double2 a = amp[(int64_t)r * n_ports + port_of_tap[t]];
        acc.x += w_node[t] * a.x;
        acc.y += w_node[t] * a.y;
    }
    rhs[(int64_t)r * n + node_idx[q]] = cscale(acc, -inv_h2);
}

template <typename C>
__global__ void k_sample(int n_ports, int64_t n, int nrhs, const int* __restrict__ port_ptr,
                         const int* __restrict__ node_of_tap, const double* __restrict__ w_port,
                         const C* __restrict__ field, double2* traces) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (p >= n_ports || r >= nrhs) return;
    double2 acc = make_double2(0.0, 0.0);
    for (int t = port_ptr[p]; t < port_ptr[p + 1]; ++t) {
        C v = field[(int64_t)r * n + node_of_tap[t]];
        acc.x += w_port[t] * (double)v.x;
        acc.y += w_port[t] * (double)v.y;
    }
    traces[(int64_t)r * n_ports + p] = acc;
}

// Time every streaming shape on the plan's grid and keep the fastest.
int32_t tune_stencil(helm_plan* p, int nrhs, hipStream_t s, double* best_us) {
    Level& L = p->levels[0];
    *best_us = 0.0;
    // the conditioned streaming kernel: the conditioned operator, and the S
    // operator once its medium has given it the kernel's operands
    const bool cond = L.op == HELM_OP_COND8 || (L.op == HELM_OP_STAG8 && L.cmass32 && L.ctab32);
    if (!cond && (L.op != HELM_OP_FLUX5 || L.nx < 64)) return HELM_OK;  // no streaming kernel there
    // Candidates are timed on inputs that are not in the 256 MB last-level
    // cache (Infinity Cache): the launches rotate through `nbuf` field pairs,
    // at least 512 MB in all, as the solve's applies meet fields that the
    // other kernels of the step have pushed out.  Timing one pair again and
    // again measured the cache and ranked shapes differently.
    const size_t stride = ((size_t)nrhs * L.n + 1) & ~(size_t)1;  // float2 per field batch, 16-byte aligned
    const size_t bytes = stride * sizeof(float2);
    const int nbuf = (int)std::min<size_t>(16, std::max<size_t>(2, ((size_t)512 << 20) / (2 * bytes) + 1));
    float2* pool = nullptr;
    if (hipMalloc((void**)&pool, 2 * nbuf * bytes) != hipSuccess) {
        (void)hipGetLastError();
        set_error("helm_autotune: scratch allocation failed");
        return HELM_ERR_ALLOC;
    }
    float2* u = pool;
    float2* out = pool + (size_t)nbuf * stride;
    int turn = 0;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int32_t st = HELM_OK;
    double best = 1e30;
    int btz = cond ? 32 : 4, brb = cond ? 1 : 4;
    if (hipMemsetAsync(pool, 0, 2 * nbuf * bytes, s) != hipSuccess) st = HELM_ERR_HIP;
    // rows per run x right-hand sides per group (the conditioned kernel's runs carry a 7-row warm-up
    // and march in groups of 8 rows; the intermediate row counts give grids that fill the device's
    // wavefront slots more evenly);
    // the flux stencil's streaming form is no longer a candidate (its block form measured faster on
    // every grid from 576^2 to 2048^2) but stays selectable with helm_set_stencil_shape
    const int ntz = cond ? 8 : 0;
    const int tzs_flux[8] = {4, 8, 16, 16, 16, 16, 16, 16}, tzs_cond[8] = {8, 16, 24, 32, 40, 48, 64, 128},
              rbs[4] = {1, 2, 4, 8};
    const int nrb = cond ? 2 : 4;  // the conditioned kernel: 1 or 2 RHS (4 fits only one wavefront per SIMD)
    const int* tzs = cond ? tzs_cond : tzs_flux;
    const double k0sq = p->k0 * p->k0;
    auto run = [&]() {
        const size_t o = (size_t)(turn++ % nbuf) * stride;
        return cond ? launch_cond8_32(L, u + o, out + o, nrhs, false, nullptr, s)
                    : launch_flux32(L, k0sq, u + o, out + o, nrhs, false, nullptr, 0.f, s);
    };
    // the flux stencil's block form: runs of 2, 4 or 8 rows loaded up front (16 was never the fastest),
    // and the paired form (two columns per lane; even rows of >= 128 nodes): runs of 2 or 4 rows; each
    // in the hardware's workgroup order, the XCD-contiguous one, and that with the RHS groups of a
    // run next to each other (one XCD's L2 then serves the run's coefficients to every group)
    const int nblk = cond ? 0 : 9, tzs_blk[3] = {2, 4, 8};
    const int npair = (!cond && pair_ok(L, u, out, nullptr)) ? 6 : 0, tzs_pair[2] = {2, 4};
    bool bblk = !cond, bpair = false;
    int bxcd = 0;
    const int reps = std::max(10, nbuf);  // launches timed per candidate
    for (int a = 0; a < ntz + nblk + npair && st == HELM_OK; ++a)
        for (int b = 0; b < nrb && st == HELM_OK; ++b) {
            const bool pair = a >= ntz + nblk;
            const bool blk = a >= ntz && !pair;
            L.st_tz = pair ? tzs_pair[(a - ntz - nblk) % 2] : (blk ? tzs_blk[(a - ntz) % 3] : tzs[a]);
            L.st_rb = rbs[b];
            L.st_block = blk || pair;  // the paired form falls back to the block form
            L.st_pair = pair;
            L.st_xcd = pair ? (a - ntz - nblk) / 2 : (blk ? (a - ntz) / 3 : 0);
            for (int w = 0; w < 3 && st == HELM_OK; ++w) st = run();
            (void)hipEventRecord(e0, s);
            for (int w = 0; w < reps && st == HELM_OK; ++w) st = run();
            (void)hipEventRecord(e1, s);
            float ms = 0.f;
            if (st == HELM_OK && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
                ms / reps < best) {
                best = ms / reps;
                btz = L.st_tz;
                brb = rbs[b];
                bblk = blk || pair;
                bxcd = L.st_xcd;
                bpair = pair;
            }
        }
    L.st_tz = btz;
    L.st_rb = brb;
    L.st_block = bblk;
    L.st_xcd = bxcd;
    L.st_pair = bpair;
    *best_us = best * 1e3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(pool);
    return st;
}

// segmented FP32 apply of the batched solve: w = A z_j into V_{j+1}'s slot,
// in the single-plan default shape (block form, 4 RHS per wavefront, 4-row
// runs, hardware order) on every plan
// (shift is a kernel argument, 0 at run time, as in the single-plan launch:
// a literal 0 would let the compiler fuse the mass term into the diagonal
// differently and round it differently)
__global__ __launch_bounds__(256) void k_seg_apply32(const SegDims* __restrict__ dims, int nseg,
                                                     const SegKrylov* __restrict__ seg, int j, int conj,
                                                     float shift) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegKrylov& a = seg[s];
    const int64_t vs = (int64_t)a.nrhs * a.n;
    k_apply_block_body<4, 4, 0>(B, a.nz, a.nx, a.n, a.nrhs, a.sx32, a.sz32, a.isxe32, a.isze32, a.eta32, a.kap32,
                                a.k0sq32, a.hh32, shift, conj, a.Z + j * vs, a.V + (j + 1) * vs, nullptr, 0);
}
void seg_apply32(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, int conj, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_apply32, dim3(blocks), dim3(64, 4), 0, s, dims, nseg, k, j, conj, 0.0f);
}

}  // namespace helm

using namespace helm;

extern "C" {

int32_t helm_apply(helm_plan* p, const void* u, void* out, int32_t nrhs, int32_t precision, int32_t adjoint,
                   void* stream) {
    HELM_CHECK(p && u && out, HELM_ERR_ARG, "helm_apply: null argument");
    HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_apply: call helm_set_medium first");
    HELM_CHECK(nrhs >= 1 && (int64_t)nrhs * p->levels[0].n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_apply: nrhs out of range");
    HELM_CHECK(u != out, HELM_ERR_ARG, "helm_apply: u and out must not alias");
    hipStream_t s = (hipStream_t)stream;
    const double k0sq = p->k0 * p->k0;
    if (precision == 32)
        return launch_apply32(p->levels[0], k0sq, (const float2*)u, (float2*)out, nrhs, adjoint != 0, nullptr,
                              0.0f, s);
    if (precision == 64)
        return launch_apply64(p->levels[0], k0sq, (const double2*)u, (double2*)out, nrhs, adjoint != 0, nullptr,
                              s);
    set_error("helm_apply: precision must be 32 or 64");
    return HELM_ERR_ARG;
}

int32_t helm_inject(helm_plan* p, int32_t n_ports, int32_t n_nodes, const int32_t* node_ptr,
                    const int32_t* node_idx, const int32_t* port_of_tap, const double* w_node, const void* amp,
                    void* rhs, int32_t nrhs, void* stream) {
    HELM_CHECK(p && node_ptr && node_idx && port_of_tap && w_node && amp && rhs, HELM_ERR_ARG,
               "helm_inject: null argument");
    HELM_CHECK(n_ports >= 1 && n_nodes >= 0 && n_nodes <= p->levels[0].n && nrhs >= 1 && nrhs <= 65535,
               HELM_ERR_ARG, "helm_inject: size out of range");
    hipStream_t s = (hipStream_t)stream;
    const Level& L = p->levels[0];
    int64_t count = (int64_t)nrhs * L.n;
    hipLaunchKernelGGL(k_fill_c128, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, count, (double2*)rhs);
    if (n_nodes > 0)
        hipLaunchKernelGGL(k_inject, dim3((n_nodes + 127) / 128, nrhs), dim3(128), 0, s, n_nodes, n_ports, L.n,
                           nrhs, node_ptr, node_idx, port_of_tap, w_node, (const double2*)amp, (double2*)rhs,
                           1.0 / (L.h_m * L.h_m));
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

int32_t helm_sample(helm_plan* p, int32_t n_ports, const int32_t* port_ptr, const int32_t* node_of_tap,
                    const double* w_port, const void* field, int32_t precision, void* traces, int32_t nrhs,
                    void* stream) {
    HELM_CHECK(p && port_ptr && node_of_tap && w_port && field && traces, HELM_ERR_ARG,
               "helm_sample: null argument");
    HELM_CHECK(n_ports >= 1 && nrhs >= 1 && nrhs <= 65535, HELM_ERR_ARG, "helm_sample: size out of range");
    hipStream_t s = (hipStream_t)stream;
    const Level& L = p->levels[0];
    dim3 grid((n_ports + 127) / 128, nrhs);
    if (precision == 64)
        hipLaunchKernelGGL(k_sample<double2>, grid, dim3(128), 0, s, n_ports, L.n, nrhs, port_ptr, node_of_tap,
                           w_port, (const double2*)field, (double2*)traces);
    else if (precision == 32)
        hipLaunchKernelGGL(k_sample<float2>, grid, dim3(128), 0, s, n_ports, L.n, nrhs, port_ptr, node_of_tap,
                           w_port, (const float2*)field, (double2*)traces);
    else {
        set_error("helm_sample: precision must be 32 or 64");
        return HELM_ERR_ARG;
    }
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

}  // extern "C"

// ---- material sensitivities for the adjoint gradient ---------------------
// For fields u_r and adjoint fields lam_r (A^H lam_r = P^T g_r):
//   s_kap[q] = sum_r conj(lam_rq) (dA/dkappa_q u_r)  = sum_r conj(lam_rq) k0^2 sz_i sx_j u_rq
//   s_eta[q] = sum_r lam_r^H (dA/deta_q) u_r
//            = sum_r sum_{edges e at q} stretch_e/(2h^2) conj(lam_q - lam_nb)(u_nb - u_q)
// (ghost neighbours: lam = u = 0).  The gradient of L(u(m)) is then
// dL/dm = -Re(sum_q s_eta[q] deta_q/dm + s_kap[q] dkappa_q/dm), the discrete
// adjoint-state formula of nsrno/dense_wave.py:240-256 (there obtained by
// autograd of -Re<lam, A u - s>) and of the Born Jacobian products of
// nsrno/theranostic/wave.py:372-388.
namespace helm {
__global__ void k_material_sens(int nz, int nx, int64_t n, int nrhs, const double2* __restrict__ sx,
                                const double2* __restrict__ sz, const double2* __restrict__ isxe,
                                const double2* __restrict__ isze, double k0sq, double hh,
                                const double2* __restrict__ u, const double2* __restrict__ lam,
                                double2* __restrict__ s_eta, double2* __restrict__ s_kap) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nz || j >= nx) return;
    const int64_t q = (int64_t)i * nx + j;
    const double2 szi = sz[i], sxj = sx[j];
    const double2 sw = cscale(cmul(szi, isxe[j]), hh), se = cscale(cmul(szi, isxe[j + 1]), hh);
    const double2 sn = cscale(cmul(sxj, isze[i]), hh), ss = cscale(cmul(sxj, isze[i + 1]), hh);
    const double2 mass = cscale(cmul(szi, sxj), k0sq);
    double2 ae = make_double2(0.0, 0.0), ak = make_double2(0.0, 0.0);
    for (int r = 0; r < nrhs; ++r) {
        const double2* ur = u + (int64_t)r * n;
        const double2* lr = lam + (int64_t)r * n;
        const double2 uq = ur[q], lq = cconj(lr[q]);
        cfma(ak, lq, cmul(mass, uq));
        auto edge = [&](bool inside, int64_t nb, double2 st) {
            const double2 un = inside ? ur[nb] : make_double2(0.0, 0.0);
            const double2 ln = inside ? cconj(lr[nb]) : make_double2(0.0, 0.0);
            cfma(ae, cmul(st, csub(lq, ln)), csub(un, uq));
        };
        edge(j > 0, q - 1, sw);
        edge(j < nx - 1, q + 1, se);
        edge(i > 0, q - nx, sn);
        edge(i < nz - 1, q + nx, ss);
    }
    s_eta[q] = ae;
    s_kap[q] = ak;
}
}  // namespace helm

extern "C" int32_t helm_material_sensitivity(helm_plan* p, const void* u, const void* lam, int32_t nrhs,
                                             void* s_eta, void* s_kap, void* stream) {
    HELM_CHECK(p && u && lam && s_eta && s_kap, HELM_ERR_ARG, "helm_material_sensitivity: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_FLUX5, HELM_ERR_STATE,
               "helm_material_sensitivity: flux-form (HELM_OP_FLUX5) plans only");
    HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_material_sensitivity: call helm_set_medium first");
    HELM_CHECK(nrhs >= 1 && (int64_t)nrhs * p->levels[0].n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_material_sensitivity: nrhs out of range");
    const Level& L = p->levels[0];
    dim3 block(64, 4), grid((L.nx + 63) / 64, (L.nz + 3) / 4);
    hipLaunchKernelGGL(k_material_sens, grid, block, 0, (hipStream_t)stream, L.nz, L.nx, L.n, nrhs, L.sx64, L.sz64,
                       L.isxe64, L.isze64, p->k0 * p->k0, 0.5 / (L.h_m * L.h_m), (const double2*)u,
                       (const double2*)lam, (double2*)s_eta, (double2*)s_kap);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

extern "C" int32_t helm_autotune(helm_plan* p, int32_t nrhs, helm_tuning* result, void* stream) {
    HELM_CHECK(p, HELM_ERR_ARG, "helm_autotune: null plan");
    HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_autotune: set the medium first");
    HELM_CHECK(nrhs >= 1 && nrhs <= 4096 && (int64_t)nrhs * p->levels[0].n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_autotune: nrhs out of range");
    hipStream_t s = (hipStream_t)stream;
    double st_us = 0.0, mg_us = 0.0;
    HELM_TRY(tune_stencil(p, nrhs, s, &st_us));
    HELM_TRY(tune_mg(p, nrhs, s, &mg_us));
    p->generation++;  // captured cycles hold the old launch shapes
    if (result) {
        const Level& L = p->levels[0];
        result->stencil_rows = L.st_tz;
        result->stencil_rhs = L.st_rb;
        const int xo[3] = {0, 1, 2};  // order -> form offset: hardware, XCD-contiguous, RHS groups fastest
        result->stencil_block = L.st_pair ? (L.st_xcd == 2 ? 6 : 3 + xo[L.st_xcd])
                                          : (L.st_block ? (L.st_xcd == 2 ? 5 : 1 + xo[L.st_xcd]) : 0);
        result->mg_workgroups = p->mg_wg_target;
        result->stencil_us = st_us;
        result->vcycle_us = mg_us;
    }
    return HELM_OK;
}

extern "C" int32_t helm_set_stencil_shape(helm_plan* p, int32_t rows, int32_t rhs, int32_t form) {
    HELM_CHECK(p, HELM_ERR_ARG, "helm_set_stencil_shape: null plan");
    Level& L = p->levels[0];
    if (rows == 0) {  // back to the built-in defaults
        L.st_tz = L.st_rb = 0;
        L.st_block = L.st_pair = false;
        L.st_xcd = 0;
    } else {
        HELM_CHECK(rhs == 1 || rhs == 2 || rhs == 4 || (rhs == 8 && form >= 1), HELM_ERR_ARG,
                   "helm_set_stencil_shape: rhs must be 1, 2, 4 (or 8 in the block and paired forms)");
        HELM_CHECK(form >= 0 && form <= 6, HELM_ERR_ARG, "helm_set_stencil_shape: form must be 0 .. 6");
        const bool pair = form == 3 || form == 4 || form == 6;
        const bool rows_ok = form == 0 ? rows >= 2 && rows <= 1024
                             : !pair   ? rows == 2 || rows == 4 || rows == 8 || rows == 16
                                       : rows == 2 || rows == 4 || rows == 8;
        HELM_CHECK(rows_ok && (form == 0 || rows * rhs <= 64), HELM_ERR_ARG,
                   "helm_set_stencil_shape: rows out of range for the form (rows x rhs <= 64)");
        L.st_tz = rows;
        L.st_rb = rhs;
        L.st_block = form >= 1;
        L.st_pair = pair;  // falls back to the block form where the paired one cannot run
        L.st_xcd = (form == 2 || form == 4) ? 1 : (form >= 5 ? 2 : 0);
    }
    p->generation++;  // captured cycles hold the old launch shapes
    return HELM_OK;
}
