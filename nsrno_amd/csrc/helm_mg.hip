// Complex-shifted-Laplacian multigrid preconditioner.
//
// M = A with kappa_c -> kappa_c (1 + i beta) (the same stretched-PML flux
// operator, damped by the shift), approximately inverted by one V-cycle:
//   * vertex-centred coarsening: coarse node J sits on fine node 2J, so an
//     axis of n nodes coarsens to (n + 1) / 2 nodes at twice the spacing; the
//     coarse operator is re-discretised from the same 1-D stretch profiles
//     (evaluated at the coarse nodes and edge midpoints) and from fully
//     weighted eta_rel / kappa_c;
//   * damped Jacobi smoothing, full-weighting restriction and bilinear
//     prolongation with the zero exterior ghosts of the fine operator;
//   * a fixed number of Jacobi sweeps on the coarsest level.
// It replaces the reference's FFT shifted-Laplacian multiplier
// 1 / (1 - q^2 - i shift) (nsrno/wave_operator.py:93-96, applied in
// wave_solver.py:246-247), which is exact only for a periodic
// constant-coefficient box and so does not see the PML or the medium.
// For adjoint solves every level uses conj(M) (M is complex symmetric).
//
// All kernels are HBM-bound: per sweep they read u, f (c64), the node
// coefficients (f32 eta + c64 kappa_c, shared by all right-hand sides) and
// write u; transfer kernels read one level and write the other.
#include <algorithm>
#include <cstdlib>
#include <cmath>

#include <hip/hip_ext.h>

#include "helm_coef.h"
#include "helm_internal.h"

namespace helm {

// Per-node complex Jacobi weight.  Local Fourier analysis of one sweep at a
// node with stencil (dg, fw, fe, fn, fs): a mode with cos(theta_x) = t1,
// cos(theta_z) = t2 is multiplied by g = 1 - w z,
//   z = 1 + t1 a + t2 b,   a = (fw + fe)/dg,  b = (fn + fs)/dg.
// |g| is convex in z, so its maximum over a rectangle of (t1, t2) sits at a
// corner.  The weight is picked from 33 x 32 candidates w = r e^{i phi}
// (phi in [-pi/2, pi/2], r in (0, 2]): on smoothing levels it minimises the
// worst |g| over the oscillatory modes (t1 <= 0 or t2 <= 0) among weights
// that amplify no mode by more than 10%; on the coarsest level (and where no
// weight meets the 10% cap) it minimises the worst |g| over all modes.  A
// real weight cannot do this for Helmholtz: modes below the wavenumber have
// Re z < 0, and the PML's complex stretch makes the stencil anisotropic, so
// plain damped Jacobi diverges there and on coarse grids.
__global__ __launch_bounds__(256) void k_node_weights(int nz, int nx, const float2* __restrict__ sx,
                                                      const float2* __restrict__ sz, const float2* __restrict__ isxe,
                                                      const float2* __restrict__ isze, const float* __restrict__ eta,
                                                      const float2* __restrict__ kap, float k0sq, float hh,
                                                      float shift, int coarsest, float2* __restrict__ wgt) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nz || j >= nx) return;
    Coef<float2> c;
    node_coef<float, float2>(i, j, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, false, c);
    const float dd = c.dg.x * c.dg.x + c.dg.y * c.dg.y;
    const float2 idg = make_float2(c.dg.x / dd, -c.dg.y / dd);
    const float2 a = cmul(cadd(c.fw, c.fe), idg), b = cmul(cadd(c.fn, c.fs), idg);
    // corners: 4 of the full square, then the oscillatory L-shaped region
    const float t1s[11] = {-1, -1, 1, 1, -1, -1, 0, 0, -1, 1, 1};
    const float t2s[11] = {-1, 1, -1, 1, -1, 1, -1, 1, 0, -1, 0};
    // |1 - w z|^2 = 1 + r (r |z|^2 - 2 Re(e^{i phi} z)) for w = r e^{i phi}: per angle the 11 projections
    // are formed once, and each radius costs two FMAs per mode
    float2 z[11];
    float m2[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) {
        z[k] = make_float2(1.f + t1s[k] * a.x + t2s[k] * b.x, t1s[k] * a.y + t2s[k] * b.y);
        m2[k] = z[k].x * z[k].x + z[k].y * z[k].y;
    }
    float best = 1e30f;
    float2 bw = make_float2(0.5f, 0.f);
    for (int ip = 0; ip <= 32; ++ip) {
        float sphi, cphi;
        sincosf(-1.5707963f + 3.1415927f * ip / 32.f, &sphi, &cphi);
        float q[11];  // -2 Re(e^{i phi} z_k)
#pragma unroll
        for (int k = 0; k < 11; ++k) q[k] = -2.f * (cphi * z[k].x - sphi * z[k].y);
        for (int ir = 1; ir <= 32; ++ir) {
            const float r = 2.f * ir / 32.f;
            float full = 0.f, hi = 0.f;
#pragma unroll
            for (int k = 0; k < 11; ++k) {
                const float g2 = fmaf(r, fmaf(r, m2[k], q[k]), 1.f);
                if (k < 4) full = fmaxf(full, g2);
                else hi = fmaxf(hi, g2);
            }
            hi = fmaxf(hi, 0.f);
            const float score = coarsest ? full : (full <= 1.21f ? hi : 1e3f + full);
            if (score < best) {
                best = score;
                bw = make_float2(r * cphi, r * sphi);
            }
        }
    }
    wgt[(int64_t)i * nx + j] = bw;
}

__global__ void k_fill_weight(int64_t n, float omega, float2* __restrict__ wgt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wgt[i] = make_float2(omega, 0.f);
}

// u_out = u_in + w (f - M u_in) / dg with the node's complex weight w
// (u_in == nullptr: zero start).  blockIdx.z selects a group of rb
// right-hand sides: large levels use few groups (coefficients are computed
// once and reused across the group), small levels one RHS per group (more
// workgroups in flight).
__global__ __launch_bounds__(256) void k_jacobi(int nz, int nx, int64_t n, int nrhs, int rb,
                                                const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                const float* __restrict__ eta, const float2* __restrict__ kap,
                                                float k0sq, float hh, float shift, int conj,
                                                const float2* __restrict__ wgt, const float2* __restrict__ u,
                                                const float2* __restrict__ f, float2* __restrict__ out) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nz || j >= nx) return;
    const int r0 = blockIdx.z * rb, r1 = min(nrhs, r0 + rb);
    Coef<float2> c;
    node_coef<float, float2>(i, j, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, conj != 0, c);
    const float dd = c.dg.x * c.dg.x + c.dg.y * c.dg.y;
    const int64_t idx = (int64_t)i * nx + j;
    const float2 wn = wgt[idx];
    const float2 idg = cmul(conj ? cconj(wn) : wn, make_float2(c.dg.x / dd, -c.dg.y / dd));
    const bool hw = j > 0, he = j < nx - 1, hn = i > 0, hs = i < nz - 1;
    for (int r = r0; r < r1; ++r) {
        const int64_t base = (int64_t)r * n;
        const float2 fv = f[base + idx];
        if (u == nullptr) {
            out[base + idx] = cmul(idg, fv);
            continue;
        }
        const float2* ur = u + base;
        const float2 uc = ur[idx];
        float2 acc = cmul(c.dg, uc);
        if (hw) cfma(acc, c.fw, ur[idx - 1]);
        if (he) cfma(acc, c.fe, ur[idx + 1]);
        if (hn) cfma(acc, c.fn, ur[idx - nx]);
        if (hs) cfma(acc, c.fs, ur[idx + nx]);
        float2 nu = uc;
        cfma(nu, idg, csub(fv, acc));
        out[base + idx] = nu;
    }
}

__device__ __forceinline__ float2 cc(float2 a, bool conj) { return conj ? cconj(a) : a; }

// ---- single-workgroup V-cycle over the small coarse levels ---------------
// The coarse levels hold too few nodes to fill the GPU, so launching one
// kernel per sweep is pure latency (tens of launches per V-cycle).  k_tail
// runs the whole V-cycle from level t down to the coarsest and back inside
// one kernel: one 1024-thread workgroup per right-hand side.  The iterate and
// its Jacobi partner of every tail level live in LDS (a stack, 16 B per node
// per level: 110 KB for 72^2 + 36^2 + 18^2); right-hand sides and residuals
// stay in global memory (L2); the stencil coefficients of the shifted
// operator and the Jacobi factor w/dg are precomputed per node at set-up
// (k_tail_coef), so a sweep is 6 coefficient loads, 5 LDS reads and one f load
// per node with no dependent chains.
// first level handled by k_tail: the largest level with at most TAIL_NODES
// nodes (its LDS stack, 16 B per node over the tail levels, must fit 160 KB)
constexpr int64_t TAIL_NODES = 80 * 80;  // LDS capacity bound
constexpr int64_t TAIL_NODES_DEFAULT = 40 * 40;  // measured best on 576^2 (11 RHS)

// tuning knobs, read once from the environment: HELM_MG_WG_TARGET (workgroups
// a grid-level launch aims for) and HELM_MG_TAIL_NODES (largest level the
// LDS-resident tail takes)
static int64_t env_int(const char* name, int64_t dflt, int64_t lo, int64_t hi) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    const long long x = std::atoll(v);
    return x < lo ? lo : (x > hi ? hi : x);
}
static int64_t wg_target() {
    static const int64_t t = env_int("HELM_MG_WG_TARGET", 2048, 64, 1 << 20);
    return t;
}
// fused level kernels for pre = post = 2 (HELM_MG_FUSED=0 selects the four-kernel path)
static bool fused_levels() {
    static const bool f = env_int("HELM_MG_FUSED", 1, 0, 1) != 0;
    return f;
}
// XCD-aware tile order in the fused kernels, per kernel and level
// (HELM_XCD_ORDER bits: 1 pre-smoothing on the finest level, 2 pre-smoothing
// on coarser levels, 4 post-smoothing on the finest level, 8 post-smoothing
// on coarser levels; 15: all; 0: the hardware order everywhere).  Default 15
// with the right-hand-side groups fastest: at 576^2 x 11 the V-cycle takes
// the same time in every order (126.9-128.6 us; the kernels are bound by
// latency), and this one moves the fewest HBM bytes (finest pre-smoothing
// 55 MB per launch against 105 MB in the hardware order).
static int xcd_order() {
    static const int t = (int)env_int("HELM_XCD_ORDER", 15, 0, 15);
    return t;
}
// in the XCD-aware order the right-hand-side groups of a tile follow each
// other (one XCD, one L2 for the tile's coefficients); HELM_XCD_RHS_FAST=0:
// the groups run slowest
static int xcd_rhs_fast() {
    static const int t = (int)env_int("HELM_XCD_RHS_FAST", 1, 0, 1);
    return t;
}
// kernel argument: 0 hardware order, 1 XCD-contiguous, 2 XCD-contiguous with
// the right-hand-side group fastest
static int xcd_pre(int l) { return ((xcd_order() >> (l == 0 ? 0 : 1)) & 1) * (1 + xcd_rhs_fast()); }
static int xcd_post(int l) { return ((xcd_order() >> (l == 0 ? 2 : 3)) & 1) * (1 + xcd_rhs_fast()); }
// post-smoothing kernel variant: without prefetching (default) it fits three
// waves per SIMD, which measured 25% faster at 576^2 x 11 than the prefetching
// variant's two (HELM_MG_POST_PREFETCH=1)
static bool post2_prefetch() {
    static const bool t = env_int("HELM_MG_POST_PREFETCH", 0, 0, 1) != 0;
    return t;
}
// pre-smoothing kernel variant (HELM_MG_PRE_PREFETCH=1: next right-hand side prefetched)
static bool pre2_prefetch() {
    static const bool t = env_int("HELM_MG_PRE_PREFETCH", 0, 0, 1) != 0;
    return t;
}
// stencil table on the finest grid level too (HELM_MG_TABLE0=1)
static bool table0() {
    static const bool t = env_int("HELM_MG_TABLE0", 0, 0, 1) != 0;
    return t;
}
static int64_t tail_nodes() {
    static const int64_t t = env_int("HELM_MG_TAIL_NODES", TAIL_NODES_DEFAULT, 16, TAIL_NODES);
    return t;
}



__global__ void k_tail_coef(int nz, int nx, const float2* __restrict__ sx, const float2* __restrict__ sz,
                            const float2* __restrict__ isxe, const float2* __restrict__ isze,
                            const float* __restrict__ eta, const float2* __restrict__ kap, float k0sq, float hh,
                            float shift, const float2* __restrict__ wgt, float2* __restrict__ cf) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nz * nx) return;
    const int i = (int)(q / nx), j = (int)(q % nx);
    Coef<float2> c;
    node_coef<float, float2>(i, j, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, false, c);
    const float dd = c.dg.x * c.dg.x + c.dg.y * c.dg.y;
    float2* o = cf + q * 6;
    o[0] = cmul(wgt[q], make_float2(c.dg.x / dd, -c.dg.y / dd));
    o[1] = c.dg;
    o[2] = c.fw;
    o[3] = c.fe;
    o[4] = c.fn;
    o[5] = c.fs;
}

// M u at node q with u in LDS
__device__ __forceinline__ float2 tail_mu(const LevelView& V, int64_t q, const float2* cf, const float2* u,
                                          bool conj) {
    const int i = (int)(q / V.nx), j = (int)(q % V.nx);
    float2 acc = cmul(cc(cf[1], conj), u[q]);
    if (j > 0) cfma(acc, cc(cf[2], conj), u[q - 1]);
    if (j < V.nx - 1) cfma(acc, cc(cf[3], conj), u[q + 1]);
    if (i > 0) cfma(acc, cc(cf[4], conj), u[q - V.nx]);
    if (i < V.nz - 1) cfma(acc, cc(cf[5], conj), u[q + V.nx]);
    return acc;
}

// sweeps of damped Jacobi from a zero start (zero_start) or from u, ending
// in u; t is the LDS partner buffer
__device__ void tail_sweeps(const LevelView& V, int sweeps, bool zero_start, bool conj, const float2* f, float2* u,
                            float2* t) {
    const float2* src = zero_start ? nullptr : u;
    if (!zero_start && sweeps % 2 == 1) {  // odd count from u: start from a copy in t
        for (int64_t q = threadIdx.x; q < V.n; q += blockDim.x) t[q] = u[q];
        __syncthreads();
        src = t;
    }
    for (int k = 0; k < sweeps; ++k) {
        float2* dst = zero_start ? ((sweeps - k) % 2 == 1 ? u : t) : (src == u ? t : u);
        for (int64_t q = threadIdx.x; q < V.n; q += blockDim.x) {
            const float2* cf = V.cf + q * 6;
            const float2 fv = f[q];
            if (src == nullptr) {
                dst[q] = cmul(cc(cf[0], conj), fv);
            } else {
                float2 nu = src[q];
                cfma(nu, cc(cf[0], conj), csub(fv, tail_mu(V, q, cf, src, conj)));
                dst[q] = nu;
            }
        }
        __syncthreads();
        src = dst;
    }
    if (zero_start && sweeps == 0) {
        for (int64_t q = threadIdx.x; q < V.n; q += blockDim.x) u[q] = make_float2(0.f, 0.f);
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_tail(const LevelView* __restrict__ views, int t, int last, int conj_i,
                                               int pre, int post, int coarse, const float2* __restrict__ f_top,
                                               float2* __restrict__ u_top, const int* __restrict__ act) {
    extern __shared__ float2 lds[];
    const int r = blockIdx.x;
    if (act && !act[r]) return;
    const bool conj = conj_i != 0;
    int64_t off[16];
    {
        int64_t o = 0;
        for (int l = t; l <= last; ++l) {
            off[l - t] = o;
            o += 2 * views[l].n;
        }
    }
    // down: smooth, residual, restrict
    for (int l = t; l <= last; ++l) {
        const LevelView& V = views[l];
        const float2* f = (l == t ? f_top : V.f) + (int64_t)r * V.n;
        float2* u = lds + off[l - t];
        float2* tt = u + V.n;
        if (l == last) {
            tail_sweeps(V, coarse, true, conj, f, u, tt);
            break;
        }
        tail_sweeps(V, pre, true, conj, f, u, tt);
        float2* res = V.r + (int64_t)r * V.n;
        for (int64_t q = threadIdx.x; q < V.n; q += blockDim.x) res[q] = csub(f[q], tail_mu(V, q, V.cf + q * 6, u, conj));
        __syncthreads();
        const LevelView& C = views[l + 1];
        float2* fc = C.f + (int64_t)r * C.n;
        for (int64_t q = threadIdx.x; q < C.n; q += blockDim.x) {
            const int I = (int)(q / C.nx), J = (int)(q % C.nx);
            float2 acc = make_float2(0.f, 0.f);
            for (int a = -1; a <= 1; ++a) {
                const int i = 2 * I + a;
                if (i < 0 || i >= V.nz) continue;
                for (int b = -1; b <= 1; ++b) {
                    const int j = 2 * J + b;
                    if (j < 0 || j >= V.nx) continue;
                    const float w = (a == 0 ? 0.5f : 0.25f) * (b == 0 ? 0.5f : 0.25f);
                    const float2 v = res[(int64_t)i * V.nx + j];
                    acc.x += w * v.x;
                    acc.y += w * v.y;
                }
            }
            fc[q] = acc;
        }
        __syncthreads();
    }
    // up: prolong-add, post-smooth
    for (int l = last - 1; l >= t; --l) {
        const LevelView& V = views[l];
        const LevelView& C = views[l + 1];
        const float2* f = (l == t ? f_top : V.f) + (int64_t)r * V.n;
        float2* u = lds + off[l - t];
        const float2* uc = lds + off[l + 1 - t];
        for (int64_t q = threadIdx.x; q < V.n; q += blockDim.x) {
            const int i = (int)(q / V.nx), j = (int)(q % V.nx);
            const int I0 = i >> 1, J0 = j >> 1;
            float2 v = make_float2(0.f, 0.f);
            for (int di = 0; di <= (i & 1); ++di)
                for (int dj = 0; dj <= (j & 1); ++dj) {
                    const int I = I0 + di, J = J0 + dj;
                    if (I < C.nz && J < C.nx) {
                        const float w = ((i & 1) ? 0.5f : 1.f) * ((j & 1) ? 0.5f : 1.f);
                        const float2 cv = uc[(int64_t)I * C.nx + J];
                        v.x += w * cv.x;
                        v.y += w * cv.y;
                    }
                }
            u[q] = cadd(u[q], v);
        }
        __syncthreads();
        tail_sweeps(V, post, false, conj, f, u, u + V.n);
    }
    // result of level t to global memory
    const LevelView& T = views[t];
    for (int64_t q = threadIdx.x; q < T.n; q += blockDim.x) u_top[(int64_t)r * T.n + q] = lds[q];
}

// ---- the same V-cycle with every operand in LDS --------------------------
// k_tail's sweeps each wait on global (L2) loads of the six coefficients and
// f per node, so at 37^2 with 10 coarse sweeps the launch is ~20 us of load
// latency for under 1 MB of data.  This form copies the coefficients of every
// tail level (as six planes, so a wavefront reads consecutive words) and the
// top level's f into LDS once, keeps the residual and the restricted
// right-hand sides there too, and touches global memory only for that copy
// and the final u: 72 B of LDS per node over the tail levels (u, partner, f,
// six coefficients), used whenever that fits 160 KB.
// row of node q on a level at most TAIL_NODES large: floor((q + 1/2) / nx) in
// FP32 (the rounding error, under 1e-3, stays below the 1/(2 nx) margin)
__device__ __forceinline__ int tail_row(int q, float inx) { return (int)(((float)q + 0.5f) * inx); }

__device__ __forceinline__ float2 tail_mu_p(int nz, int nx, int n, int q, int i, const float2* cf, const float2* u,
                                            bool conj) {
    const int j = q - i * nx;
    float2 acc = cmul(cc(cf[n + q], conj), u[q]);
    if (j > 0) cfma(acc, cc(cf[2 * n + q], conj), u[q - 1]);
    if (j < nx - 1) cfma(acc, cc(cf[3 * n + q], conj), u[q + 1]);
    if (i > 0) cfma(acc, cc(cf[4 * n + q], conj), u[q - nx]);
    if (i < nz - 1) cfma(acc, cc(cf[5 * n + q], conj), u[q + nx]);
    return acc;
}

// k_tail's sweep sequence (same ping-pong order) on plane-layout coefficients.
// One workgroup runs a right-hand side's tail on one CU, so a sweep is bound
// by that CU's LDS bandwidth: each thread keeps the stencil, Jacobi factor and
// f of its nodes (up to TS_NPT) in registers for all the sweeps of the call,
// and a sweep reads only the iterate's five stencil values from LDS (12 words
// per node before).  Same operations in the same order as tail_mu_p.
constexpr int TS_NPT = 3;
__device__ void tail_sweeps_p(int nz, int nx, int n, const float2* cf, int sweeps, bool zero_start, bool conj,
                              const float2* f, float2* u, float2* t) {
    const float inx = 1.f / (float)nx;
    const int bd = (int)blockDim.x;
    const float2* src = zero_start ? nullptr : u;
    if (!zero_start && sweeps % 2 == 1) {
        for (int q = threadIdx.x; q < n; q += bd) t[q] = u[q];
        __syncthreads();
        src = t;
    }
    if (n <= TS_NPT * bd) {
        float2 w[TS_NPT], dg[TS_NPT], fw[TS_NPT], fe[TS_NPT], fn[TS_NPT], fs[TS_NPT], fv[TS_NPT];
        int qs[TS_NPT];
        bool hw[TS_NPT], he[TS_NPT], hn[TS_NPT], hs[TS_NPT];
#pragma unroll
        for (int k = 0; k < TS_NPT; ++k) {
            const int q = threadIdx.x + k * bd;
            qs[k] = q < n ? q : -1;
            const int qc = q < n ? q : 0;
            const int i = tail_row(qc, inx), j = qc - i * nx;
            hw[k] = j > 0;
            he[k] = j < nx - 1;
            hn[k] = i > 0;
            hs[k] = i < nz - 1;
            w[k] = cc(cf[qc], conj);
            dg[k] = cc(cf[n + qc], conj);
            fw[k] = cc(cf[2 * n + qc], conj);
            fe[k] = cc(cf[3 * n + qc], conj);
            fn[k] = cc(cf[4 * n + qc], conj);
            fs[k] = cc(cf[5 * n + qc], conj);
            fv[k] = f[qc];
        }
        for (int k = 0; k < sweeps; ++k) {
            float2* dst = zero_start ? ((sweeps - k) % 2 == 1 ? u : t) : (src == u ? t : u);
#pragma unroll
            for (int e = 0; e < TS_NPT; ++e) {
                const int q = qs[e];
                if (q < 0) continue;
                if (src == nullptr) {
                    dst[q] = cmul(w[e], fv[e]);
                } else {
                    float2 acc = cmul(dg[e], src[q]);
                    if (hw[e]) cfma(acc, fw[e], src[q - 1]);
                    if (he[e]) cfma(acc, fe[e], src[q + 1]);
                    if (hn[e]) cfma(acc, fn[e], src[q - nx]);
                    if (hs[e]) cfma(acc, fs[e], src[q + nx]);
                    float2 nu = src[q];
                    cfma(nu, w[e], csub(fv[e], acc));
                    dst[q] = nu;
                }
            }
            __syncthreads();
            src = dst;
        }
    } else {
        for (int k = 0; k < sweeps; ++k) {
            float2* dst = zero_start ? ((sweeps - k) % 2 == 1 ? u : t) : (src == u ? t : u);
            for (int q = threadIdx.x; q < n; q += bd) {
                const float2 fq = f[q];
                if (src == nullptr) {
                    dst[q] = cmul(cc(cf[q], conj), fq);
                } else {
                    float2 nu = src[q];
                    cfma(nu, cc(cf[q], conj), csub(fq, tail_mu_p(nz, nx, n, q, tail_row(q, inx), cf, src, conj)));
                    dst[q] = nu;
                }
            }
            __syncthreads();
            src = dst;
        }
    }
    if (zero_start && sweeps == 0) {
        for (int q = threadIdx.x; q < n; q += bd) u[q] = make_float2(0.f, 0.f);
        __syncthreads();
    }
}

__device__ __forceinline__ void k_tail_lds_body(const VBlock& B, const LevelView* __restrict__ views, int t, int last, int conj_i,
                                                   int pre, int post, int coarse, const float2* __restrict__ f_top,
                                                   float2* __restrict__ u_top, const int* __restrict__ act) {
    extern __shared__ float2 lds[];
    const int r = B.x;
    if (act && !act[r]) return;
    const bool conj = conj_i != 0;
    int off[16];  // level l's block: u, partner, f, 6 coefficient planes (9 n float2)
    {
        int o = 0;
        for (int l = t; l <= last; ++l) {
            off[l - t] = o;
            o += 9 * (int)views[l].n;
        }
    }
    // Copy f and the coefficient tables ([node][6] -> six planes) into LDS.
    // Every load of a batch is issued before its first LDS store (a plain
    // copy loop waits for each load): f's loads first, then each level's
    // table in batches of TAIL_CB values per thread, f's stores last.  The
    // tail levels hold at most 160 KB / 72 B = 2275 nodes, so one batch of
    // TAIL_FPT f values and of TAIL_CB table values per thread covers them.
    {
        constexpr int TAIL_FPT = 3, TAIL_CB = 14;
        const int bd = (int)blockDim.x, tid = (int)threadIdx.x;
        const int nt = (int)views[t].n;
        const float2* fs = f_top + (int64_t)r * nt;
        float2 fv[TAIL_FPT];
#pragma unroll
        for (int i = 0; i < TAIL_FPT; ++i) {
            const int k = i * bd + tid;
            fv[i] = ld_global(fs + (k < nt ? k : 0));
        }
        int o = 0;
        for (int l = t; l <= last; ++l) {
            const int n = (int)views[l].n;
            const float2* src = views[l].cf;
            float2* cf = lds + o + 3 * n;
            for (int base = 0; base < 6 * n; base += TAIL_CB * bd) {
                float2 v[TAIL_CB];
#pragma unroll
                for (int i = 0; i < TAIL_CB; ++i) {
                    const int k = base + i * bd + tid;
                    v[i] = ld_global(src + (k < 6 * n ? k : 0));
                }
#pragma unroll
                for (int i = 0; i < TAIL_CB; ++i) {
                    // out-of-range values go to u[0] of the level, which the
                    // sweeps write before they read it: a store behind a
                    // branch would let the compiler sink its load to it
                    const int k = base + i * bd + tid;
                    const int q = k / 6, c = k - 6 * q;
                    *(k < 6 * n ? cf + c * n + q : lds + o) = v[i];
                }
            }
            o += 9 * n;
        }
        float2* f = lds + 2 * nt;  // level t's block starts at offset 0
#pragma unroll
        for (int i = 0; i < TAIL_FPT; ++i) {
            const int k = i * bd + tid;
            if (k < nt) f[k] = fv[i];
        }
        for (int k = TAIL_FPT * bd + tid; k < nt; k += bd) f[k] = ld_global(fs + k);
    }
    __syncthreads();
    // down: smooth, residual (into the partner buffer), restrict into the coarse level's f
    for (int l = t; l <= last; ++l) {
        const LevelView& V = views[l];
        const int n = (int)V.n;
        float2* u = lds + off[l - t];
        float2* tt = u + n;
        const float2* f = u + 2 * n;
        const float2* cf = u + 3 * n;
        if (l == last) {
            tail_sweeps_p(V.nz, V.nx, n, cf, coarse, true, conj, f, u, tt);
            break;
        }
        tail_sweeps_p(V.nz, V.nx, n, cf, pre, true, conj, f, u, tt);
        const float inx = 1.f / (float)V.nx;
        for (int q = threadIdx.x; q < n; q += blockDim.x)
            tt[q] = csub(f[q], tail_mu_p(V.nz, V.nx, n, q, tail_row(q, inx), cf, u, conj));
        __syncthreads();
        const LevelView& C = views[l + 1];
        const int nc = (int)C.n;
        const float inxc = 1.f / (float)C.nx;
        float2* fc = lds + off[l + 1 - t] + 2 * nc;
        for (int q = threadIdx.x; q < nc; q += blockDim.x) {
            const int I = tail_row(q, inxc), J = q - I * C.nx;
            float2 acc = make_float2(0.f, 0.f);
            for (int a = -1; a <= 1; ++a) {
                const int i = 2 * I + a;
                if (i < 0 || i >= V.nz) continue;
                for (int b = -1; b <= 1; ++b) {
                    const int j = 2 * J + b;
                    if (j < 0 || j >= V.nx) continue;
                    const float w = (a == 0 ? 0.5f : 0.25f) * (b == 0 ? 0.5f : 0.25f);
                    const float2 v = tt[i * V.nx + j];
                    acc.x += w * v.x;
                    acc.y += w * v.y;
                }
            }
            fc[q] = acc;
        }
        __syncthreads();
    }
    // up: prolong-add, post-smooth
    for (int l = last - 1; l >= t; --l) {
        const LevelView& V = views[l];
        const LevelView& C = views[l + 1];
        const int n = (int)V.n;
        float2* u = lds + off[l - t];
        const float2* f = u + 2 * n;
        const float2* cf = u + 3 * n;
        const float2* uc = lds + off[l + 1 - t];
        const float inx = 1.f / (float)V.nx;
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            const int i = tail_row(q, inx), j = q - i * V.nx;
            const int I0 = i >> 1, J0 = j >> 1;
            float2 v = make_float2(0.f, 0.f);
            for (int di = 0; di <= (i & 1); ++di)
                for (int dj = 0; dj <= (j & 1); ++dj) {
                    const int I = I0 + di, J = J0 + dj;
                    if (I < C.nz && J < C.nx) {
                        const float w = ((i & 1) ? 0.5f : 1.f) * ((j & 1) ? 0.5f : 1.f);
                        const float2 cv = uc[I * C.nx + J];
                        v.x += w * cv.x;
                        v.y += w * cv.y;
                    }
                }
            u[q] = cadd(u[q], v);
        }
        __syncthreads();
        tail_sweeps_p(V.nz, V.nx, n, cf, post, false, conj, f, u, u + n);
    }
    const int n = (int)views[t].n;
    for (int q = threadIdx.x; q < n; q += blockDim.x) u_top[(int64_t)r * n + q] = lds[off[0] + q];
}
__global__ __launch_bounds__(1024) void k_tail_lds(const LevelView* __restrict__ views, int t, int last, int conj_i,
                                                   int pre, int post, int coarse, const float2* __restrict__ f_top,
                                                   float2* __restrict__ u_top, const int* __restrict__ act) {
    k_tail_lds_body(hw_block(), views, t, last, conj_i, pre, post, coarse, f_top, u_top, act);
}

// ---- fused grid-level V-cycle kernels ----------------------------------
// A level above the LDS tail costs four passes over its fields per V-cycle:
//   pre     u = S(S(0))         one pass: the first sweep from zero is the
//                               pointwise u1 = (w/dg) f, so the second sweep
//                               needs only f and w/dg at the 5 stencil nodes
//   down    f_c = R (f - M u)   residual on a fine tile in LDS, then full
//                               weighting; the fine residual never reaches HBM
//   up      t = S(u + P u_c)    bilinear prolongation evaluated at the 5
//                               stencil nodes inside the first post-sweep
//   post    u = S(t)            the remaining sweeps
// Bytes per node per RHS: 16 + 18 + 26 + 24 = 84 (the unfused sequence of
// 2 + 1 + 1 + 1 + 2 kernels moves 140).

// per-node Jacobi factor of the shifted operator: wid = w / dg
__global__ void k_node_wid(int nz, int nx, const float2* __restrict__ sx, const float2* __restrict__ sz,
                           const float2* __restrict__ isxe, const float2* __restrict__ isze,
                           const float* __restrict__ eta, const float2* __restrict__ kap, float k0sq, float hh,
                           float shift, const float2* __restrict__ wgt, float2* __restrict__ wid) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nz || j >= nx) return;
    Coef<float2> c;
    node_coef<float, float2>(i, j, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, false, c);
    const float dd = c.dg.x * c.dg.x + c.dg.y * c.dg.y;
    const int64_t q = (int64_t)i * nx + j;
    wid[q] = cmul(wgt[q], make_float2(c.dg.x / dd, -c.dg.y / dd));
}

// The grid-level kernels work on 16 x 64 tiles (block 64 x 4, four rows per
// thread).  The field a sweep reads is staged in LDS on the tile plus a
// one-node halo with the operator's zero ghosts already in place, so the
// stencil needs no bounds tests and every global load is independent and
// coalesced; each thread keeps its four nodes' stencil in registers for all
// right-hand sides of its group.
constexpr int MT_TX = 64, MT_TZ = 16, MT_RPT = MT_TZ / 4;
constexpr int MT_SW = MT_TX + 2, MT_SH = MT_TZ + 2, MT_NS = MT_SW * MT_SH;
constexpr int MT_SPT = (MT_NS + 255) / 256;  // staged nodes per thread

struct TileCoef {
    Coef<float2> c[MT_RPT];
    float2 w[MT_RPT];  // Jacobi factor w/dg
    bool in[MT_RPT];
};

__device__ __forceinline__ void tile_coef(int z0, int x0, int nz, int nx, const float2* __restrict__ sx,
                                          const float2* __restrict__ sz, const float2* __restrict__ isxe,
                                          const float2* __restrict__ isze, const float* __restrict__ eta,
                                          const float2* __restrict__ kap, float k0sq, float hh, float shift, bool cj,
                                          const float2* __restrict__ wid, TileCoef& t) {
#pragma unroll
    for (int k = 0; k < MT_RPT; ++k) {
        const int i = z0 + threadIdx.y + 4 * k, j = x0 + threadIdx.x;
        t.in[k] = i < nz && j < nx;
        if (t.in[k]) {
            node_coef<float, float2>(i, j, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, cj, t.c[k]);
            t.w[k] = cc(wid[(int64_t)i * nx + j], cj);
        }
    }
}

// LDS index of tile node (row, col) of the staged block
__device__ __forceinline__ int st_idx(int row, int col) { return (row + 1) * MT_SW + col + 1; }

// M v at the thread's k-th node from the staged field
__device__ __forceinline__ float2 tile_mv(const Coef<float2>& c, const float2* sv, int row, int col) {
    const int q = st_idx(row, col);
    float2 acc = cmul(c.dg, sv[q]);
    cfma(acc, c.fw, sv[q - 1]);
    cfma(acc, c.fe, sv[q + 1]);
    cfma(acc, c.fn, sv[q - MT_SW]);
    cfma(acc, c.fs, sv[q + MT_SW]);
    return acc;
}

// Pipelining shared by the tile kernels: the loads of right-hand side r+1
// (its staged halo values and the f values of the thread's own nodes) are
// issued into registers before RHS r is computed from LDS, and the first
// RHS's loads before the stencil coefficients are formed, so each workgroup
// waits for about one memory round trip instead of two per RHS.
struct TileIdx {  // node indices within one field (a level has < 2^31 nodes)
    int st[MT_SPT];   // each staged node (-1: ghost or not staged)
    int own[MT_RPT];  // each own node (-1: outside the grid)
};

__device__ __forceinline__ void tile_idx(int z0, int x0, int nz, int nx, TileIdx& t) {
    const int tid = threadIdx.y * 64 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < MT_SPT; ++q) {
        const int e = tid + 256 * q;
        t.st[q] = -1;
        if (e < MT_NS) {
            const int gi = z0 - 1 + e / MT_SW, gj = x0 - 1 + e % MT_SW;
            if (gi >= 0 && gi < nz && gj >= 0 && gj < nx) t.st[q] = gi * nx + gj;
        }
    }
#pragma unroll
    for (int k = 0; k < MT_RPT; ++k) {
        const int i = z0 + threadIdx.y + 4 * k, j = x0 + threadIdx.x;
        t.own[k] = (i < nz && j < nx) ? i * nx + j : -1;
    }
}

__device__ __forceinline__ void tile_fetch(const TileIdx& t, const float2* __restrict__ src,
                                           const float2* __restrict__ fr, float2 (&st)[MT_SPT],
                                           float2 (&fv)[MT_RPT]) {
    const float2 zero = make_float2(0.f, 0.f);
#pragma unroll
    for (int q = 0; q < MT_SPT; ++q) st[q] = t.st[q] >= 0 ? src[t.st[q]] : zero;
#pragma unroll
    for (int k = 0; k < MT_RPT; ++k) fv[k] = t.own[k] >= 0 ? fr[t.own[k]] : zero;
}

__device__ __forceinline__ void tile_store_lds(const float2 (&st)[MT_SPT], float2* sv) {
    const int tid = threadIdx.y * 64 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < MT_SPT; ++q) {
        const int e = tid + 256 * q;
        if (e < MT_NS) sv[e] = st[q];
    }
}

// one damped-Jacobi sweep out = u + w (f - M u)
__global__ __launch_bounds__(256) void k_mgt_jacobi(int nz, int nx, int64_t n, int nrhs, int rb,
                                                    const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                    const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                    const float* __restrict__ eta, const float2* __restrict__ kap,
                                                    float k0sq, float hh, float shift, int conj,
                                                    const float2* __restrict__ wid, const float2* __restrict__ u,
                                                    const float2* __restrict__ f, float2* __restrict__ out) {
    __shared__ float2 sv[MT_NS];
    const int z0 = blockIdx.y * MT_TZ, x0 = blockIdx.x * MT_TX;
    const int r0 = blockIdx.z * rb, r1 = min(nrhs, r0 + rb);
    TileIdx ix;
    tile_idx(z0, x0, nz, nx, ix);
    float2 st[MT_SPT], fv[MT_RPT];
    tile_fetch(ix, u + (int64_t)r0 * n, f + (int64_t)r0 * n, st, fv);
    TileCoef t;
    tile_coef(z0, x0, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, conj != 0, wid, t);
    for (int r = r0; r < r1; ++r) {
        __syncthreads();  // the previous RHS is done with sv
        tile_store_lds(st, sv);
        float2 fc[MT_RPT];
#pragma unroll
        for (int k = 0; k < MT_RPT; ++k) fc[k] = fv[k];
        if (r + 1 < r1) tile_fetch(ix, u + (int64_t)(r + 1) * n, f + (int64_t)(r + 1) * n, st, fv);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MT_RPT; ++k) {
            if (!t.in[k]) continue;
            const int row = threadIdx.y + 4 * k, col = threadIdx.x;
            float2 v = sv[st_idx(row, col)];
            cfma(v, t.w[k], csub(fc[k], tile_mv(t.c[k], sv, row, col)));
            out[(int64_t)r * n + ix.own[k]] = v;
        }
    }
}

// out = S(S(0)) = u1 + w (f - M u1) with u1 = (w/dg) f staged on the halo
__global__ __launch_bounds__(256) void k_mgt_pre2(int nz, int nx, int64_t n, int nrhs, int rb,
                                                  const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                  const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                  const float* __restrict__ eta, const float2* __restrict__ kap,
                                                  float k0sq, float hh, float shift, int conj,
                                                  const float2* __restrict__ wid, const float2* __restrict__ f,
                                                  float2* __restrict__ out) {
    __shared__ float2 sv[MT_NS];
    const int z0 = blockIdx.y * MT_TZ, x0 = blockIdx.x * MT_TX;
    const bool cj = conj != 0;
    const int r0 = blockIdx.z * rb, r1 = min(nrhs, r0 + rb);
    TileIdx ix;
    tile_idx(z0, x0, nz, nx, ix);
    float2 st[MT_SPT], fv[MT_RPT];
    tile_fetch(ix, f + (int64_t)r0 * n, f + (int64_t)r0 * n, st, fv);
    float2 ws[MT_SPT];  // w/dg at the staged nodes (0 outside the grid)
#pragma unroll
    for (int q = 0; q < MT_SPT; ++q) ws[q] = ix.st[q] >= 0 ? cc(wid[ix.st[q]], cj) : make_float2(0.f, 0.f);
    TileCoef t;
    tile_coef(z0, x0, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, cj, wid, t);
    for (int r = r0; r < r1; ++r) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MT_SPT; ++q) st[q] = cmul(ws[q], st[q]);  // u1 = (w/dg) f
        tile_store_lds(st, sv);
        float2 fc[MT_RPT];
#pragma unroll
        for (int k = 0; k < MT_RPT; ++k) fc[k] = fv[k];
        if (r + 1 < r1) tile_fetch(ix, f + (int64_t)(r + 1) * n, f + (int64_t)(r + 1) * n, st, fv);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MT_RPT; ++k) {
            if (!t.in[k]) continue;
            const int row = threadIdx.y + 4 * k, col = threadIdx.x;
            float2 v = sv[st_idx(row, col)];
            cfma(v, t.w[k], csub(fc[k], tile_mv(t.c[k], sv, row, col)));
            out[(int64_t)r * n + ix.own[k]] = v;
        }
    }
}

// out = S(u + P u_c): the coarse patch under the tile and its halo is staged,
// the corrected field v = u + P u_c is formed on the halo in LDS (bilinear
// weights without branches), then one sweep; out must not alias u
constexpr int MP_CH = MT_TZ / 2 + 3, MP_CW = MT_TX / 2 + 3, MP_NC = MP_CH * MP_CW;
constexpr int MP_CPT = (MP_NC + 255) / 256;
__global__ __launch_bounds__(256) void k_mgt_prolong_smooth(int nz, int nx, int64_t n, int nzc, int nxc, int nrhs,
                                                            int rb, const float2* __restrict__ sx,
                                                            const float2* __restrict__ sz,
                                                            const float2* __restrict__ isxe,
                                                            const float2* __restrict__ isze,
                                                            const float* __restrict__ eta,
                                                            const float2* __restrict__ kap, float k0sq, float hh,
                                                            float shift, int conj, const float2* __restrict__ wid,
                                                            const float2* __restrict__ uc,
                                                            const float2* __restrict__ u,
                                                            const float2* __restrict__ f, float2* __restrict__ out) {
    __shared__ float2 sv[MT_NS];
    __shared__ float2 scv[MP_NC];
    const int z0 = blockIdx.y * MT_TZ, x0 = blockIdx.x * MT_TX, tid = threadIdx.y * 64 + threadIdx.x;
    const int cz0 = z0 / 2 - 1, cx0 = x0 / 2 - 1;
    const int64_t ncn = (int64_t)nzc * nxc;
    const int r0 = blockIdx.z * rb, r1 = min(nrhs, r0 + rb);
    TileIdx ix;
    tile_idx(z0, x0, nz, nx, ix);
    int cix[MP_CPT];
#pragma unroll
    for (int q = 0; q < MP_CPT; ++q) {
        const int e = tid + 256 * q;
        cix[q] = -1;
        if (e < MP_NC) {
            const int I = cz0 + e / MP_CW, J = cx0 + e % MP_CW;
            if (I >= 0 && I < nzc && J >= 0 && J < nxc) cix[q] = I * nxc + J;
        }
    }
    float2 st[MT_SPT], fv[MT_RPT], cs[MP_CPT];
    auto fetch = [&](int r) {
        tile_fetch(ix, u + (int64_t)r * n, f + (int64_t)r * n, st, fv);
#pragma unroll
        for (int q = 0; q < MP_CPT; ++q) cs[q] = cix[q] >= 0 ? uc[(int64_t)r * ncn + cix[q]] : make_float2(0.f, 0.f);
    };
    fetch(r0);
    TileCoef t;
    tile_coef(z0, x0, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, conj != 0, wid, t);
    for (int r = r0; r < r1; ++r) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MP_CPT; ++q) {
            const int e = tid + 256 * q;
            if (e < MP_NC) scv[e] = cs[q];
        }
        __syncthreads();
        // v = u + P u_c on the staged halo
#pragma unroll
        for (int q = 0; q < MT_SPT; ++q) {
            const int e = tid + 256 * q;
            if (e < MT_NS) {
                float2 v = make_float2(0.f, 0.f);
                if (ix.st[q] >= 0) {
                    const int gi = z0 - 1 + e / MT_SW, gj = x0 - 1 + e % MT_SW;
                    const int I = (gi >> 1) - cz0, J = (gj >> 1) - cx0;
                    const float wz1 = (gi & 1) ? 0.5f : 0.f, wz0 = 1.f - wz1;
                    const float wx1 = (gj & 1) ? 0.5f : 0.f, wx0 = 1.f - wx1;
                    const float2* c0 = scv + I * MP_CW + J;
                    v = st[q];
                    v.x += wz0 * (wx0 * c0[0].x + wx1 * c0[1].x) + wz1 * (wx0 * c0[MP_CW].x + wx1 * c0[MP_CW + 1].x);
                    v.y += wz0 * (wx0 * c0[0].y + wx1 * c0[1].y) + wz1 * (wx0 * c0[MP_CW].y + wx1 * c0[MP_CW + 1].y);
                }
                sv[e] = v;
            }
        }
        float2 fc[MT_RPT];
#pragma unroll
        for (int k = 0; k < MT_RPT; ++k) fc[k] = fv[k];
        if (r + 1 < r1) fetch(r + 1);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MT_RPT; ++k) {
            if (!t.in[k]) continue;
            const int row = threadIdx.y + 4 * k, col = threadIdx.x;
            float2 v = sv[st_idx(row, col)];
            cfma(v, t.w[k], csub(fc[k], tile_mv(t.c[k], sv, row, col)));
            out[(int64_t)r * n + ix.own[k]] = v;
        }
    }
}

// f_c = R (f - M u): one workgroup per 8 x 32 coarse tile; u is staged on the
// (2*8+3) x (2*32+3) fine nodes around it, the residual on the
// (2*8+1) x (2*32+1) nodes under it is formed in LDS and fully weighted
constexpr int RR_TZ = 8, RR_TX = 32;
constexpr int RR_FZ = 2 * RR_TZ + 1, RR_FX = 2 * RR_TX + 1;
constexpr int RR_NF = RR_FZ * RR_FX;
constexpr int RR_PT = (RR_NF + 255) / 256;
constexpr int RR_UZ = RR_FZ + 2, RR_UX = RR_FX + 2, RR_NU = RR_UZ * RR_UX;
__global__ __launch_bounds__(256) void k_mgt_restrict_res(int nzf, int nxf, int nzc, int nxc, int nrhs, int rb,
                                                          const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                          const float2* __restrict__ isxe,
                                                          const float2* __restrict__ isze,
                                                          const float* __restrict__ eta,
                                                          const float2* __restrict__ kap, float k0sq, float hh,
                                                          float shift, int conj, const float2* __restrict__ u,
                                                          const float2* __restrict__ f, float2* __restrict__ fcoarse) {
    __shared__ float2 su[RR_NU];
    __shared__ float2 rs[RR_NF];
    const int tid = threadIdx.y * RR_TX + threadIdx.x;
    const int I0 = blockIdx.y * RR_TZ, J0 = blockIdx.x * RR_TX;
    const int fi0 = 2 * I0 - 1, fj0 = 2 * J0 - 1;  // first residual node
    const int64_t nf = (int64_t)nzf * nxf, nc = (int64_t)nzc * nxc;
    const bool cj = conj != 0;
    Coef<float2> c[RR_PT];
    int64_t node[RR_PT];
#pragma unroll
    for (int q = 0; q < RR_PT; ++q) {
        const int e = tid + 256 * q;
        node[q] = -1;
        if (e < RR_NF) {
            const int fi = fi0 + e / RR_FX, fj = fj0 + e % RR_FX;
            if (fi >= 0 && fi < nzf && fj >= 0 && fj < nxf) {
                node[q] = (int64_t)fi * nxf + fj;
                node_coef<float, float2>(fi, fj, nzf, nxf, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, cj, c[q]);
            }
        }
    }
    const int I = I0 + threadIdx.y, J = J0 + threadIdx.x;
    const int r0 = blockIdx.z * rb, r1 = min(nrhs, r0 + rb);
    for (int r = r0; r < r1; ++r) {
        const float2* ur = u + (int64_t)r * nf;
        const float2* fr = f + (int64_t)r * nf;
        __syncthreads();
        for (int e = tid; e < RR_NU; e += 256) {
            const int gi = fi0 - 1 + e / RR_UX, gj = fj0 - 1 + e % RR_UX;
            su[e] = (gi >= 0 && gi < nzf && gj >= 0 && gj < nxf) ? ur[(int64_t)gi * nxf + gj] : make_float2(0.f, 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RR_PT; ++q) {
            const int e = tid + 256 * q;
            if (e < RR_NF) {
                float2 v = make_float2(0.f, 0.f);
                if (node[q] >= 0) {
                    const int qu = (e / RR_FX + 1) * RR_UX + e % RR_FX + 1;
                    float2 acc = cmul(c[q].dg, su[qu]);
                    cfma(acc, c[q].fw, su[qu - 1]);
                    cfma(acc, c[q].fe, su[qu + 1]);
                    cfma(acc, c[q].fn, su[qu - RR_UX]);
                    cfma(acc, c[q].fs, su[qu + RR_UX]);
                    v = csub(fr[node[q]], acc);
                }
                rs[e] = v;
            }
        }
        __syncthreads();
        if (I < nzc && J < nxc) {
            const float2* base = rs + (2 * threadIdx.y + 1) * RR_FX + 2 * threadIdx.x + 1;
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll
            for (int a = -1; a <= 1; ++a)
#pragma unroll
                for (int b = -1; b <= 1; ++b) {
                    const float w = (a == 0 ? 0.5f : 0.25f) * (b == 0 ? 0.5f : 0.25f);
                    const float2 v = base[a * RR_FX + b];
                    acc.x += w * v.x;
                    acc.y += w * v.y;
                }
            fcoarse[(int64_t)r * nc + (int64_t)I * nxc + J] = acc;
        }
    }
}

// Stencil and Jacobi factor of a level node: from the level's precomputed
// table (levels >= 1: [w/dg, dg, fw, fe, fn, fs] per node, TAB) or formed from
// the medium (level 0, where the table would cost more bytes than the
// formation saves).  The fused level kernels load the operands of all of a
// thread's nodes first (level_raw_load, addresses clamped, no branches) and
// form the stencils after (level_raw_form), so the set-up waits for memory
// once rather than once or twice per node.
template <bool TAB>
struct LevelRaw;
template <>
struct LevelRaw<true> {
    float4 a, b, c;  // [w/dg, dg], [fw, fe], [fn, fs]
};
template <>
struct LevelRaw<false> {
    NodeRaw m;
    float2 w;
};
// ok false: an in-range node is loaded and the result must not be used
template <bool TAB>
__device__ __forceinline__ void level_raw_load(int gi, int gj, bool ok, int nz, int nx,
                                               const float2* __restrict__ tcf, const float2* __restrict__ sx,
                                               const float2* __restrict__ sz, const float2* __restrict__ isxe,
                                               const float2* __restrict__ isze, const float* __restrict__ eta,
                                               const float2* __restrict__ kap, const float2* __restrict__ wid,
                                               LevelRaw<TAB>& r) {
    const int64_t node = ok ? (int64_t)gi * nx + gj : 0;
    if constexpr (TAB) {
        const float4* t = reinterpret_cast<const float4*>(tcf + node * 6);  // 48 B rows of a 256 B aligned table
        r.a = t[0];
        r.b = t[1];
        r.c = t[2];
    } else {
        node_raw_load(gi, gj, ok, nz, nx, sx, sz, isxe, isze, eta, kap, r.m);
        r.w = wid[node];
    }
}
template <bool TAB>
__device__ __forceinline__ void level_raw_form(const LevelRaw<TAB>& r, int gi, int gj, int nz, int nx, float k0sq,
                                               float hh, float shift, bool cj, Coef<float2>& c, float2& w) {
    if constexpr (TAB) {
        w = cc(make_float2(r.a.x, r.a.y), cj);
        c.dg = cc(make_float2(r.a.z, r.a.w), cj);
        c.fw = cc(make_float2(r.b.x, r.b.y), cj);
        c.fe = cc(make_float2(r.b.z, r.b.w), cj);
        c.fn = cc(make_float2(r.c.x, r.c.y), cj);
        c.fs = cc(make_float2(r.c.z, r.c.w), cj);
    } else {
        node_raw_form(r.m, gi, gj, nz, nx, k0sq, hh, shift, cj, c);
        w = cc(r.w, cj);
    }
}
// Jacobi factor alone (node clamped by the caller)
template <bool TAB>
__device__ __forceinline__ float2 level_w_load(int64_t node, const float2* __restrict__ tcf,
                                               const float2* __restrict__ wid) {
    return TAB ? tcf[node * 6] : wid[node];
}

// false when act (per RHS, nullable) marks every RHS of [r0, r1) inactive: the
// workgroup then has nothing to do (a converged RHS's multigrid output is
// unused).  Lane k loads the flag of RHS r0 + k and a ballot combines them:
// one memory round trip, where a loop that stops at the first active RHS
// waits for each flag in turn.  (Testing the flags after the kernel's set-up
// loads instead measured slower, V-cycle 131 against 127.5 us: the extra live
// state split those loads into several rounds.)  Groups larger than a
// wavefront are taken as active.
__device__ __forceinline__ bool group_active(const int* __restrict__ act, int r0, int r1) {
    if (!act) return true;
    const int k = __lane_id(), n = r1 - r0;
    const int v = act[r0 + (k < n ? k : 0)];
    return n > 64 || __ballot(k < n && v != 0) != 0;
}

// ---- fused level kernels (pre = post = 2) ---------------------------------
// k_mgt_pre2_rr: u2 = S(S(0)) and f_c = R (f - M u2) in one pass.  Works on
// the 8 x 32 coarse tile of k_mgt_restrict_res: f is staged on the fine
// region that the residual's full weighting needs plus a 2-node halo
// (21 x 69), u1 = (w/dg) f there, u2 on the region plus a 1-node halo
// (19 x 67), the residual on the 17 x 65 region; each fine node of the 16 x 64
// block [2 I0, 2 I0 + 16) x [2 J0, 2 J0 + 64) writes its u2 (these blocks
// partition the fine grid).  Each thread keeps the stencil of its u2-region
// nodes in registers and uses it for both the sweep and the residual.
constexpr int PR_RZ = 2 * RR_TZ + 1, PR_RX = 2 * RR_TX + 1;  // residual region 17 x 65
constexpr int PR_2Z = PR_RZ + 2, PR_2X = PR_RX + 2, PR_2N = PR_2Z * PR_2X;  // u2 region 19 x 67
constexpr int PR_1Z = PR_RZ + 4, PR_1X = PR_RX + 4, PR_1N = PR_1Z * PR_1X;  // f, u1 region 21 x 69
constexpr int PR_1PT = (PR_1N + 255) / 256, PR_2PT = (PR_2N + 255) / 256;
template <bool PF, bool TAB>
__device__ __forceinline__ void k_mgt_pre2_rr_body(const VBlock& B, int nzf, int nxf, int nzc, int nxc, int nrhs,
                                                               int rb,
                                                     const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                     const float2* __restrict__ isxe,
                                                     const float2* __restrict__ isze, const float* __restrict__ eta,
                                                     const float2* __restrict__ kap, float k0sq, float hh,
                                                     float shift, int conj, const float2* __restrict__ wid,
                                                     const float2* __restrict__ tcf, const float2* __restrict__ f,
                                                     const double* __restrict__ fscale, float2* __restrict__ u,
                                                     float2* __restrict__ fcoarse, int xcd,
                                                     const int* __restrict__ act) {
    __shared__ float2 sf[PR_1N];
    __shared__ float2 s1[PR_1N];  // u1, then the residual (17 x 65) once u2 is formed
    __shared__ float2 s2[PR_2N];
    __shared__ float2 sw[PR_1N];  // w/dg on the f / u1 region (in LDS: as registers it spilled)
    const int tid = threadIdx.y * RR_TX + threadIdx.x;
    int bx, by, bz;
    xcd_tile(B, bx, by, bz, xcd != 0, xcd == 2);
    if (!group_active(act, bz * rb, min(nrhs, bz * rb + rb))) return;
    const int I0 = by * RR_TZ, J0 = bx * RR_TX;
    const int fi0 = 2 * I0 - 1, fj0 = 2 * J0 - 1;  // first residual node
    const int64_t nf = (int64_t)nzf * nxf, nc = (int64_t)nzc * nxc;
    const bool cj = conj != 0;
    // Set-up: the Jacobi factors of the f / u1 region nodes staged by this
    // thread and the stencils of its u2 region nodes, every load issued before
    // the first is used, then the first right-hand side's f with them
    int i1[PR_1PT];
    float2 w1[PR_1PT];
#pragma unroll
    for (int q = 0; q < PR_1PT; ++q) {
        const int e = tid + 256 * q;
        const int gi = fi0 - 2 + e / PR_1X, gj = fj0 - 2 + e % PR_1X;
        const bool ok = e < PR_1N && gi >= 0 && gi < nzf && gj >= 0 && gj < nxf;
        i1[q] = ok ? gi * nxf + gj : -1;
        w1[q] = level_w_load<TAB>(ok ? i1[q] : 0, tcf, wid);
    }
    int i2[PR_2PT];
    LevelRaw<TAB> raw[PR_2PT];
#pragma unroll
    for (int q = 0; q < PR_2PT; ++q) {
        const int e = tid + 256 * q;
        const int gi = fi0 - 1 + e / PR_2X, gj = fj0 - 1 + e % PR_2X;
        const bool ok = e < PR_2N && gi >= 0 && gi < nzf && gj >= 0 && gj < nxf;
        i2[q] = ok ? gi * nxf + gj : -1;
        level_raw_load<TAB>(gi, gj, ok, nzf, nxf, tcf, sx, sz, isxe, isze, eta, kap, wid, raw[q]);
    }
    const int r0 = bz * rb, r1 = min(nrhs, r0 + rb);
    float2 fq[PR_1PT];
    auto fetch = [&](int r) {
#pragma unroll
        for (int q = 0; q < PR_1PT; ++q) fq[q] = i1[q] >= 0 ? f[(int64_t)r * nf + i1[q]] : make_float2(0.f, 0.f);
        if (fscale) {
            const float fs = (float)fscale[r];
#pragma unroll
            for (int q = 0; q < PR_1PT; ++q) fq[q] = make_float2(fq[q].x * fs, fq[q].y * fs);
        }
    };
    fetch(r0);
#pragma unroll
    for (int q = 0; q < PR_1PT; ++q) {
        const int e = tid + 256 * q;
        if (e < PR_1N) sw[e] = i1[q] >= 0 ? cc(w1[q], cj) : make_float2(0.f, 0.f);  // 0 on ghosts
    }
    Coef<float2> c2[PR_2PT];
    float2 w2[PR_2PT];
#pragma unroll
    for (int q = 0; q < PR_2PT; ++q) {
        const int e = tid + 256 * q;
        level_raw_form<TAB>(raw[q], fi0 - 1 + e / PR_2X, fj0 - 1 + e % PR_2X, nzf, nxf, k0sq, hh, shift, cj, c2[q],
                            w2[q]);
    }
    for (int r = r0; r < r1; ++r) {
        if (!PF && r > r0) fetch(r);
        lds_barrier();  // the previous RHS is done with the LDS arrays
#pragma unroll
        for (int q = 0; q < PR_1PT; ++q) {
            const int e = tid + 256 * q;
            if (e < PR_1N) {
                sf[e] = fq[q];
                s1[e] = cmul(sw[e], fq[q]);  // u1 = (w/dg) f (0 on ghosts)
            }
        }
        if (PF && r + 1 < r1) fetch(r + 1);
        lds_barrier();
        // u2 = u1 + (w/dg) (f - M u1) on the u2 region
        float2 v2[PR_2PT];
#pragma unroll
        for (int q = 0; q < PR_2PT; ++q) {
            const int e = tid + 256 * q;
            v2[q] = make_float2(0.f, 0.f);
            if (e < PR_2N && i2[q] >= 0) {
                const int k = (e / PR_2X + 1) * PR_1X + e % PR_2X + 1;  // same node in the 21 x 69 layout
                float2 acc = cmul(c2[q].dg, s1[k]);
                cfma(acc, c2[q].fw, s1[k - 1]);
                cfma(acc, c2[q].fe, s1[k + 1]);
                cfma(acc, c2[q].fn, s1[k - PR_1X]);
                cfma(acc, c2[q].fs, s1[k + PR_1X]);
                float2 v = s1[k];
                cfma(v, w2[q], csub(sf[k], acc));
                v2[q] = v;
            }
            if (e < PR_2N) s2[e] = v2[q];
        }
        lds_barrier();
        // residual f - M u2 on the 17 x 65 region (into s1) and the owned u2 outputs
#pragma unroll
        for (int q = 0; q < PR_2PT; ++q) {
            const int e = tid + 256 * q;
            if (e >= PR_2N) continue;
            const int row = e / PR_2X, col = e % PR_2X;  // region-2 coordinates
            const int gi = fi0 - 1 + row, gj = fj0 - 1 + col;
            if (row >= 1 && row <= PR_RZ && col >= 1 && col <= PR_RX) {
                float2 res = make_float2(0.f, 0.f);
                if (i2[q] >= 0) {
                    float2 acc = cmul(c2[q].dg, s2[e]);
                    cfma(acc, c2[q].fw, s2[e - 1]);
                    cfma(acc, c2[q].fe, s2[e + 1]);
                    cfma(acc, c2[q].fn, s2[e - PR_2X]);
                    cfma(acc, c2[q].fs, s2[e + PR_2X]);
                    res = csub(sf[(row + 1) * PR_1X + col + 1], acc);
                }
                s1[(row - 1) * PR_RX + col - 1] = res;
            }
            if (i2[q] >= 0 && gi >= 2 * I0 && gi < 2 * I0 + 2 * RR_TZ && gj >= 2 * J0 && gj < 2 * J0 + 2 * RR_TX)
                u[(int64_t)r * nf + i2[q]] = v2[q];
        }
        lds_barrier();
        const int I = I0 + threadIdx.y, J = J0 + threadIdx.x;
        if (I < nzc && J < nxc) {
            const float2* base = s1 + (2 * threadIdx.y + 1) * PR_RX + 2 * threadIdx.x + 1;
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll
            for (int a = -1; a <= 1; ++a)
#pragma unroll
                for (int b = -1; b <= 1; ++b) {
                    const float w = (a == 0 ? 0.5f : 0.25f) * (b == 0 ? 0.5f : 0.25f);
                    const float2 v = base[a * PR_RX + b];
                    acc.x += w * v.x;
                    acc.y += w * v.y;
                }
            fcoarse[(int64_t)r * nc + (int64_t)I * nxc + J] = acc;
        }
    }
}
template <bool PF, bool TAB>
__global__ __launch_bounds__(256, 3) void k_mgt_pre2_rr(int nzf, int nxf, int nzc, int nxc, int nrhs,
                                                               int rb,
                                                     const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                     const float2* __restrict__ isxe,
                                                     const float2* __restrict__ isze, const float* __restrict__ eta,
                                                     const float2* __restrict__ kap, float k0sq, float hh,
                                                     float shift, int conj, const float2* __restrict__ wid,
                                                     const float2* __restrict__ tcf, const float2* __restrict__ f,
                                                     const double* __restrict__ fscale, float2* __restrict__ u,
                                                     float2* __restrict__ fcoarse, int xcd,
                                                     const int* __restrict__ act) {
    k_mgt_pre2_rr_body<PF, TAB>(hw_block(), nzf, nxf, nzc, nxc, nrhs, rb, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, conj, wid, tcf, f, fscale, u, fcoarse, xcd, act);
}

// k_mgt_post2: out = S(S(u + P u_c)) in one pass on the 16 x 64 fine tile:
// v = u + P u_c is formed on the tile plus a 2-node halo (20 x 68), the first
// sweep on the tile plus a 1-node halo (the staged layout of the other tile
// kernels, 18 x 66), the second on the tile.  Each thread keeps the stencil of
// its halo-1 nodes and uses it for both sweeps.  out must not alias u.
constexpr int P2_X = MT_TX + 4, P2_Z = MT_TZ + 4, P2_N = P2_X * P2_Z;

constexpr int P2_PT = (P2_N + 255) / 256;
// PF: the loads of the next right-hand side are issued before this one is
// computed (more registers, two waves per SIMD); without, three waves per SIMD
template <bool PF, bool TAB>
__device__ __forceinline__ void k_mgt_post2_body(const VBlock& B, int nz, int nx, int64_t n, int nzc, int nxc, int nrhs,
                                                             int rb,
                                                   const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                   const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                   const float* __restrict__ eta, const float2* __restrict__ kap,
                                                   float k0sq, float hh, float shift, int conj,
                                                   const float2* __restrict__ wid, const float2* __restrict__ tcf,
                                                   const float2* __restrict__ uc, const float2* __restrict__ u,
                                                   const float2* __restrict__ f, const double* __restrict__ fscale,
                                                   float2* __restrict__ out, int xcd, const int* __restrict__ act) {
    __shared__ float2 sv2[P2_N];
    __shared__ float2 sv1[MT_NS];
    __shared__ float2 scv[MP_NC];
    int bx, by, bz;
    xcd_tile(B, bx, by, bz, xcd != 0, xcd == 2);
    if (!group_active(act, bz * rb, min(nrhs, bz * rb + rb))) return;
    const int z0 = by * MT_TZ, x0 = bx * MT_TX, tid = threadIdx.y * 64 + threadIdx.x;
    const int cz0 = z0 / 2 - 1, cx0 = x0 / 2 - 1;
    const int64_t ncn = (int64_t)nzc * nxc;
    const bool cj = conj != 0;
    int ih[P2_PT];  // halo-2 nodes staged by this thread
#pragma unroll
    for (int q = 0; q < P2_PT; ++q) {
        const int e = tid + 256 * q;
        ih[q] = -1;
        if (e < P2_N) {
            const int gi = z0 - 2 + e / P2_X, gj = x0 - 2 + e % P2_X;
            if (gi >= 0 && gi < nz && gj >= 0 && gj < nx) ih[q] = gi * nx + gj;
        }
    }
    int cix[MP_CPT];
#pragma unroll
    for (int q = 0; q < MP_CPT; ++q) {
        const int e = tid + 256 * q;
        cix[q] = -1;
        if (e < MP_NC) {
            const int I = cz0 + e / MP_CW, J = cx0 + e % MP_CW;
            if (I >= 0 && I < nzc && J >= 0 && J < nxc) cix[q] = I * nxc + J;
        }
    }
    // halo-1 nodes swept by this thread: every node's stencil operands are
    // loaded before the first is formed, the first right-hand side's fields
    // with them
    int i1[MT_SPT];
    LevelRaw<TAB> raw[MT_SPT];
#pragma unroll
    for (int q = 0; q < MT_SPT; ++q) {
        const int e = tid + 256 * q;
        const int gi = z0 - 1 + e / MT_SW, gj = x0 - 1 + e % MT_SW;
        const bool ok = e < MT_NS && gi >= 0 && gi < nz && gj >= 0 && gj < nx;
        i1[q] = ok ? gi * nx + gj : -1;
        level_raw_load<TAB>(gi, gj, ok, nz, nx, tcf, sx, sz, isxe, isze, eta, kap, wid, raw[q]);
    }
    const int r0 = bz * rb, r1 = min(nrhs, r0 + rb);
    float2 su[P2_PT], cs[MP_CPT], fv[MT_SPT];
    auto fetch = [&](int r) {
#pragma unroll
        for (int q = 0; q < P2_PT; ++q) su[q] = ih[q] >= 0 ? u[(int64_t)r * n + ih[q]] : make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 0; q < MP_CPT; ++q) cs[q] = cix[q] >= 0 ? uc[(int64_t)r * ncn + cix[q]] : make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 0; q < MT_SPT; ++q) fv[q] = i1[q] >= 0 ? f[(int64_t)r * n + i1[q]] : make_float2(0.f, 0.f);
        if (fscale) {
            const float fs = (float)fscale[r];
#pragma unroll
            for (int q = 0; q < MT_SPT; ++q) fv[q] = make_float2(fv[q].x * fs, fv[q].y * fs);
        }
    };
    fetch(r0);
    Coef<float2> c1[MT_SPT];
    float2 w1[MT_SPT];
#pragma unroll
    for (int q = 0; q < MT_SPT; ++q) {
        const int e = tid + 256 * q;
        level_raw_form<TAB>(raw[q], z0 - 1 + e / MT_SW, x0 - 1 + e % MT_SW, nz, nx, k0sq, hh, shift, cj, c1[q], w1[q]);
    }
    for (int r = r0; r < r1; ++r) {
        if (!PF && r > r0) fetch(r);
        lds_barrier();
#pragma unroll
        for (int q = 0; q < MP_CPT; ++q) {
            const int e = tid + 256 * q;
            if (e < MP_NC) scv[e] = cs[q];
        }
        lds_barrier();
        // v = u + P u_c on the halo-2 block (bilinear weights without branches)
#pragma unroll
        for (int q = 0; q < P2_PT; ++q) {
            const int e = tid + 256 * q;
            if (e < P2_N) {
                float2 v = make_float2(0.f, 0.f);
                if (ih[q] >= 0) {
                    const int gi = z0 - 2 + e / P2_X, gj = x0 - 2 + e % P2_X;
                    const int I = (gi >> 1) - cz0, J = (gj >> 1) - cx0;
                    const float wz1 = (gi & 1) ? 0.5f : 0.f, wz0 = 1.f - wz1;
                    const float wx1 = (gj & 1) ? 0.5f : 0.f, wx0 = 1.f - wx1;
                    const float2* c0 = scv + I * MP_CW + J;
                    v = su[q];
                    v.x += wz0 * (wx0 * c0[0].x + wx1 * c0[1].x) + wz1 * (wx0 * c0[MP_CW].x + wx1 * c0[MP_CW + 1].x);
                    v.y += wz0 * (wx0 * c0[0].y + wx1 * c0[1].y) + wz1 * (wx0 * c0[MP_CW].y + wx1 * c0[MP_CW + 1].y);
                }
                sv2[e] = v;
            }
        }
        float2 fc[MT_SPT];
#pragma unroll
        for (int q = 0; q < MT_SPT; ++q) fc[q] = fv[q];
        if (PF && r + 1 < r1) fetch(r + 1);
        lds_barrier();
        // first sweep on the halo-1 block
#pragma unroll
        for (int q = 0; q < MT_SPT; ++q) {
            const int e = tid + 256 * q;
            float2 v1 = make_float2(0.f, 0.f);
            if (e < MT_NS && i1[q] >= 0) {
                const int k = (e / MT_SW + 1) * P2_X + e % MT_SW + 1;  // same node in the halo-2 layout
                float2 acc = cmul(c1[q].dg, sv2[k]);
                cfma(acc, c1[q].fw, sv2[k - 1]);
                cfma(acc, c1[q].fe, sv2[k + 1]);
                cfma(acc, c1[q].fn, sv2[k - P2_X]);
                cfma(acc, c1[q].fs, sv2[k + P2_X]);
                v1 = sv2[k];
                cfma(v1, w1[q], csub(fc[q], acc));
            }
            if (e < MT_NS) sv1[e] = v1;
        }
        lds_barrier();
        // second sweep on the tile
#pragma unroll
        for (int q = 0; q < MT_SPT; ++q) {
            const int e = tid + 256 * q;
            if (e >= MT_NS || i1[q] < 0) continue;
            const int row = e / MT_SW - 1, col = e % MT_SW - 1;
            if (row < 0 || row >= MT_TZ || col < 0 || col >= MT_TX) continue;
            float2 acc = cmul(c1[q].dg, sv1[e]);
            cfma(acc, c1[q].fw, sv1[e - 1]);
            cfma(acc, c1[q].fe, sv1[e + 1]);
            cfma(acc, c1[q].fn, sv1[e - MT_SW]);
            cfma(acc, c1[q].fs, sv1[e + MT_SW]);
            float2 v = sv1[e];
            cfma(v, w1[q], csub(fc[q], acc));
            out[(int64_t)r * n + i1[q]] = v;
        }
    }
}
template <bool PF, bool TAB>
__global__ __launch_bounds__(256, PF ? 2 : 3) void k_mgt_post2(int nz, int nx, int64_t n, int nzc, int nxc, int nrhs,
                                                             int rb,
                                                   const float2* __restrict__ sx, const float2* __restrict__ sz,
                                                   const float2* __restrict__ isxe, const float2* __restrict__ isze,
                                                   const float* __restrict__ eta, const float2* __restrict__ kap,
                                                   float k0sq, float hh, float shift, int conj,
                                                   const float2* __restrict__ wid, const float2* __restrict__ tcf,
                                                   const float2* __restrict__ uc, const float2* __restrict__ u,
                                                   const float2* __restrict__ f, const double* __restrict__ fscale,
                                                   float2* __restrict__ out, int xcd, const int* __restrict__ act) {
    k_mgt_post2_body<PF, TAB>(hw_block(), nz, nx, n, nzc, nxc, nrhs, rb, sx, sz, isxe, isze, eta, kap, k0sq, hh, shift, conj, wid, tcf, uc, u, f, fscale, out, xcd, act);
}

// bilinear prolongation of the coarse correction at fine node (i, j) (zero
// coarse ghosts)
__device__ __forceinline__ float2 prolong_at(const float2* __restrict__ uc, int nzc, int nxc, int i, int j) {
    const int I0 = i >> 1, J0 = j >> 1;
    const bool oi = i & 1, oj = j & 1;
    auto at = [&](int I, int J) -> float2 {
        if (I >= nzc || J >= nxc) return make_float2(0.f, 0.f);
        return uc[(int64_t)I * nxc + J];
    };
    if (!oi && !oj) return at(I0, J0);
    if (oi && !oj) return cscale(cadd(at(I0, J0), at(I0 + 1, J0)), 0.5f);
    if (!oi && oj) return cscale(cadd(at(I0, J0), at(I0, J0 + 1)), 0.5f);
    return cscale(cadd(cadd(at(I0, J0), at(I0 + 1, J0)), cadd(at(I0, J0 + 1), at(I0 + 1, J0 + 1))), 0.25f);
}

// u += P u_c (only when no post-smoothing follows)
__global__ void k_prolong_add(int nz, int nx, int nzc, int nxc, int nrhs, const float2* __restrict__ uc,
                              float2* __restrict__ u) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int r = blockIdx.z;
    if (j >= nx || i >= nz || r >= nrhs) return;
    float2* dst = u + (int64_t)r * nz * nx + (int64_t)i * nx + j;
    *dst = cadd(*dst, prolong_at(uc + (int64_t)r * nzc * nxc, nzc, nxc, i, j));
}

// coarse coefficients by full weighting (exterior values replicate the edge)
__global__ void k_coarsen_medium(int nzf, int nxf, int nzc, int nxc, const double* __restrict__ etaf,
                                 const double2* __restrict__ kapf, double* eta64, double2* kap64, float* eta32,
                                 float2* kap32) {
    const int J = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = blockIdx.y;
    if (J >= nxc || I >= nzc) return;
    const double wz[3] = {0.25, 0.5, 0.25};
    double e = 0.0;
    double2 k = make_double2(0.0, 0.0);
    for (int a = -1; a <= 1; ++a) {
        const int i = min(max(2 * I + a, 0), nzf - 1);
        for (int b = -1; b <= 1; ++b) {
            const int j = min(max(2 * J + b, 0), nxf - 1);
            const double w = wz[a + 1] * wz[b + 1];
            e += w * etaf[(int64_t)i * nxf + j];
            const double2 kv = kapf[(int64_t)i * nxf + j];
            k.x += w * kv.x;
            k.y += w * kv.y;
        }
    }
    const int64_t idx = (int64_t)I * nxc + J;
    eta64[idx] = e;
    kap64[idx] = k;
    eta32[idx] = (float)e;
    kap32[idx] = make_float2((float)k.x, (float)k.y);
}

__global__ void k_fill_c64(int64_t count, float2* p) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = make_float2(0.f, 0.f);
}

static int32_t level_alloc_work(Level& L, int nrhs) {
    if (L.mg_cap >= nrhs) return HELM_OK;
    const size_t bytes = (size_t)nrhs * L.n * sizeof(float2) + 256;
    float2** bufs[4] = {&L.mg_u, &L.mg_f, &L.mg_r, &L.mg_t};
    for (float2** b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
        hipError_t e = hipMalloc((void**)b, bytes);
        if (e != hipSuccess) {
            *b = nullptr;
            L.mg_cap = 0;
            set_error(std::string("hipMalloc (multigrid): ") + hipGetErrorString(e));
            return HELM_ERR_ALLOC;
        }
    }
    L.mg_cap = nrhs;
    return HELM_OK;
}

int32_t mg_build(helm_plan* p, hipStream_t s) {
    for (size_t l = 1; l < p->levels.size(); ++l) {
        Level& F = p->levels[l - 1];
        Level& C = p->levels[l];
        dim3 grid((C.nx + 63) / 64, C.nz);
        hipLaunchKernelGGL(k_coarsen_medium, grid, dim3(64), 0, s, F.nz, F.nx, C.nz, C.nx, F.eta64, F.kap64, C.eta64,
                           C.kap64, C.eta32, C.kap32);
        HELM_HIP(hipGetLastError());
    }
    for (size_t l = 0; l < p->levels.size(); ++l) {
        Level& L = p->levels[l];
        if (!L.wgt) HELM_HIP(hipMalloc((void**)&L.wgt, L.n * sizeof(float2) + 256));
        if (p->mg.omega > 0) {
            hipLaunchKernelGGL(k_fill_weight, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, L.n,
                               (float)p->mg.omega, L.wgt);
        } else {
            dim3 block(64, 4), grid((L.nx + 63) / 64, (L.nz + 3) / 4);
            hipLaunchKernelGGL(k_node_weights, grid, block, 0, s, L.nz, L.nx, L.sx32, L.sz32, L.isxe32, L.isze32,
                               L.eta32, L.kap32, (float)(p->k0 * p->k0), (float)(0.5 / (L.h_m * L.h_m)),
                               (float)p->mg.shift, (int)(l + 1 == p->levels.size()), L.wgt);
        }
        HELM_HIP(hipGetLastError());
        if (!L.wid) HELM_HIP(hipMalloc((void**)&L.wid, L.n * sizeof(float2) + 256));
        {
            dim3 block(64, 4), grid((L.nx + 63) / 64, (L.nz + 3) / 4);
            hipLaunchKernelGGL(k_node_wid, grid, block, 0, s, L.nz, L.nx, L.sx32, L.sz32, L.isxe32, L.isze32, L.eta32,
                               L.kap32, (float)(p->k0 * p->k0), (float)(0.5 / (L.h_m * L.h_m)), (float)p->mg.shift,
                               L.wgt, L.wid);
            HELM_HIP(hipGetLastError());
        }
        if (l > 0 || table0() || L.n <= tail_nodes()) {  // read by the tail and the fused kernels
            if (!L.tcf) HELM_HIP(hipMalloc((void**)&L.tcf, 6 * L.n * sizeof(float2) + 256));
            hipLaunchKernelGGL(k_tail_coef, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, L.nz, L.nx, L.sx32,
                               L.sz32, L.isxe32, L.isze32, L.eta32, L.kap32, (float)(p->k0 * p->k0),
                               (float)(0.5 / (L.h_m * L.h_m)), (float)p->mg.shift, L.wgt, L.tcf);
            HELM_HIP(hipGetLastError());
        }
    }
    return HELM_OK;
}

// right-hand sides per workgroup so that a level launches >= wg_target() workgroups
static int rhs_group(const helm_plan* p, int64_t blocks_xy, int nrhs) {
    const int64_t target = p->mg_wg_target ? p->mg_wg_target : wg_target();
    const int64_t groups = std::max<int64_t>(1, std::min<int64_t>(nrhs, (target + blocks_xy - 1) / blocks_xy));
    return (int)((nrhs + groups - 1) / groups);
}

static int32_t jacobi(helm_plan* p, int l, double shift, bool adj, const float2* u, const float2* f,
                      float2* out, int nrhs, hipStream_t s) {
    const Level& L = p->levels[l];  // conj(M) uses conj(w): the same contraction
    if (u != nullptr && L.wid) {  // sweep from a start vector: LDS-tiled kernel
        const int64_t tiles = (int64_t)((L.nx + MT_TX - 1) / MT_TX) * ((L.nz + MT_TZ - 1) / MT_TZ);
        const int rb = rhs_group(p, tiles, nrhs);
        const dim3 grid((L.nx + MT_TX - 1) / MT_TX, (L.nz + MT_TZ - 1) / MT_TZ, (nrhs + rb - 1) / rb);
        hipLaunchKernelGGL(k_mgt_jacobi, grid, dim3(64, 4), 0, s, L.nz, L.nx, L.n, nrhs, rb, L.sx32, L.sz32, L.isxe32,
                           L.isze32, L.eta32, L.kap32, (float)(p->k0 * p->k0), (float)(0.5 / (L.h_m * L.h_m)),
                           (float)shift, (int)adj, L.wid, u, f, out);
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    const int64_t bxy = (int64_t)((L.nx + 63) / 64) * ((L.nz + 3) / 4);
    const int rb = rhs_group(p, bxy, nrhs);
    dim3 block(64, 4), grid((L.nx + 63) / 64, (L.nz + 3) / 4, (nrhs + rb - 1) / rb);
    const double k0sq = p->k0 * p->k0;
    hipLaunchKernelGGL(k_jacobi, grid, block, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.sx32, L.sz32, L.isxe32, L.isze32,
                       L.eta32, L.kap32, (float)k0sq, (float)(0.5 / (L.h_m * L.h_m)), (float)shift, (int)adj, L.wgt,
                       u, f, out);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}


static size_t tail_lds_bytes(const helm_plan* p, int t) {
    size_t b = 0;
    for (size_t l = t; l < p->levels.size(); ++l) b += 2 * sizeof(float2) * (size_t)p->levels[l].n;
    return b;
}
// LDS of the all-in-LDS tail (k_tail_lds): 9 float2 per node over the tail levels
static size_t tail_lds_bytes_full(const helm_plan* p, int t) {
    size_t b = 0;
    for (size_t l = t; l < p->levels.size(); ++l) b += 9 * sizeof(float2) * (size_t)p->levels[l].n;
    return b;
}
// HELM_MG_TAIL_LDS=0 keeps the coefficients and right-hand sides of the tail in global memory
static bool tail_all_lds() {
    static const bool t = env_int("HELM_MG_TAIL_LDS", 1, 0, 1) != 0;
    return t;
}

static int32_t tail_views(helm_plan* p) {
    const int nl = (int)p->levels.size();
    std::vector<LevelView> h(nl);
    for (int l = 0; l < nl; ++l) {
        const Level& L = p->levels[l];
        h[l] = LevelView{L.nz, L.nx, L.n, L.tcf, L.mg_f, L.mg_r};
    }
    if (!p->mg_views) HELM_HIP(hipMalloc(&p->mg_views, 16 * sizeof(LevelView)));
    HELM_TRY(upload_h2d(p->mg_views, h.data(), nl * sizeof(LevelView)));
    return HELM_OK;
}

// the fused level kernels for this level: stencils from the level's table
// (tab) or formed from the medium; HELM_MG_*_PREFETCH picks the variants that
// load the next right-hand side ahead
static decltype(&k_mgt_pre2_rr<false, false>) pre2_kernel(bool tab) {
    if (pre2_prefetch()) return tab ? k_mgt_pre2_rr<true, true> : k_mgt_pre2_rr<true, false>;
    return tab ? k_mgt_pre2_rr<false, true> : k_mgt_pre2_rr<false, false>;
}
static decltype(&k_mgt_post2<false, false>) post2_kernel(bool tab) {
    if (post2_prefetch()) return tab ? k_mgt_post2<true, true> : k_mgt_post2<true, false>;
    return tab ? k_mgt_post2<false, true> : k_mgt_post2<false, false>;
}

// Solve M_l u = f approximately.  f and u are level-l arrays that must not
// alias the level's scratch: on level 0 they are the caller's input and
// output, on coarser levels L.mg_f and L.mg_u.  L.mg_r holds the residual,
// L.mg_t is the Jacobi ping-pong partner of u.
static int32_t vcycle(helm_plan* p, int l, const float2* f, float2* u, int nrhs, bool adj, hipStream_t s,
                      const double* fscale = nullptr, const int* act = nullptr) {
    Level& L = p->levels[l];
    const MGConfig& c = p->mg;
    HELM_CHECK(!fscale || (l == 0 && mg_scaled_input_ok(p)), HELM_ERR_STATE,
               "multigrid: scaled input needs the fused finest-level kernels");
    if (L.n <= tail_nodes()) {
        const size_t full = tail_lds_bytes_full(p, l);
        if (tail_all_lds() && full <= 160 * 1024)
            hipLaunchKernelGGL(k_tail_lds, dim3(nrhs), dim3(1024), full, s, (const LevelView*)p->mg_views, l,
                               (int)p->levels.size() - 1, (int)adj, c.pre, c.post, c.coarse, f, u, act);
        else
            hipLaunchKernelGGL(k_tail, dim3(nrhs), dim3(1024), tail_lds_bytes(p, l), s,
                               (const LevelView*)p->mg_views, l, (int)p->levels.size() - 1, (int)adj, c.pre, c.post,
                               c.coarse, f, u, act);
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    const bool coarsest = l + 1 == (int)p->levels.size();
    const int sweeps = coarsest ? c.coarse : c.pre;
    const float k0sq = (float)(p->k0 * p->k0), hh = (float)(0.5 / (L.h_m * L.h_m)), sh = (float)c.shift;
    const int64_t tiles = (int64_t)((L.nx + MT_TX - 1) / MT_TX) * ((L.nz + MT_TZ - 1) / MT_TZ);
    const int rb = rhs_group(p, tiles, nrhs);
    const dim3 blk(64, 4), grd((L.nx + MT_TX - 1) / MT_TX, (L.nz + MT_TZ - 1) / MT_TZ, (nrhs + rb - 1) / rb);
    if (!coarsest && c.pre == 2 && c.post == 2 && fused_levels()) {
        // two launches per level: smoothing + residual + restriction, then
        // prolongation + smoothing (the smoothed field lives in mg_t between)
        Level& C = p->levels[l + 1];
        const int64_t bc = (int64_t)((C.nx + RR_TX - 1) / RR_TX) * ((C.nz + RR_TZ - 1) / RR_TZ);
        const int rbc = rhs_group(p, bc, nrhs);
        const dim3 g((C.nx + RR_TX - 1) / RR_TX, (C.nz + RR_TZ - 1) / RR_TZ, (nrhs + rbc - 1) / rbc);
        const float2* tab = (l > 0 || table0()) ? (const float2*)L.tcf : nullptr;
        // algorithmic bytes per launch (timing inside solves): each field read
        // or written once, the stencil (table or eta, kappa, w/dg) once
        Profiler* pf = (p->prof && p->prof->on) ? p->prof : nullptr;
        const double nf = (double)L.n, ncn = (double)C.n, coef = (tab ? 48.0 : 20.0) * nf;
        hipEvent_t ea, eb;
        // timed as every level's launch, or as the finest level's only
        const int kpre = (pf && (pf->mask >> PK_MG_PRE & 1)) ? PK_MG_PRE : (l == 0 ? PK_MG_PRE0 : -1);
        const int kpost = (pf && (pf->mask >> PK_MG_POST & 1)) ? PK_MG_POST : (l == 0 ? PK_MG_POST0 : -1);
        if (kpre >= 0 && prof_pair(pf, kpre, &ea, &eb)) {
            hipExtLaunchKernelGGL(pre2_kernel(tab != nullptr), g, dim3(RR_TX, RR_TZ),
                                  0, s, ea, eb, 0, L.nz, L.nx, C.nz, C.nx, nrhs,
                                  rbc, (const float2*)L.sx32, (const float2*)L.sz32, (const float2*)L.isxe32,
                                  (const float2*)L.isze32, (const float*)L.eta32, (const float2*)L.kap32, k0sq, hh,
                                  sh, (int)adj, (const float2*)L.wid, tab, f, fscale, L.mg_t, C.mg_f, xcd_pre(l), act);
            prof_push(pf, kpre, ea, eb, nrhs * (16.0 * nf + 8.0 * ncn) + coef);
        } else {
            hipLaunchKernelGGL(pre2_kernel(tab != nullptr), g, dim3(RR_TX, RR_TZ), 0,
                               s, L.nz, L.nx, C.nz, C.nx, nrhs, rbc, L.sx32,
                               L.sz32, L.isxe32, L.isze32, L.eta32, L.kap32, k0sq, hh, sh, (int)adj, L.wid, tab, f,
                               fscale, L.mg_t, C.mg_f, xcd_pre(l), act);
        }
        HELM_HIP(hipGetLastError());
        HELM_TRY(vcycle(p, l + 1, C.mg_f, C.mg_u, nrhs, adj, s, nullptr, act));
        if (kpost >= 0 && prof_pair(pf, kpost, &ea, &eb)) {
            hipExtLaunchKernelGGL(post2_kernel(tab != nullptr), grd, blk, 0, s, ea, eb, 0,
                                  L.nz, L.nx, L.n, C.nz, C.nx, nrhs, rb,
                                  (const float2*)L.sx32, (const float2*)L.sz32, (const float2*)L.isxe32,
                                  (const float2*)L.isze32, (const float*)L.eta32, (const float2*)L.kap32, k0sq, hh,
                                  sh, (int)adj, (const float2*)L.wid, tab, (const float2*)C.mg_u,
                                  (const float2*)L.mg_t, f, fscale, u, xcd_post(l), act);
            prof_push(pf, kpost, ea, eb, nrhs * (24.0 * nf + 8.0 * ncn) + coef);
        } else {
            hipLaunchKernelGGL(post2_kernel(tab != nullptr), grd, blk, 0, s, L.nz, L.nx,
                               L.n, C.nz, C.nx, nrhs, rb, L.sx32, L.sz32,
                               L.isxe32, L.isze32, L.eta32, L.kap32, k0sq, hh, sh, (int)adj, L.wid, tab, C.mg_u,
                               L.mg_t, f, fscale, u, xcd_post(l), act);
        }
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    // pre-smoothing from a zero guess, ending in u
    if (sweeps == 0) {
        const int64_t cnt = (int64_t)nrhs * L.n;
        hipLaunchKernelGGL(k_fill_c64, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, cnt, u);
        HELM_HIP(hipGetLastError());
    } else if (sweeps == 1) {
        HELM_TRY(jacobi(p, l, c.shift, adj, nullptr, f, u, nrhs, s));
    } else {
        const int rem = sweeps - 2;
        float2* dst = rem % 2 == 0 ? u : L.mg_t;
        hipLaunchKernelGGL(k_mgt_pre2, grd, blk, 0, s, L.nz, L.nx, L.n, nrhs, rb, L.sx32, L.sz32, L.isxe32, L.isze32,
                           L.eta32, L.kap32, k0sq, hh, sh, (int)adj, L.wid, f, dst);
        HELM_HIP(hipGetLastError());
        const float2* src = dst;
        for (int k = 0; k < rem; ++k) {
            dst = src == u ? L.mg_t : u;
            HELM_TRY(jacobi(p, l, c.shift, adj, src, f, dst, nrhs, s));
            src = dst;
        }
    }
    if (coarsest) return HELM_OK;
    Level& C = p->levels[l + 1];
    // coarse right-hand side R (f - M u), the fine residual kept in LDS
    {
        const int64_t bc = (int64_t)((C.nx + RR_TX - 1) / RR_TX) * ((C.nz + RR_TZ - 1) / RR_TZ);
        const int rbc = rhs_group(p, bc, nrhs);
        const dim3 g((C.nx + RR_TX - 1) / RR_TX, (C.nz + RR_TZ - 1) / RR_TZ, (nrhs + rbc - 1) / rbc);
        hipLaunchKernelGGL(k_mgt_restrict_res, g, dim3(RR_TX, RR_TZ), 0, s, L.nz, L.nx, C.nz, C.nx, nrhs, rbc, L.sx32,
                           L.sz32, L.isxe32, L.isze32, L.eta32, L.kap32, k0sq, hh, sh, (int)adj, u, f, C.mg_f);
        HELM_HIP(hipGetLastError());
    }
    HELM_TRY(vcycle(p, l + 1, C.mg_f, C.mg_u, nrhs, adj, s));
    if (c.post == 0) {
        hipLaunchKernelGGL(k_prolong_add, dim3((L.nx + 63) / 64, L.nz, nrhs), dim3(64), 0, s, L.nz, L.nx, C.nz, C.nx,
                           nrhs, C.mg_u, u);
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    // first post-sweep on u + P u_c into mg_t, the rest ping-pong, ending in u
    hipLaunchKernelGGL(k_mgt_prolong_smooth, grd, blk, 0, s, L.nz, L.nx, L.n, C.nz, C.nx, nrhs, rb, L.sx32, L.sz32,
                       L.isxe32, L.isze32, L.eta32, L.kap32, k0sq, hh, sh, (int)adj, L.wid, C.mg_u, u, f, L.mg_t);
    HELM_HIP(hipGetLastError());
    const float2* src = L.mg_t;
    for (int k = 1; k < c.post; ++k) {
        float2* dst = src == u ? L.mg_t : u;
        HELM_TRY(jacobi(p, l, c.shift, adj, src, f, dst, nrhs, s));
        src = dst;
    }
    if (src != u)
        HELM_HIP(hipMemcpyAsync(u, src, (size_t)nrhs * L.n * sizeof(float2), hipMemcpyDeviceToDevice, s));
    return HELM_OK;
}

// ---- batched V-cycle (helm_batch_solve) ------------------------------------
// The fused level kernels and the LDS tail over the plans of a batch, one
// launch per level and kernel: level l's launch covers the plans for which
// l is a grid level above the tail.  Plans are independent, so running every
// plan's pre-smoothing of level l before any plan's level l + 1 keeps each
// plan's own order (pre 0..t-1, tail at t, post t-1..0).
// TAB: the launch's level takes its stencils from the level table (every
// plan of a launch is at the same level index, so all or none do)
template <bool TAB>
__global__ __launch_bounds__(256, 3) void k_seg_mg_pre(const SegDims* __restrict__ dims, int nseg,
                                                       const SegLevel* __restrict__ lv, int j, int conj) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegLevel& a = lv[s];
    k_mgt_pre2_rr_body<false, TAB>(B, a.nz, a.nx, a.nzc, a.nxc, a.nrhs, a.rb_pre, a.sx, a.sz, a.isxe, a.isze, a.eta,
                                   a.kap, a.k0sq, a.hh, a.shift, conj, a.wid, a.tcf, a.f + j * a.f_step, a.fscale,
                                   a.t, a.fc, a.xcd_pre, a.act);
}
template <bool TAB>
__global__ __launch_bounds__(256, 3) void k_seg_mg_post(const SegDims* __restrict__ dims, int nseg,
                                                        const SegLevel* __restrict__ lv, int j, int conj) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegLevel& a = lv[s];
    k_mgt_post2_body<false, TAB>(B, a.nz, a.nx, a.n, a.nzc, a.nxc, a.nrhs, a.rb_post, a.sx, a.sz, a.isxe, a.isze,
                                 a.eta, a.kap, a.k0sq, a.hh, a.shift, conj, a.wid, a.tcf, a.uc, a.t,
                                 a.f + j * a.f_step, a.fscale, a.u + j * a.u_step, a.xcd_post, a.act);
}
__global__ __launch_bounds__(1024) void k_seg_mg_tail(const SegDims* __restrict__ dims, int nseg,
                                                      const SegTail* __restrict__ tv, int conj) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegTail& a = tv[s];
    k_tail_lds_body(B, a.views, a.t, a.last, conj, a.pre, a.post, a.coarse, a.f, a.u, a.act);
}
void seg_mg_pre(const SegDims* dims, int nseg, int blocks, const SegLevel* lv, int j, int conj, int level,
                hipStream_t s) {
    hipLaunchKernelGGL((level > 0 || table0()) ? k_seg_mg_pre<true> : k_seg_mg_pre<false>, dim3(blocks),
                       dim3(RR_TX, RR_TZ), 0, s, dims, nseg, lv, j, conj);
}
void seg_mg_post(const SegDims* dims, int nseg, int blocks, const SegLevel* lv, int j, int conj, int level,
                 hipStream_t s) {
    hipLaunchKernelGGL((level > 0 || table0()) ? k_seg_mg_post<true> : k_seg_mg_post<false>, dim3(blocks),
                       dim3(64, 4), 0, s, dims, nseg, lv, j, conj);
}
void seg_mg_tail(const SegDims* dims, int nseg, int blocks, size_t lds, const SegTail* tv, int conj, hipStream_t s) {
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)k_seg_mg_tail, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)attr;
    hipLaunchKernelGGL(k_seg_mg_tail, dim3(blocks), dim3(1024), lds, s, dims, nseg, tv, conj);
}

// work arrays and tail views for nrhs right-hand sides (as mg_vcycle does before a cycle)
int32_t mg_prepare(helm_plan* p, int nrhs) {
    HELM_CHECK(p->mg.levels > 0, HELM_ERR_STATE, "multigrid not set up");
    bool grew = false;
    for (auto& L : p->levels) {
        grew = grew || L.mg_cap < nrhs;
        HELM_TRY(level_alloc_work(L, nrhs));
    }
    if (grew || !p->mg_views_valid) {
        HELM_TRY(tail_views(p));
        p->mg_views_valid = true;
        p->generation++;
    }
    return HELM_OK;
}
// first level of the LDS-resident tail, or -1 when the V-cycle does not end
// in it (a coarsest level above the tail size, or a tail too large for LDS)
int mg_tail_level(const helm_plan* p) {
    const int nl = (int)p->levels.size();
    for (int l = 0; l < nl; ++l)
        if (p->levels[l].n <= tail_nodes())
            return tail_all_lds() && tail_lds_bytes_full(p, l) <= 160 * 1024 ? l : -1;
    return -1;
}
size_t mg_tail_lds(const helm_plan* p, int t) { return tail_lds_bytes_full(p, t); }
bool mg_fused_ok(const helm_plan* p) { return p->mg.pre == 2 && p->mg.post == 2 && fused_levels(); }
SegLevel mg_seg_level(helm_plan* p, int l, int nrhs, const float2* f, int64_t f_step, float2* u, int64_t u_step,
                      const double* fscale, const int* act, int64_t wg_target, int pre_grid[3], int post_grid[3]) {
    Level& L = p->levels[l];
    Level& C = p->levels[l + 1];
    SegLevel a{};
    a.nz = L.nz;
    a.nx = L.nx;
    a.nzc = C.nz;
    a.nxc = C.nx;
    a.nrhs = nrhs;
    a.n = L.n;
    auto group = [&](int64_t tiles) {
        const int64_t groups = std::max<int64_t>(1, std::min<int64_t>(nrhs, (wg_target + tiles - 1) / tiles));
        return (int)((nrhs + groups - 1) / groups);
    };
    const int64_t tiles = (int64_t)((L.nx + MT_TX - 1) / MT_TX) * ((L.nz + MT_TZ - 1) / MT_TZ);
    const int64_t bc = (int64_t)((C.nx + RR_TX - 1) / RR_TX) * ((C.nz + RR_TZ - 1) / RR_TZ);
    a.rb_post = wg_target > 0 ? group(tiles) : rhs_group(p, tiles, nrhs);
    a.rb_pre = wg_target > 0 ? group(bc) : rhs_group(p, bc, nrhs);
    a.xcd_pre = xcd_pre(l);
    a.xcd_post = xcd_post(l);
    a.sx = L.sx32;
    a.sz = L.sz32;
    a.isxe = L.isxe32;
    a.isze = L.isze32;
    a.kap = L.kap32;
    a.eta = L.eta32;
    a.wid = L.wid;
    a.tcf = (l > 0 || table0()) ? (const float2*)L.tcf : nullptr;
    a.k0sq = (float)(p->k0 * p->k0);
    a.hh = (float)(0.5 / (L.h_m * L.h_m));
    a.shift = (float)p->mg.shift;
    a.f = f;
    a.f_step = f_step;
    a.u = u;
    a.u_step = u_step;
    a.fscale = fscale;
    a.t = L.mg_t;
    a.fc = C.mg_f;
    a.uc = C.mg_u;
    a.act = act;
    pre_grid[0] = (C.nx + RR_TX - 1) / RR_TX;
    pre_grid[1] = (C.nz + RR_TZ - 1) / RR_TZ;
    pre_grid[2] = (nrhs + a.rb_pre - 1) / a.rb_pre;
    post_grid[0] = (L.nx + MT_TX - 1) / MT_TX;
    post_grid[1] = (L.nz + MT_TZ - 1) / MT_TZ;
    post_grid[2] = (nrhs + a.rb_post - 1) / a.rb_post;
    return a;
}
SegTail mg_seg_tail(helm_plan* p, int t, const int* act) {
    SegTail a{};
    a.views = (const LevelView*)p->mg_views;
    a.t = t;
    a.last = (int)p->levels.size() - 1;
    a.pre = p->mg.pre;
    a.post = p->mg.post;
    a.coarse = p->mg.coarse;
    a.f = p->levels[t].mg_f;
    a.u = p->levels[t].mg_u;
    a.act = act;
    return a;
}

bool mg_scaled_input_ok(const helm_plan* p) {
    return p->levels.size() > 1 && fused_levels() && p->mg.pre == 2 && p->mg.post == 2 &&
           p->levels[0].n > tail_nodes();
}

int32_t mg_vcycle(helm_plan* p, const float2* v, float2* z, int nrhs, bool adjoint, hipStream_t s,
                  const double* fscale, const int* active) {
    HELM_CHECK(p->mg.levels > 0, HELM_ERR_STATE, "multigrid not set up");
    bool grew = false;
    for (auto& L : p->levels) {
        grew = grew || L.mg_cap < nrhs;
        HELM_TRY(level_alloc_work(L, nrhs));
    }
    if (grew || !p->mg_views_valid) {
        HELM_TRY(tail_views(p));
        p->mg_views_valid = true;
        p->generation++;
    }
    return vcycle(p, 0, v, z, nrhs, adjoint, s, fscale, active);
}

// Time one V-cycle for each workgroup target and keep the fastest.
int32_t tune_mg(helm_plan* p, int nrhs, hipStream_t s, double* best_us) {
    *best_us = 0.0;
    if (p->mg.levels == 0) return HELM_OK;
    const size_t bytes = (size_t)nrhs * p->levels[0].n * sizeof(float2);
    float2 *v = nullptr, *z = nullptr;
    HELM_HIP(hipMalloc((void**)&v, bytes));
    if (hipMalloc((void**)&z, bytes) != hipSuccess) {
        (void)hipFree(v);
        set_error("helm_autotune: scratch allocation failed");
        return HELM_ERR_ALLOC;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int32_t st = hipMemsetAsync(v, 0, bytes, s) == hipSuccess ? HELM_OK : HELM_ERR_HIP;
    const int64_t targets[5] = {256, 512, 1024, 2048, 4096};
    double best = 1e30;
    int64_t bt = 2048;
    for (int a = 0; a < 5 && st == HELM_OK; ++a) {
        p->mg_wg_target = targets[a];
        st = mg_vcycle(p, v, z, nrhs, false, s);
        (void)hipEventRecord(e0, s);
        for (int w = 0; w < 3 && st == HELM_OK; ++w) st = mg_vcycle(p, v, z, nrhs, false, s);
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (st == HELM_OK && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
            ms / 3.0 < best) {
            best = ms / 3.0;
            bt = targets[a];
        }
    }
    p->mg_wg_target = bt;
    *best_us = best * 1e3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(v);
    (void)hipFree(z);
    return st;
}

int32_t precond_apply(helm_plan* p, int kind, const float2* v, float2* z, int nrhs, bool adjoint, hipStream_t s) {
    if (kind == HELM_PC_SPECTRAL || kind == HELM_PC_SPECTRAL_PLAIN)
        return spectral_precond(p, v, z, nrhs, adjoint, s, kind == HELM_PC_SPECTRAL);
    if (p->levels[0].op == HELM_OP_COND8) {
        HELM_CHECK(kind == HELM_PC_MG, HELM_ERR_ARG, "conditioned plans take the multigrid preconditioner only");
        return cond_precond(p, v, z, nrhs, adjoint, s);
    }
    if (kind == HELM_PC_MG) return mg_vcycle(p, v, z, nrhs, adjoint, s);
    if (kind == HELM_PC_LEARNED) {
        HELM_CHECK(p->net, HELM_ERR_STATE, "learned preconditioner: no net attached (helm_plan_set_net)");
        HELM_CHECK(!adjoint, HELM_ERR_ARG, "learned preconditioner: adjoint solves use HELM_PC_MG");
        HELM_TRY(mg_vcycle(p, v, z, nrhs, adjoint, s));
        return net_correction(p, p->net, v, z, nrhs, s);
    }
    set_error("preconditioner kind not available");
    return HELM_ERR_STATE;
}

}  // namespace helm

using namespace helm;

extern "C" int32_t helm_mg_setup(helm_plan* p, int32_t levels, double shift, int32_t pre, int32_t post,
                                 int32_t coarse_sweeps, double omega, void* stream) {
    HELM_CHECK(p, HELM_ERR_ARG, "helm_mg_setup: null plan");
    HELM_CHECK(levels >= 1 && levels <= 12, HELM_ERR_ARG, "helm_mg_setup: levels must be in [1, 12]");
    HELM_CHECK(std::isfinite(shift) && shift >= 0, HELM_ERR_ARG, "helm_mg_setup: shift must be >= 0");
    HELM_CHECK(pre >= 0 && post >= 0 && coarse_sweeps >= 0 && pre + post + coarse_sweeps > 0, HELM_ERR_ARG,
               "helm_mg_setup: sweep counts must be >= 0 and not all zero");
    HELM_CHECK(std::isfinite(omega) && omega >= 0 && omega <= 1.5, HELM_ERR_ARG,
               "helm_mg_setup: omega must be 0 (automatic complex weights) or in (0, 1.5]");
    hipStream_t s = (hipStream_t)stream;
    // the same configuration again: the levels exist and helm_set_medium* has
    // already rebuilt their coefficients and weights for the current medium
    if (p->mg.levels > 0 && p->mg.requested == levels && p->mg.shift == shift && p->mg.pre == pre &&
        p->mg.post == post && p->mg.coarse == coarse_sweeps && p->mg.omega == omega && p->medium_set)
        return HELM_OK;
    // drop old coarse levels, build new ones down to >= 3 nodes per axis
    while (p->levels.size() > 1) {
        Level& L = p->levels.back();
        void* ptrs[] = {L.sx64, L.sz64, L.isxe64, L.isze64, L.sx32, L.sz32, L.isxe32, L.isze32,
                        L.eta64, L.kap64, L.eta32, L.kap32, L.mg_u, L.mg_f, L.mg_r, L.mg_t, L.wgt, L.tcf, L.wid};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        p->levels.pop_back();
    }
    const helm_desc& d = p->desc;
    for (int l = 1; l < levels; ++l) {
        const Level& F = p->levels.back();
        const int nzc = (F.nz + 1) / 2, nxc = (F.nx + 1) / 2;
        if (nzc < 3 || nxc < 3) break;
        p->levels.emplace_back();
        const double hc = d.h_mm * (double)(1 << l);
        HELM_TRY(build_level_profiles(p, p->levels.back(), d.x0_mm, d.z0_mm, hc, nzc, nxc));
    }
    p->mg.levels = (int)p->levels.size();
    p->mg_views_valid = false;
    p->generation++;
    {
        int t = 0;
        while (t < (int)p->levels.size() && p->levels[t].n > tail_nodes()) ++t;
        const size_t lds = t < (int)p->levels.size() ? tail_lds_bytes(p, t) : 0;
        HELM_CHECK(lds <= 160 * 1024, HELM_ERR_ARG, "helm_mg_setup: coarse levels exceed the LDS budget");
        // the attribute is per kernel, not per plan: allow the full 160 KB so
        // plans set up on other threads never lower it under each other
        HELM_HIP(hipFuncSetAttribute((const void*)k_tail, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        HELM_HIP(
            hipFuncSetAttribute((const void*)k_tail_lds, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    }
    p->mg.requested = levels;
    p->mg.shift = shift;
    p->mg.pre = pre;
    p->mg.post = post;
    p->mg.coarse = coarse_sweeps;
    p->mg.omega = omega;
    if (p->medium_set) HELM_TRY(mg_build(p, s));
    return HELM_OK;
}

extern "C" int32_t helm_precond_apply(helm_plan* p, int32_t precond, const void* v, void* z, int32_t nrhs,
                                      void* stream) {
    HELM_CHECK(p && v && z, HELM_ERR_ARG, "helm_precond_apply: null argument");
    HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_precond_apply: call helm_set_medium first");
    HELM_CHECK(nrhs >= 1 && nrhs <= 65535, HELM_ERR_ARG, "helm_precond_apply: nrhs out of range");
    HELM_CHECK(v != z, HELM_ERR_ARG, "helm_precond_apply: v and z must not alias");
    return precond_apply(p, precond, (const float2*)v, (float2*)z, nrhs, false, (hipStream_t)stream);
}

extern "C" int32_t helm_precond_apply_ex(helm_plan* p, int32_t precond, const void* v, void* z, int32_t nrhs,
                                         int32_t adjoint, void* stream) {
    HELM_CHECK(p && v && z, HELM_ERR_ARG, "helm_precond_apply_ex: null argument");
    HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_precond_apply_ex: call helm_set_medium first");
    HELM_CHECK(nrhs >= 1 && nrhs <= 65535, HELM_ERR_ARG, "helm_precond_apply_ex: nrhs out of range");
    HELM_CHECK(v != z, HELM_ERR_ARG, "helm_precond_apply_ex: v and z must not alias");
    return precond_apply(p, precond, (const float2*)v, (float2*)z, nrhs, adjoint != 0, (hipStream_t)stream);
}
