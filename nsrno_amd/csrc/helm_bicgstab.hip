// Right-preconditioned BiCGSTAB with FP64 refinement.
//
// Algorithm: nsrno/dense_wave.py:118-148 (short recurrence, shadow residual
// r~ = r0, breakdown when |rho| or |omega| or |<r~, v>| vanish, the half-step
// exit on ||s||, and every 40 iterations the residual recomputed from x with
// 'diverged' when it is not finite or above 1e5 times the pass's scale)
// inside the mixed FP32-inner / FP64-outer refinement of
// dense_wave.py:167-192: outer FP64 true-residual corrections, each running a
// complex64 BiCGSTAB on the current residual to an inner tolerance.
// Differences: every right-hand side runs its own recurrence (per-RHS scalars
// on the device, lock-step kernels, converged, broken-down or diverged RHS
// are skipped) and keeps its own reason, all inner products are accumulated
// in FP64, the 40-step refresh recomputes b - A x in FP64, and the host
// synchronises only every CHECK steps.
#include <cmath>

#include "helm_internal.h"

namespace helm {

constexpr int BCHUNK = 4096;
constexpr int CHECK = 8;  // host looks at the flags every CHECK iterations

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block-reduce NQ complex doubles; thread 0 writes partial[(r*NQ + q)*nch + chunk]
template <int NQ>
__device__ __forceinline__ void block_write(double2 (&acc)[NQ], double* lds, double2* partial, int r, int nch,
                                            int chunk) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double a = wsum(acc[q].x), b = wsum(acc[q].y);
        if (lane == 0) {
            lds[(q * 4 + wid) * 2] = a;
            lds[(q * 4 + wid) * 2 + 1] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        double2 s = make_double2(0.0, 0.0);
        for (int k = 0; k < 4; ++k) {
            s.x += lds[(q * 4 + k) * 2];
            s.y += lds[(q * 4 + k) * 2 + 1];
        }
        partial[((int64_t)r * NQ + q) * nch + chunk] = s;
    }
}

__device__ __forceinline__ double2 cdotc(float2 a, float2 b) {  // conj(a) * b in double
    return make_double2((double)a.x * b.x + (double)a.y * b.y, (double)a.x * b.y - (double)a.y * b.x);
}

// Per-RHS scalar slots in sc[r*8 + k]
enum { S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_BETA, S_RNORM, S_STATE, S_SCALE };
// state: 0 running, 1 converged, 2 breakdown, 3 diverged

// p = r + beta (p - omega v)
__global__ void k_bi_p(int64_t n, const float2* __restrict__ r, float2* __restrict__ p, const float2* __restrict__ v,
                       const double2* __restrict__ sc) {
    const int rr = blockIdx.y;
    const double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = (int64_t)rr * n + i;
    const float2 beta = make_float2((float)s[S_BETA].x, (float)s[S_BETA].y);
    const float2 om = make_float2((float)s[S_OMEGA].x, (float)s[S_OMEGA].y);
    float2 t = p[o];
    float2 vv = v[o];
    t.x -= om.x * vv.x - om.y * vv.y;
    t.y -= om.x * vv.y + om.y * vv.x;
    float2 res = r[o];
    res.x += beta.x * t.x - beta.y * t.y;
    res.y += beta.x * t.y + beta.y * t.x;
    p[o] = res;
}

// partial dots of NQ (a_q, b_q) pairs: conj(a) . b
template <int NQ>
__global__ __launch_bounds__(256) void k_bi_dots(int64_t n, int nch, const float2* a0, const float2* b0,
                                                 const float2* a1, const float2* b1, const double2* __restrict__ sc,
                                                 double2* __restrict__ partial) {
    __shared__ double lds[NQ * 8];
    const int chunk = blockIdx.x, rr = blockIdx.y;
    if (sc[rr * 8 + S_STATE].x != 0.0) return;
    double2 acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = make_double2(0.0, 0.0);
    const int64_t beg = (int64_t)chunk * BCHUNK, end = min(beg + BCHUNK, n);
    const int64_t off = (int64_t)rr * n;
    for (int64_t i = beg + threadIdx.x; i < end; i += 256) {
        acc[0] = cadd(acc[0], cdotc(a0[off + i], b0[off + i]));
        if (NQ > 1) acc[NQ - 1] = cadd(acc[NQ - 1], cdotc(a1[off + i], b1[off + i]));
    }
    block_write<NQ>(acc, lds, partial, rr, nch, chunk);
}

// s = r - alpha v ; partial |s|^2
__global__ __launch_bounds__(256) void k_bi_s(int64_t n, int nch, const float2* __restrict__ r,
                                              const float2* __restrict__ v, float2* __restrict__ s_,
                                              const double2* __restrict__ sc, double2* __restrict__ partial) {
    __shared__ double lds[8];
    const int chunk = blockIdx.x, rr = blockIdx.y;
    const double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    const float2 al = make_float2((float)s[S_ALPHA].x, (float)s[S_ALPHA].y);
    const int64_t beg = (int64_t)chunk * BCHUNK, end = min(beg + BCHUNK, n);
    const int64_t off = (int64_t)rr * n;
    double2 acc[1] = {make_double2(0.0, 0.0)};
    for (int64_t i = beg + threadIdx.x; i < end; i += 256) {
        float2 vv = v[off + i];
        float2 t = r[off + i];
        t.x -= al.x * vv.x - al.y * vv.y;
        t.y -= al.x * vv.y + al.y * vv.x;
        s_[off + i] = t;
        acc[0].x += (double)t.x * t.x + (double)t.y * t.y;
    }
    block_write<1>(acc, lds, partial, rr, nch, chunk);
}

// x += alpha phat + omega shat (FP64); r = s - omega t; partials |r|^2, <shadow, r>
__global__ __launch_bounds__(256) void k_bi_xr(int64_t n, int nch, double2* __restrict__ x,
                                               const float2* __restrict__ phat, const float2* __restrict__ shat,
                                               const float2* __restrict__ s_, const float2* __restrict__ t_,
                                               float2* __restrict__ r, const float2* __restrict__ shadow,
                                               const double2* __restrict__ sc, double2* __restrict__ partial) {
    __shared__ double lds[16];
    const int chunk = blockIdx.x, rr = blockIdx.y;
    const double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    const double2 al = s[S_ALPHA], om = s[S_OMEGA];
    const float2 omf = make_float2((float)om.x, (float)om.y);
    const int64_t beg = (int64_t)chunk * BCHUNK, end = min(beg + BCHUNK, n);
    const int64_t off = (int64_t)rr * n;
    double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    for (int64_t i = beg + threadIdx.x; i < end; i += 256) {
        const float2 ph = phat[off + i], sh = shat[off + i];
        double2 xv = x[off + i];
        xv.x += al.x * ph.x - al.y * ph.y + om.x * sh.x - om.y * sh.y;
        xv.y += al.x * ph.y + al.y * ph.x + om.x * sh.y + om.y * sh.x;
        x[off + i] = xv;
        const float2 tv = t_[off + i];
        float2 rv = s_[off + i];
        rv.x -= omf.x * tv.x - omf.y * tv.y;
        rv.y -= omf.x * tv.y + omf.y * tv.x;
        r[off + i] = rv;
        acc[0].x += (double)rv.x * rv.x + (double)rv.y * rv.y;
        acc[1] = cadd(acc[1], cdotc(shadow[off + i], rv));
    }
    block_write<2>(acc, lds, partial, rr, nch, chunk);
}

// x += alpha phat (early exit when |s| is small enough)
__global__ void k_bi_xonly(int64_t n, double2* __restrict__ x, const float2* __restrict__ phat,
                           const double2* __restrict__ sc, const int* __restrict__ flag) {
    const int rr = blockIdx.y;
    if (!flag[rr]) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 al = sc[rr * 8 + S_ALPHA];
    const float2 ph = phat[(int64_t)rr * n + i];
    double2 xv = x[(int64_t)rr * n + i];
    xv.x += al.x * ph.x - al.y * ph.y;
    xv.y += al.x * ph.y + al.y * ph.x;
    x[(int64_t)rr * n + i] = xv;
}

__device__ double2 red(const double2* src, int nch, double* lds) {
    double2 a = make_double2(0.0, 0.0);
    for (int c = threadIdx.x; c < nch; c += blockDim.x) a = cadd(a, src[c]);
    a.x = wsum(a.x);
    a.y = wsum(a.y);
    return a;  // block of 64 threads: full sum in every lane
}

__device__ double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

__device__ double cabs2(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }

// stage 0: rho = <shadow, r> (partials q0), beta, breakdown test
__global__ void k_bi_scalar_rho(int nch, const double2* __restrict__ partial, double2* __restrict__ sc) {
    const int rr = blockIdx.x;
    double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    double2 rho = red(partial + (int64_t)rr * nch, nch, nullptr);
    if (threadIdx.x) return;
    if (!isfinite(rho.x) || !isfinite(rho.y) || cabs2(rho) < 1e-30 || cabs2(s[S_OMEGA]) < 1e-30) {
        s[S_STATE].x = 2.0;
        return;
    }
    s[S_BETA] = cmul(cdiv(rho, s[S_RHO_OLD]), cdiv(s[S_ALPHA], s[S_OMEGA]));
    s[S_RHO] = rho;
}

// stage 1: alpha = rho / <shadow, v>
__global__ void k_bi_scalar_alpha(int nch, const double2* __restrict__ partial, double2* __restrict__ sc,
                                  int* __restrict__ apps) {
    const int rr = blockIdx.x;
    double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    double2 den = red(partial + (int64_t)rr * nch, nch, nullptr);
    if (threadIdx.x) return;
    apps[rr] += 1;  // v = A phat
    if (cabs2(den) < 1e-30 || !isfinite(den.x)) {
        s[S_STATE].x = 2.0;
        return;
    }
    s[S_ALPHA] = cdiv(s[S_RHO], den);
}

// stage 2: |s| test (converged after the half step -> flag for x += alpha phat)
__global__ void k_bi_scalar_s(int nch, double tol_abs_scale, const double2* __restrict__ partial,
                              double2* __restrict__ sc, int* __restrict__ half, int* __restrict__ its) {
    const int rr = blockIdx.x;
    double2* s = sc + rr * 8;
    if (threadIdx.x == 0) half[rr] = 0;
    if (s[S_STATE].x != 0.0) return;
    double2 ss = red(partial + (int64_t)rr * nch, nch, nullptr);
    if (threadIdx.x) return;
    const double sn = sqrt(ss.x);
    if (sn < s[S_SCALE].x * tol_abs_scale) {
        half[rr] = 1;
        s[S_RNORM].x = sn;
        s[S_STATE].x = 1.0;
        its[rr] += 1;  // the step counts (dense_wave.py:135-136)
    }
}

// stage 3: omega = <t, s> / <t, t>
__global__ void k_bi_scalar_omega(int nch, const double2* __restrict__ partial, double2* __restrict__ sc,
                                  int* __restrict__ apps) {
    const int rr = blockIdx.x;
    double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    double2 ts = red(partial + ((int64_t)rr * 2 + 0) * nch, nch, nullptr);
    double2 tt = red(partial + ((int64_t)rr * 2 + 1) * nch, nch, nullptr);
    if (threadIdx.x) return;
    apps[rr] += 1;  // t = A shat
    // denominator clamped at 1e-30 (dense_wave.py:138)
    s[S_OMEGA] = make_double2(ts.x / fmax(tt.x, 1e-30), ts.y / fmax(tt.x, 1e-30));
}

// stage 4: |r| and next rho; convergence
__global__ void k_bi_scalar_r(int nch, double tol_abs_scale, const double2* __restrict__ partial,
                              double2* __restrict__ sc, int* __restrict__ its) {
    const int rr = blockIdx.x;
    double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    double2 r2 = red(partial + ((int64_t)rr * 2 + 0) * nch, nch, nullptr);
    if (threadIdx.x) return;
    its[rr] += 1;
    s[S_RHO_OLD] = s[S_RHO];
    const double rn = sqrt(r2.x);
    s[S_RNORM].x = rn;
    if (!isfinite(rn)) s[S_STATE].x = 3.0;
    else if (rn < s[S_SCALE].x * tol_abs_scale) s[S_STATE].x = 1.0;
}

// 40-step refresh (dense_wave.py:140-142): r = c64(b - A x) for the running
// RHS (r64 holds b - A x in FP64), partial |r|^2
__global__ __launch_bounds__(256) void k_bi_refresh(int64_t n, int nch, const double2* __restrict__ r64,
                                                    float2* __restrict__ r, const double2* __restrict__ sc,
                                                    double2* __restrict__ partial) {
    __shared__ double lds[8];
    const int chunk = blockIdx.x, rr = blockIdx.y;
    if (sc[rr * 8 + S_STATE].x != 0.0) return;
    const int64_t beg = (int64_t)chunk * BCHUNK, end = min(beg + BCHUNK, n);
    const int64_t off = (int64_t)rr * n;
    double2 acc[1] = {make_double2(0.0, 0.0)};
    for (int64_t i = beg + threadIdx.x; i < end; i += 256) {
        const double2 a = r64[off + i];
        r[off + i] = make_float2((float)a.x, (float)a.y);
        acc[0].x += a.x * a.x + a.y * a.y;
    }
    block_write<1>(acc, lds, partial, rr, nch, chunk);
}

// diverged when the recomputed residual is not finite or exceeds 1e5 x the
// pass's scale; converged when it meets the inner tolerance
__global__ void k_bi_scalar_refresh(int nch, double tol_abs_scale, const double2* __restrict__ partial,
                                    double2* __restrict__ sc, int* __restrict__ apps) {
    const int rr = blockIdx.x;
    double2* s = sc + rr * 8;
    if (s[S_STATE].x != 0.0) return;
    double2 r2 = red(partial + (int64_t)rr * nch, nch, nullptr);
    if (threadIdx.x) return;
    apps[rr] += 1;
    const double rn = sqrt(r2.x);
    s[S_RNORM].x = rn;
    if (!isfinite(rn) || rn > 1e5 * s[S_SCALE].x) s[S_STATE].x = 3.0;
    else if (rn < s[S_SCALE].x * tol_abs_scale) s[S_STATE].x = 1.0;
}

__global__ void k_bi_izero(int count, int* p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// initialise the inner recurrence: r = shadow = c64(r64), p = v = 0, scalars
__global__ void k_bi_init(int64_t n, const double2* __restrict__ r64, float2* __restrict__ r,
                          float2* __restrict__ shadow, float2* __restrict__ p, float2* __restrict__ v,
                          const int* __restrict__ active) {
    const int rr = blockIdx.y;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = (int64_t)rr * n + i;
    const double2 a = r64[o];
    const float2 f = active[rr] ? make_float2((float)a.x, (float)a.y) : make_float2(0.f, 0.f);
    r[o] = f;
    shadow[o] = f;
    p[o] = make_float2(0.f, 0.f);
    v[o] = make_float2(0.f, 0.f);
}

__global__ void k_bi_scalars_init(int nrhs, const double* __restrict__ beta, const int* __restrict__ active,
                                  double2* __restrict__ sc) {
    const int rr = blockIdx.x * blockDim.x + threadIdx.x;
    if (rr >= nrhs) return;
    double2* s = sc + rr * 8;
    const double b = beta[rr];
    s[S_RHO_OLD] = make_double2(1.0, 0.0);
    s[S_ALPHA] = make_double2(1.0, 0.0);
    s[S_OMEGA] = make_double2(1.0, 0.0);
    s[S_RNORM] = make_double2(b, 0.0);
    s[S_SCALE] = make_double2(b, 0.0);  // inner tolerances are relative to the outer residual
    s[S_STATE] = make_double2(active[rr] && b > 0 ? 0.0 : 1.0, 0.0);
}

__global__ void k_bi_norm_partial(int64_t n, int nch, const double2* __restrict__ v, double2* __restrict__ partial) {
    __shared__ double lds[8];
    const int chunk = blockIdx.x, rr = blockIdx.y;
    const int64_t beg = (int64_t)chunk * BCHUNK, end = min(beg + BCHUNK, n);
    double2 acc[1] = {make_double2(0.0, 0.0)};
    for (int64_t i = beg + threadIdx.x; i < end; i += 256) {
        const double2 a = v[(int64_t)rr * n + i];
        acc[0].x += a.x * a.x + a.y * a.y;
    }
    block_write<1>(acc, lds, partial, rr, nch, chunk);
}

// outer: beta = ||r64||, rel, active flags (first call also sets ||b||)
__global__ void k_bi_outer(int nch, double tol, int first, const double2* __restrict__ partial,
                           double* __restrict__ bnorm, double* __restrict__ beta, double* __restrict__ rel,
                           int* __restrict__ active) {
    const int rr = blockIdx.x;
    double2 a = red(partial + (int64_t)rr * nch, nch, nullptr);
    if (threadIdx.x) return;
    const double nrm = sqrt(a.x);
    if (first) {
        bnorm[rr] = fmax(nrm, 1e-30);
        active[rr] = 1;
    }
    beta[rr] = nrm;
    rel[rr] = nrm / bnorm[rr];
    if (!(rel[rr] >= tol) || !isfinite(rel[rr])) active[rr] = 0;
}

__global__ void k_bi_zero(int64_t count, double2* p) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = make_double2(0.0, 0.0);
}

static int32_t bi_alloc(float2** p, size_t count) {
    if (*p) return HELM_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(float2) + 256);
    if (e != hipSuccess) {
        *p = nullptr;
        set_error(std::string("hipMalloc (bicgstab): ") + hipGetErrorString(e));
        return HELM_ERR_ALLOC;
    }
    return HELM_OK;
}

int32_t solve_bicgstab(helm_plan* p, const double2* b, double2* x, int nrhs, const helm_solve_opts* o,
                       helm_solve_info* info, hipStream_t s) {
    HELM_TRY(ensure_workspace(p, nrhs, std::max(o->restart, 1)));
    Workspace& W = p->ws;
    const Level& L = p->levels[0];
    const int64_t n = L.n;
    const size_t cnt = (size_t)W.nrhs * n;
    float2** vecs[] = {&W.bp, &W.bv, &W.bs, &W.bt, &W.bphat, &W.bshat, &W.bshadow, &W.br};
    if (W.bi_cap < cnt) {  // (re)allocate when the workspace grew
        for (float2** q : vecs)
            if (*q) {
                (void)hipFree(*q);
                *q = nullptr;
            }
        if (W.bsc) (void)hipFree(W.bsc);
        W.bsc = nullptr;
        W.bi_cap = 0;
    }
    for (float2** q : vecs) HELM_TRY(bi_alloc(q, cnt));
    if (!W.bsc) HELM_HIP(hipMalloc((void**)&W.bsc, (size_t)W.nrhs * 8 * sizeof(double2) + 256));
    W.bi_cap = cnt;
    const int nch = (int)((n + BCHUNK - 1) / BCHUNK);
    HELM_CHECK((int64_t)nrhs * 2 * nch <= W.partial_len, HELM_ERR_STATE, "bicgstab: partial buffer too small");
    const double k0sq = p->k0 * p->k0;
    const bool adj = o->adjoint != 0;
    const int pc = o->precond;
    const dim3 blk(256), gvec((unsigned)((n + 255) / 256), nrhs), gch(nch, nrhs);
    std::vector<double2> hsc((size_t)nrhs * 8);

    if (!o->use_x0) {
        const int64_t c = (int64_t)nrhs * n;
        hipLaunchKernelGGL(k_bi_zero, dim3((unsigned)((c + 255) / 256)), blk, 0, s, c, x);
    }
    // per-RHS counters on the device: completed steps (W.iters) and operator
    // applications inside the recurrences (W.used)
    hipLaunchKernelGGL(k_bi_izero, dim3((nrhs + 255) / 256), blk, 0, s, nrhs, W.iters);
    hipLaunchKernelGGL(k_bi_izero, dim3((nrhs + 255) / 256), blk, 0, s, nrhs, W.used);
    int total_steps = 0, outer = 0;
    // state of each RHS at the end of its last inner pass (0 running/budget,
    // 1 converged, 2 breakdown, 3 diverged); reset at every outer pass
    std::vector<int> rstate(nrhs, 0);
    for (;; ++outer) {
        HELM_TRY(launch_apply64(L, k0sq, x, W.r64, nrhs, adj, b, s));
        hipLaunchKernelGGL(k_bi_norm_partial, gch, blk, 0, s, n, nch, W.r64, W.partial);
        if (outer == 0) {  // ||b||
            hipLaunchKernelGGL(k_bi_norm_partial, gch, blk, 0, s, n, nch, b, W.partial + (size_t)nrhs * nch);
            hipLaunchKernelGGL(k_bi_outer, dim3(nrhs), dim3(64), 0, s, nch, o->tolerance, 1,
                               W.partial + (size_t)nrhs * nch, W.bnorm, W.beta, W.rel, W.active);
        }
        hipLaunchKernelGGL(k_bi_outer, dim3(nrhs), dim3(64), 0, s, nch, o->tolerance, 0, W.partial, W.bnorm, W.beta,
                           W.rel, W.active);
        HELM_HIP(hipGetLastError());
        HELM_HIP(hipMemcpyAsync(W.h_rel.data(), W.rel, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        HELM_HIP(hipMemcpyAsync(W.h_active.data(), W.active, nrhs * sizeof(int), hipMemcpyDeviceToHost, s));
        HELM_HIP(hipStreamSynchronize(s));
        if (o->callback) o->callback(o->user, outer + total_steps * 2, W.h_rel.data(), nrhs);
        int n_active = 0;
        for (int r = 0; r < nrhs; ++r) n_active += W.h_active[r];
        if (n_active == 0 || total_steps >= o->max_iterations || outer >= 5 + o->max_iterations / 400) break;
        for (int r = 0; r < nrhs; ++r)
            if (W.h_active[r]) rstate[r] = 0;
        // inner tolerance relative to the current residual: reach tol*||b|| with margin
        double worst = 0;
        for (int r = 0; r < nrhs; ++r)
            if (W.h_active[r]) worst = std::max(worst, W.h_rel[r]);
        const double inner = std::min(0.3, std::max(2e-6, 0.5 * o->tolerance / worst));
        hipLaunchKernelGGL(k_bi_init, gvec, blk, 0, s, n, W.r64, W.br, W.bshadow, W.bp, W.bv, W.active);
        hipLaunchKernelGGL(k_bi_scalars_init, dim3((nrhs + 63) / 64), dim3(64), 0, s, nrhs, W.beta, W.active, W.bsc);
        HELM_HIP(hipGetLastError());
        // rho_0 = <shadow, r>
        hipLaunchKernelGGL(k_bi_dots<1>, gch, blk, 0, s, n, nch, W.bshadow, W.br, nullptr, nullptr, W.bsc, W.partial);
        const int inner_budget = o->max_iterations - total_steps;
        for (int it = 0; it < inner_budget; ++it) {
            hipLaunchKernelGGL(k_bi_scalar_rho, dim3(nrhs), dim3(64), 0, s, nch, W.partial, W.bsc);
            hipLaunchKernelGGL(k_bi_p, gvec, blk, 0, s, n, W.br, W.bp, W.bv, W.bsc);
            const float2* ph = W.bp;
            if (pc != HELM_PC_NONE) {
                HELM_TRY(precond_apply(p, pc, W.bp, W.bphat, nrhs, adj, s));
                ph = W.bphat;
            } else {
                HELM_HIP(hipMemcpyAsync(W.bphat, W.bp, (size_t)nrhs * n * sizeof(float2), hipMemcpyDeviceToDevice, s));
                ph = W.bphat;
            }
            HELM_TRY(launch_apply32(L, k0sq, ph, W.bv, nrhs, adj, nullptr, 0.0f, s));
            hipLaunchKernelGGL(k_bi_dots<1>, gch, blk, 0, s, n, nch, W.bshadow, W.bv, nullptr, nullptr, W.bsc,
                               W.partial);
            hipLaunchKernelGGL(k_bi_scalar_alpha, dim3(nrhs), dim3(64), 0, s, nch, W.partial, W.bsc, W.used);
            hipLaunchKernelGGL(k_bi_s, gch, blk, 0, s, n, nch, W.br, W.bv, W.bs, W.bsc, W.partial);
            hipLaunchKernelGGL(k_bi_scalar_s, dim3(nrhs), dim3(64), 0, s, nch, inner, W.partial, W.bsc, W.cyc,
                               W.iters);
            hipLaunchKernelGGL(k_bi_xonly, gvec, blk, 0, s, n, x, W.bphat, W.bsc, W.cyc);
            if (pc != HELM_PC_NONE) {
                HELM_TRY(precond_apply(p, pc, W.bs, W.bshat, nrhs, adj, s));
            } else {
                HELM_HIP(hipMemcpyAsync(W.bshat, W.bs, (size_t)nrhs * n * sizeof(float2), hipMemcpyDeviceToDevice, s));
            }
            HELM_TRY(launch_apply32(L, k0sq, W.bshat, W.bt, nrhs, adj, nullptr, 0.0f, s));
            hipLaunchKernelGGL(k_bi_dots<2>, gch, blk, 0, s, n, nch, W.bt, W.bs, W.bt, W.bt, W.bsc, W.partial);
            hipLaunchKernelGGL(k_bi_scalar_omega, dim3(nrhs), dim3(64), 0, s, nch, W.partial, W.bsc, W.used);
            hipLaunchKernelGGL(k_bi_xr, gch, blk, 0, s, n, nch, x, W.bphat, W.bshat, W.bs, W.bt, W.br, W.bshadow,
                               W.bsc, W.partial);
            hipLaunchKernelGGL(k_bi_scalar_r, dim3(nrhs), dim3(64), 0, s, nch, inner, W.partial, W.bsc, W.iters);
            if (it % 40 == 39) {  // residual refresh and divergence test (dense_wave.py:140-142)
                HELM_TRY(launch_apply64(L, k0sq, x, W.r64, nrhs, adj, b, s));
                hipLaunchKernelGGL(k_bi_refresh, gch, blk, 0, s, n, nch, W.r64, W.br, W.bsc, W.partial);
                hipLaunchKernelGGL(k_bi_scalar_refresh, dim3(nrhs), dim3(64), 0, s, nch, inner, W.partial, W.bsc,
                                   W.used);
            }
            HELM_HIP(hipGetLastError());
            total_steps++;
            if ((it + 1) % CHECK == 0 || total_steps >= o->max_iterations) {
                HELM_HIP(hipMemcpyAsync(hsc.data(), W.bsc, hsc.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
                HELM_HIP(hipStreamSynchronize(s));
                bool any = false;
                for (int r = 0; r < nrhs; ++r) {
                    const int st = (int)hsc[(size_t)r * 8 + S_STATE].x;
                    if (st == 0) any = true;
                    if (W.h_active[r]) rstate[r] = st;
                }
                if (!any) break;
            }
            // rho for the next iteration
            hipLaunchKernelGGL(k_bi_dots<1>, gch, blk, 0, s, n, nch, W.bshadow, W.br, nullptr, nullptr, W.bsc,
                               W.partial);
        }
    }
    std::vector<int> its(nrhs), apps(nrhs);
    HELM_HIP(hipMemcpyAsync(its.data(), W.iters, nrhs * sizeof(int), hipMemcpyDeviceToHost, s));
    HELM_HIP(hipMemcpyAsync(apps.data(), W.used, nrhs * sizeof(int), hipMemcpyDeviceToHost, s));
    HELM_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < nrhs; ++r) {
        const bool conv = !(W.h_rel[r] >= o->tolerance);
        int reason = HELM_ITERATION_BUDGET;
        if (conv) reason = HELM_CONVERGED;
        else if (!std::isfinite(W.h_rel[r]) || rstate[r] == 3) reason = HELM_DIVERGED;
        else if (rstate[r] == 2) reason = HELM_BREAKDOWN;
        info[r].reason = reason;
        info[r].first_reason = reason;
        info[r].fallback = 0;
        info[r].graph_fallbacks = 0;
        info[r].iterations = its[r];
        info[r].applications = apps[r] + outer + 1;  // + one FP64 true residual per outer pass
        info[r].cycles = outer;
        info[r].relative_residual = W.h_rel[r];
    }
    return HELM_OK;
}

}  // namespace helm
