// Restarted (F)GMRES(m) with FP64 inner products, FP64 Givens least squares
// and FP64 true-residual refinement at every restart.
//
// Algorithm: nsrno/wave_solver.py:9-36 (right preconditioning, classical
// Gram-Schmidt applied twice, convergence decided on the true residual at
// each restart, reasons converged / iteration_budget), with these changes
// made for the GPU:
//   * classical Gram-Schmidt with the second pass taken only where the first
//     cancelled heavily (||w'|| < 0.7 ||w||, Kahan-Parlett); the decision is
//     made on the device per right-hand side, the second-pass kernels are
//     always launched and return at once where it is not needed, so the
//     launch sequence stays fixed.  Usually this halves the basis reads;
//   * every right-hand side has its own Krylov space; all are advanced in
//     lock step by the same kernels (grid dimension = RHS), so one launch
//     sequence serves the whole angle batch;
//   * the small least-squares problem is solved on the device by Givens
//     rotations in FP64 (the reference copies H to the CPU and calls lstsq),
//     which also gives the residual estimate |g_{j+1}| after every step: a
//     right-hand side whose estimate reaches the tolerance stops its cycle
//     early, without a host synchronisation;
//   * the true residual b - A x is recomputed in FP64 with x held in FP64
//     (iterative refinement), so the inner FP32 cycles can reach tolerances
//     below the FP32 operator's rounding floor;
//   * the host synchronises once per restart (to read the residuals and
//     decide termination), never inside a cycle.
// Directions Z_j = M^{-1} v_j are stored (flexible GMRES), so the
// preconditioner may change between steps (the learned network does).
//
// Data layout: V[k][r][n] (c64), Z[k][r][n] (c64), w[r][n], x[r][n] (c128).
// Inner products are reduced in two deterministic stages: per-chunk partial
// sums (one workgroup per 4096-node chunk, per vector group, per RHS) then a
// fixed-order sum over chunks; no atomics, so results are bit reproducible.
#include <cmath>

#include "helm_internal.h"

namespace helm {

constexpr int CHUNK = 4096;  // nodes per reduction chunk (256 threads x 16)
constexpr int KB = 8;        // basis vectors per dot-product workgroup

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of a complex double over a 256-thread block; result valid in thread 0.
__device__ __forceinline__ double2 block_sum(double2 v, double* lds) {
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        lds[2 * wid] = v.x;
        lds[2 * wid + 1] = v.y;
    }
    __syncthreads();
    double2 out = make_double2(0.0, 0.0);
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            out.x += lds[2 * k];
            out.y += lds[2 * k + 1];
        }
    return out;
}

static inline int nchunks(int64_t n) { return (int)((n + CHUNK - 1) / CHUNK); }

// partial[(r*KMAX + k)*nch + chunk] = sum_chunk conj(V_k[r]) * w[r]  for k < kcount
// Also (wnorm != nullptr, first vector group) wnorm[r*nch+chunk] = |w|^2.
// A right-hand side is skipped unless cyc[r] and (gate == nullptr or gate[r]).
__global__ __launch_bounds__(256) void k_dots(int64_t n, int nrhs, int nch, int kmax, int kcount,
                                              const float2* __restrict__ V, const float2* __restrict__ w,
                                              const int* __restrict__ cyc, const int* __restrict__ gate,
                                              double2* __restrict__ partial, double2* __restrict__ wnorm) {
    __shared__ double lds[8];
    const int chunk = blockIdx.x, r = blockIdx.z, k0 = blockIdx.y * KB;
    if (!cyc[r] || (gate && !gate[r])) return;
    const bool want_norm = wnorm != nullptr && blockIdx.y == 0;
    double wn = 0.0;
    const int kc = min(KB, kcount - k0);
    double2 acc[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) acc[q] = make_double2(0.0, 0.0);
    const int64_t beg = (int64_t)chunk * CHUNK, end = min(beg + CHUNK, n);
    const int64_t vstride = (int64_t)nrhs * n;
    const float2* wr = w + (int64_t)r * n;
    const float2* vr = V + (int64_t)k0 * vstride + (int64_t)r * n;
    for (int64_t p = beg + threadIdx.x; p < end; p += 256) {
        const float2 wv = wr[p];
        wn += (double)wv.x * wv.x + (double)wv.y * wv.y;
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            if (q < kc) {
                const float2 v = vr[(int64_t)q * vstride + p];
                acc[q].x += (double)v.x * wv.x + (double)v.y * wv.y;
                acc[q].y += (double)v.x * wv.y - (double)v.y * wv.x;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < KB; ++q) {
        if (q < kc) {
            double2 s = block_sum(acc[q], lds);
            if (threadIdx.x == 0) partial[((int64_t)r * kmax + k0 + q) * nch + chunk] = s;
        }
    }
    if (want_norm) {
        double2 s = block_sum(make_double2(wn, 0.0), lds);
        if (threadIdx.x == 0) wnorm[(int64_t)r * nch + chunk] = s;
    }
}

// coef[r][k] = sum over chunks; H column j accumulates it (pass 0 sets).
__global__ void k_reduce_dots(int nrhs, int nch, int kmax, int kcount, int m, int j, int pass,
                              const double2* __restrict__ partial, const int* __restrict__ cyc,
                              const int* __restrict__ gate, double2* __restrict__ coef, double2* __restrict__ H) {
    __shared__ double lds[8];
    const int r = blockIdx.y, k = blockIdx.x;
    if (!cyc[r] || k >= kcount || (gate && !gate[r])) return;
    double2 acc = make_double2(0.0, 0.0);
    const double2* src = partial + ((int64_t)r * kmax + k) * nch;
    for (int c = threadIdx.x; c < nch; c += blockDim.x) acc = cadd(acc, src[c]);
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) {
        coef[(int64_t)r * kmax + k] = acc;
        double2* hc = H + ((int64_t)r * m + j) * (m + 1);
        hc[k] = pass == 0 ? acc : cadd(hc[k], acc);
    }
}

// w[r] -= sum_k coef[r][k] V_k[r]; optionally partial[r*nch+chunk] = |w|^2
template <bool NORM>
__global__ __launch_bounds__(256) void k_update(int64_t n, int nrhs, int nch, int kmax, int kcount,
                                                const float2* __restrict__ V, float2* __restrict__ w,
                                                const double2* __restrict__ coef, const int* __restrict__ cyc,
                                                const int* __restrict__ gate, double2* __restrict__ partial) {
    __shared__ double lds[8];
    __shared__ float2 cf[64];
    const int chunk = blockIdx.x, r = blockIdx.y;
    if (!cyc[r] || (gate && !gate[r])) return;
    for (int k = threadIdx.x; k < kcount; k += 256) {
        double2 c = coef[(int64_t)r * kmax + k];
        cf[k] = make_float2((float)c.x, (float)c.y);
    }
    __syncthreads();
    const int64_t beg = (int64_t)chunk * CHUNK, end = min(beg + CHUNK, n);
    const int64_t vstride = (int64_t)nrhs * n;
    float2* wr = w + (int64_t)r * n;
    const float2* vr = V + (int64_t)r * n;
    double nrm = 0.0;
    for (int64_t p = beg + threadIdx.x; p < end; p += 256) {
        float2 acc = wr[p];
        for (int k = 0; k < kcount; ++k) {
            const float2 v = vr[(int64_t)k * vstride + p];
            const float2 c = cf[k];
            acc.x -= c.x * v.x - c.y * v.y;
            acc.y -= c.x * v.y + c.y * v.x;
        }
        wr[p] = acc;
        if (NORM) nrm += (double)acc.x * acc.x + (double)acc.y * acc.y;
    }
    if (NORM) {
        double2 s = block_sum(make_double2(nrm, 0.0), lds);
        if (threadIdx.x == 0) partial[(int64_t)r * nch + chunk] = s;
    }
}

// partial[r*nch+chunk] = |v[r]|^2 over the chunk (c128 or c64 vector)
template <typename C>
__global__ __launch_bounds__(256) void k_norm_partial(int64_t n, int nch, const C* __restrict__ v,
                                                      double2* __restrict__ partial) {
    __shared__ double lds[8];
    const int chunk = blockIdx.x, r = blockIdx.y;
    const int64_t beg = (int64_t)chunk * CHUNK, end = min(beg + CHUNK, n);
    const C* vr = v + (int64_t)r * n;
    double nrm = 0.0;
    for (int64_t p = beg + threadIdx.x; p < end; p += 256) {
        C a = vr[p];
        nrm += (double)a.x * a.x + (double)a.y * a.y;
    }
    double2 s = block_sum(make_double2(nrm, 0.0), lds);
    if (threadIdx.x == 0) partial[(int64_t)r * nch + chunk] = s;
}

__device__ double reduce_chunks(const double2* src, int nch, double* lds) {
    double2 acc = make_double2(0.0, 0.0);
    for (int c = threadIdx.x; c < nch; c += blockDim.x) acc = cadd(acc, src[c]);
    acc = block_sum(acc, lds);
    return acc.x;
}

// ||b_r|| (clamped like wave_solver.py:10)
__global__ void k_bnorm(int nch, const double2* __restrict__ partial, double* __restrict__ bnorm,
                        int* __restrict__ active, int* __restrict__ iters) {
    __shared__ double lds[8];
    const int r = blockIdx.x;
    double s = reduce_chunks(partial + (int64_t)r * nch, nch, lds);
    if (threadIdx.x == 0) {
        bnorm[r] = fmax(sqrt(s), 1e-30);
        active[r] = 1;
        iters[r] = 0;
    }
}

// Start of a cycle: beta = ||b - A x||, decide convergence, reset rotations.
__global__ void k_cycle_init(int nch, int m, double tol, const double2* __restrict__ partial,
                             const double* __restrict__ bnorm, double* __restrict__ beta, double* __restrict__ rel,
                             double* __restrict__ inv, int* __restrict__ active, int* __restrict__ cyc,
                             int* __restrict__ used, int* __restrict__ iters, double2* __restrict__ g) {
    __shared__ double lds[8];
    const int r = blockIdx.x;
    double s = reduce_chunks(partial + (int64_t)r * nch, nch, lds);
    if (threadIdx.x == 0) {
        double b = sqrt(s);
        beta[r] = b;
        double rr = b / bnorm[r];
        rel[r] = rr;
        iters[r] += used[r];
        used[r] = 0;
        if (!(rr >= tol)) active[r] = 0;  // converged (or NaN guard below)
        if (!isfinite(rr)) active[r] = 0;
        cyc[r] = active[r];
        inv[r] = b > 0 ? 1.0 / b : 0.0;
        g[(int64_t)r * (m + 1)] = make_double2(b, 0.0);
        for (int k = 1; k <= m; ++k) g[(int64_t)r * (m + 1) + k] = make_double2(0.0, 0.0);
    }
}

// v0[r] = r64[r] / beta_r  (c128 -> c64)
__global__ void k_scale_to32(int64_t n, const double2* __restrict__ src, float2* __restrict__ dst,
                             const double* __restrict__ inv, const int* __restrict__ cyc) {
    const int r = blockIdx.y;
    if (!cyc[r]) return;
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double2 a = src[(int64_t)r * n + p];
    double s = inv[r];
    dst[(int64_t)r * n + p] = make_float2((float)(a.x * s), (float)(a.y * s));
}

// v_{j+1}[r] = w[r] / h_{j+1,j}
__global__ void k_scale32(int64_t n, const float2* __restrict__ src, float2* __restrict__ dst,
                          const double* __restrict__ inv, const int* __restrict__ cyc) {
    const int r = blockIdx.y;
    if (!cyc[r]) return;
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float s = (float)inv[r];
    float2 a = src[(int64_t)r * n + p];
    dst[(int64_t)r * n + p] = make_float2(a.x * s, a.y * s);
}

// Kahan-Parlett "twice is enough": after one classical Gram-Schmidt pass
// w' = w - V c, a second pass is needed only if ||w'|| < 0.7 ||w|| (heavy
// cancellation).  Sets gate[r] for the second pass and stores ||w'||^2.
__global__ void k_reorth_check(int nch, const double2* __restrict__ wn_before, const double2* __restrict__ wn_after,
                               const int* __restrict__ cyc, int* __restrict__ gate, double* __restrict__ hn2) {
    __shared__ double lds[8];
    const int r = blockIdx.x;
    if (!cyc[r]) return;
    const double before = reduce_chunks(wn_before + (int64_t)r * nch, nch, lds);
    const double after = reduce_chunks(wn_after + (int64_t)r * nch, nch, lds);
    if (threadIdx.x == 0) {
        gate[r] = after < 0.49 * before ? 1 : 0;
        hn2[r] = after;
    }
}

// Finish column j: h_{j+1,j} = ||w||, apply old rotations, new rotation,
// residual estimate, early stop of the cycle.
__global__ void k_givens(int nch, int m, int j, double tol, const double2* __restrict__ partial,
                         const int* __restrict__ gate, const double* __restrict__ hn2,
                         const double* __restrict__ bnorm, double2* __restrict__ H, double* __restrict__ cs,
                         double2* __restrict__ sn, double2* __restrict__ g, double* __restrict__ est,
                         double* __restrict__ inv, int* __restrict__ cyc, int* __restrict__ used) {
    __shared__ double lds[8];
    const int r = blockIdx.x;
    if (!cyc[r]) return;
    double s = gate[r] ? reduce_chunks(partial + (int64_t)r * nch, nch, lds) : hn2[r];
    if (threadIdx.x != 0) return;
    const double hn = sqrt(s);
    double2* hc = H + ((int64_t)r * m + j) * (m + 1);
    double* c = cs + (int64_t)r * m;
    double2* sv = sn + (int64_t)r * m;
    double2* gr = g + (int64_t)r * (m + 1);
    hc[j + 1] = make_double2(hn, 0.0);
    for (int k = 0; k < j; ++k) {  // previous rotations
        double2 a = hc[k], b = hc[k + 1];
        double2 na = cadd(cscale(a, c[k]), cmul(sv[k], b));
        double2 nb = csub(cscale(b, c[k]), cmul(cconj(sv[k]), a));
        hc[k] = na;
        hc[k + 1] = nb;
    }
    double2 a = hc[j];
    double b = hn;
    double aa = sqrt(a.x * a.x + a.y * a.y);
    double rr = sqrt(aa * aa + b * b);
    double ck;
    double2 sk, rho;
    if (aa == 0.0) {
        ck = 0.0;
        sk = make_double2(1.0, 0.0);
        rho = make_double2(b, 0.0);
    } else {
        double2 ph = make_double2(a.x / aa, a.y / aa);
        ck = aa / rr;
        sk = cscale(ph, b / rr);
        rho = cscale(ph, rr);
    }
    c[j] = ck;
    sv[j] = sk;
    hc[j] = rho;
    hc[j + 1] = make_double2(0.0, 0.0);
    double2 gj = gr[j];
    gr[j] = cscale(gj, ck);
    gr[j + 1] = cscale(cmul(cconj(sk), gj), -1.0);
    double e = sqrt(gr[j + 1].x * gr[j + 1].x + gr[j + 1].y * gr[j + 1].y);
    est[r] = e / bnorm[r];
    used[r] = j + 1;
    inv[r] = hn > 0 ? 1.0 / hn : 0.0;
    // stop this right-hand side's cycle: converged estimate, happy breakdown
    // (wave_solver.py:29-30), or the basis is full
    if (e <= 0.5 * tol * bnorm[r] || hn < 1e-12 * bnorm[r] || !isfinite(e)) cyc[r] = 0;
}

// y = R^{-1} g (upper triangular, used[r] columns), stored in coef[r][k]
__global__ void k_backsolve(int nrhs, int m, int kmax, const double2* __restrict__ H,
                            const double2* __restrict__ g, const int* __restrict__ used, double2* __restrict__ y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrhs) return;
    const int u = used[r];
    const double2* Hr = H + (int64_t)r * m * (m + 1);
    double2* yr = y + (int64_t)r * kmax;
    for (int k = u - 1; k >= 0; --k) {
        double2 acc = g[(int64_t)r * (m + 1) + k];
        for (int q = k + 1; q < u; ++q) acc = csub(acc, cmul(Hr[(int64_t)q * (m + 1) + k], yr[q]));
        double2 d = Hr[(int64_t)k * (m + 1) + k];
        double dd = d.x * d.x + d.y * d.y;
        yr[k] = dd > 0 ? make_double2((acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd)
                       : make_double2(0.0, 0.0);
    }
}

// x[r] += sum_{k<used[r]} y_k Z_k[r]   (FP64 accumulate into c128 x)
__global__ __launch_bounds__(256) void k_xupdate(int64_t n, int nrhs, int kmax, const float2* __restrict__ Z,
                                                 const double2* __restrict__ y, const int* __restrict__ used,
                                                 double2* __restrict__ x) {
    __shared__ double2 yy[64];
    const int r = blockIdx.y;
    const int u = used[r];
    if (u == 0) return;
    for (int k = threadIdx.x; k < u; k += blockDim.x) yy[k] = y[(int64_t)r * kmax + k];
    __syncthreads();
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t vstride = (int64_t)nrhs * n;
    const float2* zr = Z + (int64_t)r * n;
    double2 acc = x[(int64_t)r * n + p];
    for (int k = 0; k < u; ++k) {
        float2 z = zr[(int64_t)k * vstride + p];
        double2 c = yy[k];
        acc.x += c.x * z.x - c.y * z.y;
        acc.y += c.x * z.y + c.y * z.x;
    }
    x[(int64_t)r * n + p] = acc;
}

__global__ void k_zero_c128(int64_t count, double2* p) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = make_double2(0.0, 0.0);
}

template <typename T>
static int32_t alloc(T** ptr, size_t count) {
    if (*ptr) {
        (void)hipFree(*ptr);
        *ptr = nullptr;
    }
    hipError_t e = hipMalloc((void**)ptr, count * sizeof(T) + 256);
    if (e != hipSuccess) {
        *ptr = nullptr;
        set_error(std::string("hipMalloc (solver workspace): ") + hipGetErrorString(e));
        return HELM_ERR_ALLOC;
    }
    return HELM_OK;
}

int32_t ensure_workspace(helm_plan* p, int nrhs, int restart) {
    Workspace& w = p->ws;
    const int64_t n = p->levels[0].n;
    if (w.V && w.nrhs >= nrhs && w.restart >= restart && w.n == n) return HELM_OK;
    nrhs = std::max(nrhs, w.nrhs);
    restart = std::max(restart, w.restart);
    const int nch = nchunks(n);
    const int kmax = restart + 1;
    w.nrhs = nrhs;
    w.restart = restart;
    w.n = n;
    HELM_TRY(alloc(&w.V, (size_t)(restart + 1) * nrhs * n));
    HELM_TRY(alloc(&w.Z, (size_t)restart * nrhs * n));
    HELM_TRY(alloc(&w.w, (size_t)nrhs * n));
    HELM_TRY(alloc(&w.r64, (size_t)nrhs * n));
    w.partial_len = (int64_t)nrhs * (kmax + 3) * nch;  // dots | |w|^2 | |w'|^2 | |w''|^2
    HELM_TRY(alloc(&w.partial, (size_t)w.partial_len));
    HELM_TRY(alloc(&w.H, (size_t)nrhs * restart * (restart + 1)));
    HELM_TRY(alloc(&w.cs, (size_t)nrhs * restart));
    HELM_TRY(alloc(&w.sn, (size_t)nrhs * restart));
    HELM_TRY(alloc(&w.g, (size_t)nrhs * (restart + 1)));
    HELM_TRY(alloc(&w.coef, (size_t)nrhs * kmax));
    HELM_TRY(alloc(&w.beta, nrhs));
    HELM_TRY(alloc(&w.bnorm, nrhs));
    HELM_TRY(alloc(&w.rel, nrhs));
    HELM_TRY(alloc(&w.inv, nrhs));
    HELM_TRY(alloc(&w.est, nrhs));
    HELM_TRY(alloc(&w.active, nrhs));
    HELM_TRY(alloc(&w.cyc, nrhs));
    HELM_TRY(alloc(&w.used, nrhs));
    HELM_TRY(alloc(&w.iters, nrhs));
    HELM_TRY(alloc(&w.gate, nrhs));
    HELM_TRY(alloc(&w.hn2, nrhs));
    w.h_rel.assign(nrhs, 0.0);
    w.h_active.assign(nrhs, 0);
    w.h_iters.assign(nrhs, 0);
    return HELM_OK;
}

// ---- host driver -----------------------------------------------------------
int32_t solve_gmres(helm_plan* p, const double2* b, double2* x, int nrhs, const helm_solve_opts* o,
                    helm_solve_info* info, hipStream_t s) {
    const int m = o->restart;
    HELM_TRY(ensure_workspace(p, nrhs, m));
    Workspace& W = p->ws;
    const Level& L = p->levels[0];
    const int64_t n = L.n;
    const int nch = nchunks(n);
    const int kmax = W.restart + 1;
    const int mm = W.restart;  // row stride of H / rotations (capacity)
    const double k0sq = p->k0 * p->k0;
    const bool adj = o->adjoint != 0;
    const int pc = o->precond;
    const dim3 blk(256);
    const dim3 gvec((unsigned)((n + 255) / 256), nrhs);
    const int64_t vstride = (int64_t)nrhs * n;
    double2* pw0 = W.partial + (int64_t)nrhs * kmax * nch;  // |w|^2 before projection
    double2* pw1 = pw0 + (int64_t)nrhs * nch;               // after the first pass
    double2* pw2 = pw1 + (int64_t)nrhs * nch;               // after the second pass

    // ||b|| and initial x
    hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nch, nrhs), blk, 0, s, n, nch, b, W.partial);
    hipLaunchKernelGGL(k_bnorm, dim3(nrhs), dim3(64), 0, s, nch, W.partial, W.bnorm, W.active, W.iters);
    // used[] must start at zero for the first k_cycle_init
    HELM_HIP(hipMemsetAsync(W.used, 0, nrhs * sizeof(int), s));
    if (!o->use_x0) {
        int64_t cnt = (int64_t)nrhs * n;
        hipLaunchKernelGGL(k_zero_c128, dim3((unsigned)((cnt + 255) / 256)), blk, 0, s, cnt, x);
    }
    HELM_HIP(hipGetLastError());

    int steps = 0, cycles = 0, applications = 0;
    bool budget = false;
    for (;;) {
        // true residual in FP64
        HELM_TRY(launch_apply64(L, k0sq, x, W.r64, nrhs, adj, b, s));
        applications++;
        hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nch, nrhs), blk, 0, s, n, nch, W.r64, W.partial);
        hipLaunchKernelGGL(k_cycle_init, dim3(nrhs), dim3(64), 0, s, nch, mm, o->tolerance, W.partial, W.bnorm,
                           W.beta, W.rel, W.inv, W.active, W.cyc, W.used, W.iters, W.g);
        HELM_HIP(hipGetLastError());
        HELM_HIP(hipMemcpyAsync(W.h_rel.data(), W.rel, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        HELM_HIP(hipMemcpyAsync(W.h_active.data(), W.active, nrhs * sizeof(int), hipMemcpyDeviceToHost, s));
        HELM_HIP(hipMemcpyAsync(W.h_iters.data(), W.iters, nrhs * sizeof(int), hipMemcpyDeviceToHost, s));
        HELM_HIP(hipStreamSynchronize(s));
        if (o->callback) o->callback(o->user, applications, W.h_rel.data(), nrhs);
        int n_active = 0;
        for (int r = 0; r < nrhs; ++r) n_active += W.h_active[r];
        if (n_active == 0) break;
        if (steps >= o->max_iterations) {
            budget = true;
            break;
        }
        cycles++;
        hipLaunchKernelGGL(k_scale_to32, gvec, blk, 0, s, n, W.r64, W.V, W.inv, W.cyc);
        const int msteps = std::min(m, o->max_iterations - steps);
        for (int j = 0; j < msteps; ++j) {
            float2* vj = W.V + (int64_t)j * vstride;
            float2* zj = (pc == HELM_PC_NONE) ? vj : W.Z + (int64_t)j * vstride;
            if (pc != HELM_PC_NONE) HELM_TRY(precond_apply(p, pc, vj, zj, nrhs, adj, s));
            HELM_TRY(launch_apply32(L, k0sq, zj, W.w, nrhs, adj, nullptr, 0.0f, s));
            applications++;
            const int kc = j + 1;
            const dim3 gd(nch, (kc + KB - 1) / KB, nrhs);
            // pass 1: c = V^H w (and |w|^2), w -= V c (and |w'|^2)
            hipLaunchKernelGGL(k_dots, gd, blk, 0, s, n, nrhs, nch, kmax, kc, W.V, W.w, W.cyc, (const int*)nullptr,
                               W.partial, pw0);
            hipLaunchKernelGGL(k_reduce_dots, dim3(kc, nrhs), dim3(64), 0, s, nrhs, nch, kmax, kc, mm, j, 0,
                               W.partial, W.cyc, (const int*)nullptr, W.coef, W.H);
            hipLaunchKernelGGL(k_update<true>, dim3(nch, nrhs), blk, 0, s, n, nrhs, nch, kmax, kc, W.V, W.w, W.coef,
                               W.cyc, (const int*)nullptr, pw1);
            hipLaunchKernelGGL(k_reorth_check, dim3(nrhs), dim3(64), 0, s, nch, pw0, pw1, W.cyc, W.gate, W.hn2);
            // pass 2 only where the first pass cancelled heavily (gate[r])
            hipLaunchKernelGGL(k_dots, gd, blk, 0, s, n, nrhs, nch, kmax, kc, W.V, W.w, W.cyc, W.gate, W.partial,
                               (double2*)nullptr);
            hipLaunchKernelGGL(k_reduce_dots, dim3(kc, nrhs), dim3(64), 0, s, nrhs, nch, kmax, kc, mm, j, 1,
                               W.partial, W.cyc, W.gate, W.coef, W.H);
            hipLaunchKernelGGL(k_update<true>, dim3(nch, nrhs), blk, 0, s, n, nrhs, nch, kmax, kc, W.V, W.w, W.coef,
                               W.cyc, W.gate, pw2);
            hipLaunchKernelGGL(k_givens, dim3(nrhs), dim3(64), 0, s, nch, mm, j, o->tolerance, pw2, W.gate, W.hn2,
                               W.bnorm, W.H, W.cs, W.sn, W.g, W.est, W.inv, W.cyc, W.used);
            if (j + 1 < m)
                hipLaunchKernelGGL(k_scale32, gvec, blk, 0, s, n, W.w, W.V + (int64_t)(j + 1) * vstride, W.inv,
                                   W.cyc);
            HELM_HIP(hipGetLastError());
            steps++;
        }
        hipLaunchKernelGGL(k_backsolve, dim3((nrhs + 63) / 64), dim3(64), 0, s, nrhs, mm, kmax, W.H, W.g, W.used,
                           W.coef);
        hipLaunchKernelGGL(k_xupdate, gvec, blk, 0, s, n, nrhs, kmax, pc == HELM_PC_NONE ? W.V : W.Z, W.coef,
                           W.used, x);
        HELM_HIP(hipGetLastError());
    }
    for (int r = 0; r < nrhs; ++r) {
        info[r].reason = W.h_active[r] ? HELM_ITERATION_BUDGET : HELM_CONVERGED;
        if (!std::isfinite(W.h_rel[r])) info[r].reason = HELM_BREAKDOWN;
        info[r].iterations = W.h_iters[r];
        info[r].applications = W.h_iters[r] + cycles + 1;
        info[r].cycles = cycles;
        info[r].relative_residual = W.h_rel[r];
    }
    (void)budget;
    return HELM_OK;
}

}  // namespace helm

using namespace helm;

extern "C" int32_t helm_solve(helm_plan* p, const void* rhs, void* x, int32_t nrhs, const helm_solve_opts* o,
                              helm_solve_info* info, void* stream) {
    HELM_CHECK(p && rhs && x && o && info, HELM_ERR_ARG, "helm_solve: null argument");
    HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_solve: call helm_set_medium first");
    HELM_CHECK(nrhs >= 1 && nrhs <= 65535 && (int64_t)nrhs * p->levels[0].n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_solve: nrhs out of range");
    HELM_CHECK(o->restart >= 1 && o->restart <= 63, HELM_ERR_ARG, "helm_solve: restart must be in [1, 63]");
    HELM_CHECK(o->max_iterations >= 1, HELM_ERR_ARG, "helm_solve: max_iterations must be positive");
    HELM_CHECK(std::isfinite(o->tolerance) && o->tolerance > 0, HELM_ERR_ARG,
               "helm_solve: tolerance must be finite and positive");
    HELM_CHECK(o->precond >= HELM_PC_NONE && o->precond <= HELM_PC_LEARNED, HELM_ERR_ARG,
               "helm_solve: unknown preconditioner");
    HELM_CHECK(o->precond != HELM_PC_MG || p->mg.levels > 0, HELM_ERR_STATE,
               "helm_solve: call helm_mg_setup before using the multigrid preconditioner");
    hipStream_t s = (hipStream_t)stream;
    if (o->method == HELM_GMRES) return solve_gmres(p, (const double2*)rhs, (double2*)x, nrhs, o, info, s);
    if (o->method == HELM_BICGSTAB) return solve_bicgstab(p, (const double2*)rhs, (double2*)x, nrhs, o, info, s);
    set_error("helm_solve: unknown method");
    return HELM_ERR_ARG;
}
