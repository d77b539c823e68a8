// Restarted (F)GMRES(m) with FP64 inner products, FP64 Givens least squares
// and FP64 true-residual refinement at every restart.
//
// Algorithm: nsrno/wave_solver.py:9-36 (right preconditioning, classical
// Gram-Schmidt applied twice, convergence decided on the true residual at
// each restart, reasons converged / iteration_budget), with these changes
// made for the GPU:
//   * delayed classical Gram-Schmidt (DCGS2): the reorthogonalisation of V_j
//     is folded into the projection of the next vector, so each step reads
//     the basis twice (one fused dot pass, one fused update pass) instead of
//     four times, with CGS2's orthogonality; the Hessenberg column of a step
//     is finalised (corrected, then rotated) one step later;
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
#include <cstdlib>
#include <cstring>

#include <hip/hip_ext.h>

#include "helm_coef.h"
#include "helm_internal.h"

namespace helm {

constexpr int CHUNK = 4096;  // nodes per reduction chunk (256 threads x 16)

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

__device__ double reduce_chunks(const double2* src, int nch, double* lds) {
    double2 acc = make_double2(0.0, 0.0);
    for (int c = threadIdx.x; c < nch; c += blockDim.x) acc = cadd(acc, src[c]);
    acc = block_sum(acc, lds);
    return acc.x;
}

// ---- delayed classical Gram-Schmidt (DCGS2) --------------------------------
// Step j holds V_0..V_{j-1} final and V_j orthogonalised once (provisional).
// One read of the basis computes both the reorthogonalisation coefficients
// of V_j, c_k = <V_k, V_j> (k < j, and <V_j, V_j>), and the projections of
// w = A M V_j, d_k = <V_k, w> (k <= j).  A second read updates both vectors:
//   V_j <- (V_j - sum c_k V_k) / alpha,          alpha^2 = |V_j|^2 - sum |c_k|^2
//   w   <- w - sum_{k<j} d_k V_k - h_jj V_j,     h_jj = (d_j - sum conj(c_k) d_k) / alpha
// and the previous Hessenberg column is corrected for the change of V_j,
//   H[k][j-1] += H[j][j-1] c_k,   H[j][j-1] *= alpha,
// after which its Givens rotation is applied (one step later than plain
// Arnoldi).  Two basis reads per step instead of four, numerically a CGS2.

// pc[(r*kmax + k)*nch + chunk] = <V_k, V_j>, pd[...] = <V_k, w>, k <= j.
// One workgroup per (chunk of DOT_CHUNK nodes, RHS): its w and V_j are read
// once into registers (8 nodes per thread), then V_0..V_{j-1} stream through with the
// next vector prefetched while the current one is reduced, so a launch reads
// exactly j + 2 vectors.  Per k the four sums are reduced within each wave
// (shuffles); the four waves' results meet in LDS once, after the k loop.
// nodes per thread: 8 (2048-node chunks, 3 waves per SIMD) or 4 (1024-node
// chunks, twice the workgroups and partial sums, more waves); HELM_DOT_NPT
static int dot_npt() {
    static const int t = [] {
        const char* v = std::getenv("HELM_DOT_NPT");
        const int x = v ? std::atoi(v) : 0;
        return x == 4 ? 4 : 8;
    }();
    return t;
}
// restart residual of the flux operator in its block form (HELM_RESID_BLOCK=0: the per-node kernel)
constexpr int RES_RB = 2, RES_TZ = 2;
static bool resid_block() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_RESID_BLOCK");
        return !(v && std::atoi(v) == 0);
    }();
    return t;
}
// V_{j+1} left unnormalised and scaled by the preconditioner as it loads it
// (multigrid on the fused path only; HELM_LAZY_BASIS=0: a scaling pass per step)
static bool lazy_basis() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_LAZY_BASIS");
        return !(v && std::atoi(v) == 0);
    }();
    return t;
}
// next restart cycle queued before the host reads the status (HELM_RESTART_AHEAD=0: after)
static bool restart_ahead() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_RESTART_AHEAD");
        return !(v && std::atoi(v) == 0);
    }();
    return t;
}
// fused reduce+update and norm+scale launches (HELM_DCGS_FUSED=0: the separate kernels)
static bool dcgs_fused() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_DCGS_FUSED");
        return !(v && std::atoi(v) == 0);
    }();
    return t;
}
static inline int ndotchunks(int64_t n) { return (int)((n + 256 * dot_npt() - 1) / (256 * dot_npt())); }
// partial-sum capacity: enough chunks for the smaller chunk size
static inline int ndotchunks_max(int64_t n) { return (int)((n + 1023) / 1024); }
template <int DOT_NPT>
__device__ __forceinline__ void k_dcgs_dots_body(const VBlock& B, int64_t n, int nrhs, int nch, int kmax,
                                                                          int j,
                                                   const float2* __restrict__ V, const float2* __restrict__ w,
                                                   const int* __restrict__ cyc, double2* __restrict__ pc,
                                                   double2* __restrict__ pd) {
    __shared__ double4 red[64][4];  // [k][wave] = (c.re, c.im, d.re, d.im)
    const int chunk = B.x, r = B.y;
    if (!cyc[r]) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t beg = (int64_t)chunk * (256 * DOT_NPT);
    const int64_t vstride = (int64_t)nrhs * n;
    const float2* wr = w + (int64_t)r * n;
    const float2* vr = V + (int64_t)r * n;
    const float2* vj = vr + (int64_t)j * vstride;
    float2 wv[DOT_NPT], jv[DOT_NPT], cur[DOT_NPT], nxt[DOT_NPT];
    const float2 zero = make_float2(0.f, 0.f);
    // every load of the chunk is issued before the first is used (addresses
    // clamped, out-of-range values zeroed after): one memory latency, where
    // loads behind bounds branches are waited for element by element.  For
    // j = 0, vr is vj, so cur = V_0 = jv as the products below need.
#pragma unroll
    for (int i = 0; i < DOT_NPT; ++i) {
        const int64_t p = beg + threadIdx.x + 256 * i;
        const int64_t pc = p < n ? p : 0;
        wv[i] = wr[pc];
        jv[i] = vj[pc];
        cur[i] = vr[pc];
    }
#pragma unroll
    for (int i = 0; i < DOT_NPT; ++i) {
        if (beg + threadIdx.x + 256 * i >= n) {
            wv[i] = zero;
            jv[i] = zero;
            cur[i] = zero;
        }
    }
    for (int k = 0; k <= j; ++k) {
        if (k + 1 < j) {
            const float2* vn = vr + (int64_t)(k + 1) * vstride;
#pragma unroll
            for (int i = 0; i < DOT_NPT; ++i) {
                const int64_t p = beg + threadIdx.x + 256 * i;
                nxt[i] = vn[p < n ? p : 0];
            }
#pragma unroll
            for (int i = 0; i < DOT_NPT; ++i)
                if (beg + threadIdx.x + 256 * i >= n) nxt[i] = zero;
        } else {
#pragma unroll
            for (int i = 0; i < DOT_NPT; ++i) nxt[i] = jv[i];
        }
        double cr = 0.0, ci = 0.0, dr = 0.0, di = 0.0;
#pragma unroll
        for (int i = 0; i < DOT_NPT; ++i) {
            const double vx = cur[i].x, vy = cur[i].y;
            cr += vx * jv[i].x + vy * jv[i].y;
            ci += vx * jv[i].y - vy * jv[i].x;
            dr += vx * wv[i].x + vy * wv[i].y;
            di += vx * wv[i].y - vy * wv[i].x;
        }
        cr = wave_sum(cr);
        ci = wave_sum(ci);
        dr = wave_sum(dr);
        di = wave_sum(di);
        if (lane == 0) red[k][wid] = make_double4(cr, ci, dr, di);
#pragma unroll
        for (int i = 0; i < DOT_NPT; ++i) cur[i] = nxt[i];
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= j; k += blockDim.x) {
        double4 t = red[k][0];
        for (int q = 1; q < 4; ++q) {
            const double4 u = red[k][q];
            t.x += u.x;
            t.y += u.y;
            t.z += u.z;
            t.w += u.w;
        }
        pc[((int64_t)r * kmax + k) * nch + chunk] = make_double2(t.x, t.y);
        pd[((int64_t)r * kmax + k) * nch + chunk] = make_double2(t.z, t.w);
    }
}
template <int DOT_NPT>
__global__ __launch_bounds__(256, DOT_NPT == 8 ? 3 : 4) void k_dcgs_dots(int64_t n, int nrhs, int nch, int kmax,
                                                                          int j,
                                                   const float2* __restrict__ V, const float2* __restrict__ w,
                                                   const int* __restrict__ cyc, double2* __restrict__ pc,
                                                   double2* __restrict__ pd) {
    k_dcgs_dots_body<DOT_NPT>(hw_block(), n, nrhs, nch, kmax, j, V, w, cyc, pc, pd);
}

// apply rotations 0..c-1 to column c, create rotation c, update g; returns |g_{c+1}|
__device__ double givens_column(double2* hc, double* cs, double2* sn, double2* g, int c) {
    for (int k = 0; k < c; ++k) {
        const double2 a = hc[k], b = hc[k + 1];
        hc[k] = cadd(cscale(a, cs[k]), cmul(sn[k], b));
        hc[k + 1] = csub(cscale(b, cs[k]), cmul(cconj(sn[k]), a));
    }
    const double2 a = hc[c];
    const double b = hc[c + 1].x;  // real, nonnegative (a norm)
    const double aa = sqrt(a.x * a.x + a.y * a.y);
    const double rr = sqrt(aa * aa + b * b);
    double ck;
    double2 sk, rho;
    if (aa == 0.0) {
        ck = 0.0;
        sk = make_double2(1.0, 0.0);
        rho = make_double2(b, 0.0);
    } else {
        const double2 ph = make_double2(a.x / aa, a.y / aa);
        ck = aa / rr;
        sk = cscale(ph, b / rr);
        rho = cscale(ph, rr);
    }
    cs[c] = ck;
    sn[c] = sk;
    hc[c] = rho;
    hc[c + 1] = make_double2(0.0, 0.0);
    const double2 gc = g[c];
    g[c] = cscale(gc, ck);
    g[c + 1] = cscale(cmul(cconj(sk), gc), -1.0);
    return sqrt(g[c + 1].x * g[c + 1].x + g[c + 1].y * g[c + 1].y);
}

// Per RHS: reduce the partials, correct and finalise column j-1 (Givens,
// residual estimate, early stop), build provisional column j and the update
// coefficients cf[r][0..j-1] = c_k, cf[r][kmax + k] = d_k, cf[r][2kmax] =
// (h_jj, 1/alpha).
__global__ void k_dcgs_reduce(int nch, int m, int kmax, int j, double tol, const double2* __restrict__ pc,
                              const double2* __restrict__ pd, const double* __restrict__ bnorm,
                              double2* __restrict__ H, double* __restrict__ cs, double2* __restrict__ sn,
                              double2* __restrict__ g, double* __restrict__ est, int* __restrict__ cyc,
                              int* __restrict__ used, double2* __restrict__ cf) {
    __shared__ double lds[8];
    __shared__ double2 sc[64], sd[64];
    const int r = blockIdx.x;
    if (!cyc[r]) return;
    (void)lds;
    // one wave per k: sum the chunk partials in a fixed order (deterministic)
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int k = threadIdx.x >> 6; k <= j; k += nw) {
        double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
        const double2* srcc = pc + ((int64_t)r * kmax + k) * nch;
        const double2* srcd = pd + ((int64_t)r * kmax + k) * nch;
        for (int c = lane; c < nch; c += 64) {
            a = cadd(a, srcc[c]);
            b = cadd(b, srcd[c]);
        }
        a.x = wave_sum(a.x);
        a.y = wave_sum(a.y);
        b.x = wave_sum(b.x);
        b.y = wave_sum(b.y);
        if (lane == 0) {
            sc[k] = a;
            sd[k] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double2* Hr = H + (int64_t)r * m * (m + 1);
    double* csr = cs + (int64_t)r * m;
    double2* snr = sn + (int64_t)r * m;
    double2* gr = g + (int64_t)r * (m + 1);
    double2* cfr = cf + (int64_t)r * (2 * kmax + 1);
    double alpha = 1.0;
    if (j > 0) {
        double c2 = 0.0;
        for (int k = 0; k < j; ++k) c2 += sc[k].x * sc[k].x + sc[k].y * sc[k].y;
        alpha = sqrt(fmax(sc[j].x - c2, 1e-300));
        // correct column j-1 for V_j = alpha V_j'' + sum c_k V_k, then finalise it
        double2* hp = Hr + (int64_t)(j - 1) * (m + 1);
        const double hn = hp[j].x;
        for (int k = 0; k < j; ++k) hp[k] = cadd(hp[k], cscale(sc[k], hn));
        hp[j] = make_double2(hn * alpha, 0.0);
        const double e = givens_column(hp, csr, snr, gr, j - 1);
        est[r] = e / bnorm[r];
        used[r] = j;
        if (e <= 0.5 * tol * bnorm[r] || !isfinite(e)) {
            cyc[r] = 0;
            return;
        }
    }
    // provisional column j w.r.t. the final V_0..V_{j-1} and V_j''
    double2* hc = Hr + (int64_t)j * (m + 1);
    double2 hjj = sd[j];
    for (int k = 0; k < j; ++k) {
        hc[k] = sd[k];
        hjj = csub(hjj, cmul(cconj(sc[k]), sd[k]));
        cfr[k] = sc[k];
        cfr[kmax + k] = sd[k];
    }
    hjj = cscale(hjj, 1.0 / alpha);
    hc[j] = hjj;
    cfr[2 * kmax] = make_double2(1.0 / alpha, 0.0);
    cfr[2 * kmax - 1] = hjj;  // slot kmax - 1 + kmax: h_jj (j < kmax - 1 always)
}

// V_j <- (V_j - sum c_k V_k)/alpha ; w <- w - sum d_k V_k - h_jj V_j ; |w|^2 partials
__global__ __launch_bounds__(256) void k_dcgs_update(int64_t n, int nrhs, int nch, int kmax, int j,
                                                     float2* __restrict__ V, float2* __restrict__ w,
                                                     const double2* __restrict__ cf, const int* __restrict__ cyc,
                                                     double2* __restrict__ partial) {
    __shared__ double lds[8];
    __shared__ float2 c_s[64], d_s[64];
    __shared__ float2 hjj_s;
    __shared__ float ialpha_s;
    const int chunk = blockIdx.x, r = blockIdx.y;
    if (!cyc[r]) return;
    const double2* cfr = cf + (int64_t)r * (2 * kmax + 1);
    for (int k = threadIdx.x; k < j; k += 256) {
        c_s[k] = make_float2((float)cfr[k].x, (float)cfr[k].y);
        d_s[k] = make_float2((float)cfr[kmax + k].x, (float)cfr[kmax + k].y);
    }
    if (threadIdx.x == 0) {
        hjj_s = make_float2((float)cfr[2 * kmax - 1].x, (float)cfr[2 * kmax - 1].y);
        ialpha_s = (float)cfr[2 * kmax].x;
    }
    __syncthreads();
    const int64_t beg = (int64_t)chunk * CHUNK, end = min(beg + CHUNK, n);
    const int64_t vstride = (int64_t)nrhs * n;
    float2* wr = w + (int64_t)r * n;
    float2* vj = V + (int64_t)j * vstride + (int64_t)r * n;
    const float2* vr = V + (int64_t)r * n;
    const float2 hjj = hjj_s;
    const float ia = ialpha_s;
    double nrm = 0.0;
    for (int64_t p = beg + threadIdx.x; p < end; p += 256) {
        float2 sc = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
        for (int k = 0; k < j; ++k) {
            const float2 v = vr[(int64_t)k * vstride + p];
            cfma(sc, c_s[k], v);
            cfma(sd, d_s[k], v);
        }
        float2 jv = csub(vj[p], sc);
        jv = make_float2(jv.x * ia, jv.y * ia);
        vj[p] = jv;
        float2 wv = csub(wr[p], sd);
        wv.x -= hjj.x * jv.x - hjj.y * jv.y;
        wv.y -= hjj.x * jv.y + hjj.y * jv.x;
        wr[p] = wv;
        nrm += (double)wv.x * wv.x + (double)wv.y * wv.y;
    }
    const double2 tot = block_sum(make_double2(nrm, 0.0), lds);
    if (threadIdx.x == 0) partial[(int64_t)r * nch + chunk] = tot;
}

// provisional H[j+1][j] = ||w||, inv = 1/||w||; a happy breakdown (w = 0)
// finalises column j at once and stops the cycle
__global__ void k_dcgs_norm(int nch, int m, int j, const double2* __restrict__ partial,
                            const double* __restrict__ bnorm, double2* __restrict__ H, double* __restrict__ cs,
                            double2* __restrict__ sn, double2* __restrict__ g, double* __restrict__ est,
                            double* __restrict__ inv, int* __restrict__ cyc, int* __restrict__ used) {
    __shared__ double lds[8];
    const int r = blockIdx.x;
    if (!cyc[r]) return;
    const double s = reduce_chunks(partial + (int64_t)r * nch, nch, lds);
    if (threadIdx.x != 0) return;
    const double hn = sqrt(s);
    double2* hc = H + ((int64_t)r * m + j) * (m + 1);
    hc[j + 1] = make_double2(hn, 0.0);
    inv[r] = hn > 0 ? 1.0 / hn : 0.0;
    if (hn < 1e-12 * bnorm[r]) {
        const double e = givens_column(hc, cs + (int64_t)r * m, sn + (int64_t)r * m, g + (int64_t)r * (m + 1), j);
        est[r] = e / bnorm[r];
        used[r] = j + 1;
        cyc[r] = 0;
    }
}

// end of a cycle: finalise the last provisional column (no reorthogonalisation
// of the unused V_{j+1}) for every RHS still in the cycle
__global__ void k_dcgs_finish(int nrhs, int m, int j, double2* __restrict__ H, double* __restrict__ cs,
                              double2* __restrict__ sn, double2* __restrict__ g, double* __restrict__ est,
                              const double* __restrict__ bnorm, int* __restrict__ cyc, int* __restrict__ used) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrhs || !cyc[r]) return;
    double2* hc = H + ((int64_t)r * m + j) * (m + 1);
    const double e = givens_column(hc, cs + (int64_t)r * m, sn + (int64_t)r * m, g + (int64_t)r * (m + 1), j);
    est[r] = e / bnorm[r];
    used[r] = j + 1;
    cyc[r] = 0;
}

// ---- fused forms: two launches fewer per step ------------------------------
// k_dcgs_reduce_update = k_dcgs_reduce + k_dcgs_update: every workgroup sums
// the dots partials of its right-hand side itself, in k_dcgs_reduce's fixed
// order (so every workgroup derives bit-identical coefficients), and the
// chunk-0 workgroup alone does the Hessenberg bookkeeping (correct and rotate
// column j-1, residual estimate, early stop, provisional column j).  A
// right-hand side that stops here is still updated by workgroups that read
// cyc before the stop: harmless, as its V_j and w are never read again (the
// x update uses Z_0..Z_{j-1}).  Extra reads per workgroup: 2 (j + 1) nchd
// partials (31 KB at j = 5, from L2).
__device__ __forceinline__ void k_dcgs_reduce_update_body(const VBlock& B, int64_t n, int nrhs, int nch, int nchd, int m, int kmax,
                                                            int j, double tol, const double2* __restrict__ pc,
                                                            const double2* __restrict__ pd,
                                                            const double* __restrict__ bnorm, double2* __restrict__ H,
                                                            double* __restrict__ cs, double2* __restrict__ sn,
                                                            double2* __restrict__ g, double* __restrict__ est,
                                                            int* __restrict__ cyc, int* __restrict__ used,
                                                            float2* __restrict__ V, float2* __restrict__ w,
                                                            double2* __restrict__ partial, int v0_raw) {
    __shared__ double lds[8];
    __shared__ double2 sc[64], sd[64];
    __shared__ float2 c_s[64], d_s[64];
    __shared__ float2 hjj_s;
    __shared__ float ialpha_s;
    const int chunk = B.x, r = B.y;
    if (!cyc[r]) return;
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int k = threadIdx.x >> 6; k <= j; k += nw) {
        double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
        const double2* srcc = pc + ((int64_t)r * kmax + k) * nchd;
        const double2* srcd = pd + ((int64_t)r * kmax + k) * nchd;
        // four strides of partials in flight per wait, summed in the order
        // c = lane, lane + 64, ... as one at a time
        for (int c0 = lane; c0 < nchd; c0 += 256) {
            double2 ta[4], tb[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = c0 + 64 * q < nchd ? c0 + 64 * q : 0;
                ta[q] = srcc[c];
                tb[q] = srcd[c];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (c0 + 64 * q < nchd) {
                    a = cadd(a, ta[q]);
                    b = cadd(b, tb[q]);
                }
        }
        a.x = wave_sum(a.x);
        a.y = wave_sum(a.y);
        b.x = wave_sum(b.x);
        b.y = wave_sum(b.y);
        if (lane == 0) {
            sc[k] = a;
            sd[k] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double alpha = 1.0;
        bool stop = false;
        double2* Hr = H + (int64_t)r * m * (m + 1);
        if (j == 0 && v0_raw) alpha = sqrt(fmax(sc[0].x, 1e-300));  // V_0 = r, not r / beta
        if (j > 0) {
            double c2 = 0.0;
            for (int k = 0; k < j; ++k) c2 += sc[k].x * sc[k].x + sc[k].y * sc[k].y;
            alpha = sqrt(fmax(sc[j].x - c2, 1e-300));
            if (chunk == 0) {
                double2* hp = Hr + (int64_t)(j - 1) * (m + 1);
                const double hn = hp[j].x;
                for (int k = 0; k < j; ++k) hp[k] = cadd(hp[k], cscale(sc[k], hn));
                hp[j] = make_double2(hn * alpha, 0.0);
                const double e = givens_column(hp, cs + (int64_t)r * m, sn + (int64_t)r * m, g + (int64_t)r * (m + 1),
                                               j - 1);
                est[r] = e / bnorm[r];
                used[r] = j;
                if (e <= 0.5 * tol * bnorm[r] || !isfinite(e)) {
                    cyc[r] = 0;
                    stop = true;
                }
            }
        }
        double2 hjj = sd[j];
        for (int k = 0; k < j; ++k) hjj = csub(hjj, cmul(cconj(sc[k]), sd[k]));
        hjj = cscale(hjj, 1.0 / alpha);
        if (chunk == 0 && !stop) {
            double2* hc = Hr + (int64_t)j * (m + 1);
            for (int k = 0; k < j; ++k) hc[k] = sd[k];
            hc[j] = hjj;
        }
        for (int k = 0; k < j; ++k) {
            c_s[k] = make_float2((float)sc[k].x, (float)sc[k].y);
            d_s[k] = make_float2((float)sd[k].x, (float)sd[k].y);
        }
        hjj_s = make_float2((float)hjj.x, (float)hjj.y);
        ialpha_s = (float)(1.0 / alpha);
    }
    __syncthreads();
    const int64_t beg = (int64_t)chunk * CHUNK, end = min(beg + CHUNK, n);
    const int64_t vstride = (int64_t)nrhs * n;
    float2* wr = w + (int64_t)r * n;
    float2* vj = V + (int64_t)j * vstride + (int64_t)r * n;
    const float2* vr = V + (int64_t)r * n;
    const float2 hjj = hjj_s;
    const float ia = ialpha_s;
    double nrm = 0.0;
    // UPD_U nodes per thread per pass: their loads are issued before the first
    // is used (the stores of V_j and w may alias the basis loads, so a plain
    // loop would load, wait and store node by node); same arithmetic and
    // accumulation order as one node at a time
    constexpr int UPD_U = 4;
    for (int64_t p0 = beg + threadIdx.x; p0 < end; p0 += 256 * UPD_U) {
        int64_t pp[UPD_U];
        float2 jvv[UPD_U], wvv[UPD_U], sc2[UPD_U], sd2[UPD_U];
#pragma unroll
        for (int q = 0; q < UPD_U; ++q) {
            pp[q] = p0 + 256 * q < end ? p0 + 256 * q : beg;
            jvv[q] = vj[pp[q]];
            wvv[q] = wr[pp[q]];
            sc2[q] = make_float2(0.f, 0.f);
            sd2[q] = make_float2(0.f, 0.f);
        }
        for (int k = 0; k < j; ++k) {
            float2 v[UPD_U];
#pragma unroll
            for (int q = 0; q < UPD_U; ++q) v[q] = vr[(int64_t)k * vstride + pp[q]];
#pragma unroll
            for (int q = 0; q < UPD_U; ++q) {
                cfma(sc2[q], c_s[k], v[q]);
                cfma(sd2[q], d_s[k], v[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < UPD_U; ++q) {
            if (p0 + 256 * q >= end) continue;
            float2 jv = csub(jvv[q], sc2[q]);
            jv = make_float2(jv.x * ia, jv.y * ia);
            vj[pp[q]] = jv;
            float2 wv = csub(wvv[q], sd2[q]);
            wv.x -= hjj.x * jv.x - hjj.y * jv.y;
            wv.y -= hjj.x * jv.y + hjj.y * jv.x;
            wr[pp[q]] = wv;
            nrm += (double)wv.x * wv.x + (double)wv.y * wv.y;
        }
    }
    const double2 tot = block_sum(make_double2(nrm, 0.0), lds);
    if (threadIdx.x == 0) partial[(int64_t)r * nch + chunk] = tot;
}
__global__ __launch_bounds__(256) void k_dcgs_reduce_update(int64_t n, int nrhs, int nch, int nchd, int m, int kmax,
                                                            int j, double tol, const double2* __restrict__ pc,
                                                            const double2* __restrict__ pd,
                                                            const double* __restrict__ bnorm, double2* __restrict__ H,
                                                            double* __restrict__ cs, double2* __restrict__ sn,
                                                            double2* __restrict__ g, double* __restrict__ est,
                                                            int* __restrict__ cyc, int* __restrict__ used,
                                                            float2* __restrict__ V, float2* __restrict__ w,
                                                            double2* __restrict__ partial, int v0_raw) {
    k_dcgs_reduce_update_body(hw_block(), n, nrhs, nch, nchd, m, kmax, j, tol, pc, pd, bnorm, H, cs, sn, g, est, cyc, used, V, w, partial, v0_raw);
}

// k_dcgs_norm + k_scale32 (last = 0) or k_dcgs_norm + k_dcgs_finish (last = 1,
// one workgroup per RHS): every workgroup sums the |w|^2 partials of its RHS
// in k_dcgs_norm's order; the chunk-0 workgroup writes H[j+1][j] and 1/||w||
// and finalises the column on a happy breakdown or at the end of the cycle;
// otherwise each workgroup writes its chunk of V_{j+1} = w / ||w||.
// Lazily scaled basis (pscale != nullptr, one workgroup per RHS): w already
// sits in V_{j+1}'s slot and stays unnormalised.  The preconditioner applies
// pscale = 1/||w|| as it loads V_{j+1}, so Z_{j+1} and the next w come out at
// the normalised scale; the next step's reorthogonalisation divides V_{j+1}
// by alpha' = ||w|| alpha and its coefficients c_k carry the factor ||w||, so
// H[j+1][j] = 1 here reproduces H[j+1][j] = ||w|| alpha and
// H[k][j] += ||w|| c_k exactly as with a normalised V_{j+1}.
__device__ __forceinline__ void k_dcgs_norm_scale_body(const VBlock& B, int64_t n, int nch, int m, int j, int last,
                                                         const double2* __restrict__ partial,
                                                         const double* __restrict__ bnorm, double2* __restrict__ H,
                                                         double* __restrict__ cs, double2* __restrict__ sn,
                                                         double2* __restrict__ g, double* __restrict__ est,
                                                         double* __restrict__ inv, int* __restrict__ cyc,
                                                         int* __restrict__ used, const float2* __restrict__ w,
                                                         float2* __restrict__ vnext, double* __restrict__ pscale) {
    __shared__ double lds[8];
    __shared__ double hn_s;
    const int chunk = B.x, r = B.y;
    if (!cyc[r]) return;
    const double s = reduce_chunks(partial + (int64_t)r * nch, nch, lds);
    if (threadIdx.x == 0) hn_s = sqrt(s);
    __syncthreads();
    const double hn = hn_s;
    const bool fin = last || hn < 1e-12 * bnorm[r];
    if (chunk == 0 && threadIdx.x == 0) {
        double2* hc = H + ((int64_t)r * m + j) * (m + 1);
        hc[j + 1] = make_double2(pscale && !fin ? 1.0 : hn, 0.0);
        inv[r] = hn > 0 ? 1.0 / hn : 0.0;
        if (pscale) pscale[r] = inv[r];
        if (fin) {
            const double e = givens_column(hc, cs + (int64_t)r * m, sn + (int64_t)r * m, g + (int64_t)r * (m + 1), j);
            est[r] = e / bnorm[r];
            used[r] = j + 1;
            cyc[r] = 0;
        }
    }
    if (fin || pscale) return;
    const float sc = (float)(hn > 0 ? 1.0 / hn : 0.0);
    const int64_t beg = (int64_t)chunk * CHUNK, end = min(beg + CHUNK, n);
    const float2* wr = w + (int64_t)r * n;
    float2* vr = vnext + (int64_t)r * n;
    for (int64_t p = beg + threadIdx.x; p < end; p += 256) {
        const float2 a = wr[p];
        vr[p] = make_float2(a.x * sc, a.y * sc);
    }
}
__global__ __launch_bounds__(256) void k_dcgs_norm_scale(int64_t n, int nch, int m, int j, int last,
                                                         const double2* __restrict__ partial,
                                                         const double* __restrict__ bnorm, double2* __restrict__ H,
                                                         double* __restrict__ cs, double2* __restrict__ sn,
                                                         double2* __restrict__ g, double* __restrict__ est,
                                                         double* __restrict__ inv, int* __restrict__ cyc,
                                                         int* __restrict__ used, const float2* __restrict__ w,
                                                         float2* __restrict__ vnext, double* __restrict__ pscale) {
    k_dcgs_norm_scale_body(hw_block(), n, nch, m, j, last, partial, bnorm, H, cs, sn, g, est, inv, cyc, used, w, vnext, pscale);
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
__device__ __forceinline__ void k_cycle_init_body(const VBlock& B, int nch, int m, double tol, const double2* __restrict__ partial,
                             const double* __restrict__ bnorm, double* __restrict__ beta, double* __restrict__ rel,
                             double* __restrict__ inv, int* __restrict__ active, int* __restrict__ cyc,
                             int* __restrict__ used, int* __restrict__ iters, double2* __restrict__ g,
                             double* __restrict__ pscale, double* __restrict__ hrel, int* __restrict__ hact,
                             int* __restrict__ hit, int pscale_v0) {
    __shared__ double lds[8];
    const int r = B.x;
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
        if (pscale) pscale[r] = pscale_v0 ? inv[r] : 1.0;  // V_0 stored unnormalised / normalised
        if (hrel) {  // the host's copy (pinned, mapped): read after the stream synchronises
            hrel[r] = rr;
            hact[r] = active[r];
            hit[r] = iters[r];
        }
        g[(int64_t)r * (m + 1)] = make_double2(b, 0.0);
        for (int k = 1; k <= m; ++k) g[(int64_t)r * (m + 1) + k] = make_double2(0.0, 0.0);
    }
}
__global__  void k_cycle_init(int nch, int m, double tol, const double2* __restrict__ partial,
                             const double* __restrict__ bnorm, double* __restrict__ beta, double* __restrict__ rel,
                             double* __restrict__ inv, int* __restrict__ active, int* __restrict__ cyc,
                             int* __restrict__ used, int* __restrict__ iters, double2* __restrict__ g,
                             double* __restrict__ pscale, double* __restrict__ hrel, int* __restrict__ hact,
                             int* __restrict__ hit, int pscale_v0) {
    k_cycle_init_body(hw_block(), nch, m, tol, partial, bnorm, beta, rel, inv, active, cyc, used, iters, g, pscale, hrel, hact, hit, pscale_v0);
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

// Restart residual of the flux-form operator in one pass (lazily scaled
// basis): r = b - A x in FP64 (x c128), written straight into V_0 as c64 and
// left unnormalised (the multigrid applies 1/beta as it loads V_0, and the
// first step's reorthogonalisation divides V_0 by its norm), with |r|^2 per
// workgroup and RHS for k_cycle_init (partial[r * nchr + workgroup]).  This
// replaces the FP64 apply, the norm pass over r and the c128 -> c64
// conversion of r/beta.  One workgroup = 64 columns x 4 rows x rb RHS.
__global__ __launch_bounds__(256) void k_residual_v0(int nz, int nx, int64_t n, int nrhs, int rb,
                                                     const double2* __restrict__ sx, const double2* __restrict__ sz,
                                                     const double2* __restrict__ isxe,
                                                     const double2* __restrict__ isze,
                                                     const double* __restrict__ eta, const double2* __restrict__ kap,
                                                     double k0sq, double hh, int conj, const double2* __restrict__ x,
                                                     const double2* __restrict__ b, float2* __restrict__ v0,
                                                     double2* __restrict__ partial) {
    __shared__ double lds[8];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * 64 + (tid & 63), i = blockIdx.y * 4 + (tid >> 6);
    const bool inb = i < nz && j < nx;
    Coef<double2> c;
    if (inb) node_coef<double, double2>(i, j, nz, nx, sx, sz, isxe, isze, eta, kap, k0sq, hh, 0.0, conj != 0, c);
    const int64_t idx = (int64_t)i * nx + j;
    const bool hw = j > 0, he = j < nx - 1, hn = i > 0, hs = i < nz - 1;
    const int nchr = gridDim.x * gridDim.y, wg = blockIdx.y * gridDim.x + blockIdx.x;
    const int r1 = min(nrhs, (int)(blockIdx.z + 1) * rb);
    for (int r = blockIdx.z * rb; r < r1; ++r) {
        double nrm = 0.0;
        if (inb) {
            const double2* xr = x + (int64_t)r * n;
            double2 acc = cmul(c.dg, xr[idx]);
            if (hw) cfma(acc, c.fw, xr[idx - 1]);
            if (he) cfma(acc, c.fe, xr[idx + 1]);
            if (hn) cfma(acc, c.fn, xr[idx - nx]);
            if (hs) cfma(acc, c.fs, xr[idx + nx]);
            const double2 res = csub(b[(int64_t)r * n + idx], acc);
            v0[(int64_t)r * n + idx] = make_float2((float)res.x, (float)res.y);
            nrm = res.x * res.x + res.y * res.y;
        }
        const double2 tot = block_sum(make_double2(nrm, 0.0), lds);  // every thread of the block takes part
        if (tid == 0) partial[(int64_t)r * nchr + wg] = tot;
    }
}

// Block form of k_residual_v0 (the block form of the FP32 stencil in FP64):
// one wavefront per (64-column strip, run of TZ rows, RB right-hand sides)
// loads every value of its run before forming the first node -- TZ + 2 rows
// of x per RHS, eta, kappa, b, and the strip's west/east halo columns packed
// one value per lane and handed to the edge lanes with readlane -- so a
// wavefront waits for one memory latency; |r|^2 is reduced per wavefront by
// shuffles and per workgroup (4 wavefronts) in LDS.  partial[r * nchr + wg]
// with nchr = gridDim.x * gridDim.y as in k_residual_v0.
__device__ __forceinline__ double readlane_d(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double2 shfl_up1d(double2 v) {
    return make_double2(__shfl_up(v.x, 1, 64), __shfl_up(v.y, 1, 64));
}
__device__ __forceinline__ double2 shfl_down1d(double2 v) {
    return make_double2(__shfl_down(v.x, 1, 64), __shfl_down(v.y, 1, 64));
}

template <int RB, int TZ>
__device__ __forceinline__ void k_residual_block_body(const VBlock& B, int nz, int nx, int64_t n, int nrhs,
                                                        const double2* __restrict__ sx,
                                                        const double2* __restrict__ sz,
                                                        const double2* __restrict__ isxe,
                                                        const double2* __restrict__ isze,
                                                        const double* __restrict__ eta,
                                                        const double2* __restrict__ kap, double k0sq, double hh,
                                                        int conj, const double2* __restrict__ x,
                                                        const double2* __restrict__ b, float2* __restrict__ v0,
                                                        double2* __restrict__ partial) {
    static_assert(RB * TZ <= 64, "halo columns are packed one value per lane");
    __shared__ double red[4][RB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x0 = B.x * 64, j = x0 + lane;
    const int i0 = __builtin_amdgcn_readfirstlane((B.y * 4 + wave) * TZ);
    const int r0 = B.z * RB;
    const int nchr = B.gx * B.gy, wg = B.y * B.gx + B.x;
    double nrm[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) nrm[q] = 0.0;
    if (i0 < nz) {  // uniform over the wavefront; every wavefront still takes part in the reduction below
        const bool colok = j < nx;
        const int jj = colok ? j : nx - 1;
        const bool hw = jj > 0, he = jj < nx - 1;
        const bool wstrip = x0 > 0, estrip = x0 + 64 < nx;
        double2 SZ[TZ], IZ[TZ + 1];
#pragma unroll
        for (int t = 0; t < TZ; ++t) SZ[t] = sz[min(i0 + t, nz - 1)];
#pragma unroll
        for (int t = 0; t <= TZ; ++t) IZ[t] = isze[min(i0 + t, nz)];
        double2 U[RB][TZ + 2], Bv[RB][TZ];
        double E[TZ + 2];
        double2 K[TZ];
#pragma unroll
        for (int t = 0; t < TZ + 2; ++t) {
            const int i = i0 - 1 + t;
            const bool rok = i >= 0 && i < nz;
#pragma unroll
            for (int q = 0; q < RB; ++q)
                U[q][t] = (rok && r0 + q < nrhs) ? x[(int64_t)(r0 + q) * n + (int64_t)i * nx + jj]
                                                 : make_double2(0.0, 0.0);
            E[t] = rok ? eta[(int64_t)i * nx + jj] : 1.0;
        }
#pragma unroll
        for (int t = 0; t < TZ; ++t) {
            const bool rok = i0 + t < nz;
            K[t] = rok ? kap[(int64_t)(i0 + t) * nx + jj] : make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < RB; ++q)
                Bv[q][t] = (rok && r0 + q < nrhs) ? b[(int64_t)(r0 + q) * n + (int64_t)(i0 + t) * nx + jj]
                                                  : make_double2(0.0, 0.0);
        }
        double2 HW = make_double2(0.0, 0.0), HE = make_double2(0.0, 0.0);
        double EH = 1.0;  // lanes 0..TZ-1: west eta of row t; lanes 32..32+TZ-1: east eta
        {
            const int q = lane / TZ, t = lane % TZ, i = i0 + t;
            if (lane < RB * TZ && i < nz && r0 + q < nrhs) {
                const int64_t o = (int64_t)(r0 + q) * n + (int64_t)i * nx;
                if (wstrip) HW = x[o + x0 - 1];
                if (estrip) HE = x[o + x0 + 64];
            }
            if (lane < TZ && i0 + lane < nz && wstrip) EH = eta[(int64_t)(i0 + lane) * nx + x0 - 1];
            if (lane >= 32 && lane < 32 + TZ && i0 + lane - 32 < nz && estrip)
                EH = eta[(int64_t)(i0 + lane - 32) * nx + x0 + 64];
        }
        const double2 sxj = sx[jj], iw = isxe[jj], ie = isxe[jj + 1];
        const bool cj = conj != 0;
#pragma unroll
        for (int t = 0; t < TZ; ++t) {
            const int i = i0 + t;
            const double en = E[t], ec = E[t + 1], es = E[t + 2];
            double eww = __shfl_up(ec, 1, 64), eee = __shfl_down(ec, 1, 64);
            const double ehw = readlane_d(EH, t), ehe = readlane_d(EH, 32 + t);
            if (lane == 0) eww = ehw;
            if (lane == 63) eee = ehe;
            if (!hw) eww = 1.0;
            if (!he) eee = 1.0;
            Coef<double2> c;
            c.fw = cscale(cmul(SZ[t], iw), hh * (ec + eww));
            c.fe = cscale(cmul(SZ[t], ie), hh * (ec + eee));
            c.fn = cscale(cmul(sxj, IZ[t]), hh * (ec + en));
            c.fs = cscale(cmul(sxj, IZ[t + 1]), hh * (ec + es));
            const double2 mass = cscale(cmul(cmul(SZ[t], sxj), K[t]), k0sq);
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
                const double2 uc = U[q][t + 1];
                double2 w = shfl_up1d(uc), e = shfl_down1d(uc);
                const double2 hwv = make_double2(readlane_d(HW.x, q * TZ + t), readlane_d(HW.y, q * TZ + t));
                const double2 hev = make_double2(readlane_d(HE.x, q * TZ + t), readlane_d(HE.y, q * TZ + t));
                if (lane == 0) w = hwv;
                if (lane == 63) e = hev;
                double2 acc = cmul(c.dg, uc);
                if (hw) cfma(acc, c.fw, w);
                if (he) cfma(acc, c.fe, e);
                if (hn) cfma(acc, c.fn, U[q][t]);
                if (hs) cfma(acc, c.fs, U[q][t + 2]);
                const int r = r0 + q;
                if (colok && r < nrhs && i < nz) {
                    const double2 res = csub(Bv[q][t], acc);
                    v0[(int64_t)r * n + (int64_t)i * nx + j] = make_float2((float)res.x, (float)res.y);
                    nrm[q] += res.x * res.x + res.y * res.y;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        const double v = wave_sum(nrm[q]);
        if (lane == 0) red[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < RB && r0 + (int)threadIdx.x < nrhs) {
        const int q = threadIdx.x;
        const double tot = red[0][q] + red[1][q] + red[2][q] + red[3][q];
        partial[(int64_t)(r0 + q) * nchr + wg] = make_double2(tot, 0.0);
    }
}
template <int RB, int TZ>
__global__ __launch_bounds__(256) void k_residual_block(int nz, int nx, int64_t n, int nrhs,
                                                        const double2* __restrict__ sx,
                                                        const double2* __restrict__ sz,
                                                        const double2* __restrict__ isxe,
                                                        const double2* __restrict__ isze,
                                                        const double* __restrict__ eta,
                                                        const double2* __restrict__ kap, double k0sq, double hh,
                                                        int conj, const double2* __restrict__ x,
                                                        const double2* __restrict__ b, float2* __restrict__ v0,
                                                        double2* __restrict__ partial) {
    k_residual_block_body<RB, TZ>(hw_block(), nz, nx, n, nrhs, sx, sz, isxe, isze, eta, kap, k0sq, hh, conj, x, b, v0, partial);
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

// x[r] += sum_{k<used[r]} y_k Z_k[r] with y = R^{-1} g (upper triangular,
// used[r] columns of the rotated Hessenberg; FP64 accumulate into c128 x).
// Thread 0 of every workgroup solves its RHS's small triangle into LDS (the
// same operations in the same order in every workgroup, so every workgroup
// holds the same y) and the workgroup applies it to its chunk of x.  The
// loads of x and Z are issued first and stay in flight (LDS-only barriers)
// while the triangle's entries, loaded into LDS by all threads at once, are
// solved.
constexpr int XU_MAX = 4;  // columns whose Z loads are issued ahead (more: loaded after the solve)
constexpr int XU_LDS = 16;  // columns whose triangle is staged in LDS (more: read from global memory)
// back substitution of the u x u triangle (entry (q, k) at t[q * ld + k]) for y
template <typename T>
__device__ __forceinline__ void xupdate_triangle(const T* __restrict__ t, int ld, const T* __restrict__ gv, int u,
                                                 double2* __restrict__ yy) {
    for (int k = u - 1; k >= 0; --k) {
        double2 a = gv[k];
        for (int q = k + 1; q < u; ++q) a = csub(a, cmul(t[q * ld + k], yy[q]));
        const double2 d = t[k * ld + k];
        const double dd = d.x * d.x + d.y * d.y;
        yy[k] = dd > 0 ? make_double2((a.x * d.x + a.y * d.y) / dd, (a.y * d.x - a.x * d.y) / dd)
                       : make_double2(0.0, 0.0);
    }
}
__device__ __forceinline__ void k_xupdate_body(const VBlock& B, int64_t n, int nrhs, int m,
                                               const double2* __restrict__ H, const double2* __restrict__ g,
                                               const float2* __restrict__ Z, const int* __restrict__ used,
                                               double2* __restrict__ x) {
    __shared__ double2 yy[64];
    __shared__ double2 sh[256], sg[256];  // one slot per thread (launched with 256 threads)
    static_assert(XU_LDS * XU_LDS <= 256, "the staged triangle takes one slot per thread");
    const int r = B.y;
    const int u = used[r];
    if (u == 0) return;
    const int64_t p = (int64_t)B.x * blockDim.x + threadIdx.x;
    const int64_t pc = p < n ? p : 0;
    const int64_t vstride = (int64_t)nrhs * n;
    const float2* zr = Z + (int64_t)r * n;
    const double2* Hr = H + (int64_t)r * m * (m + 1);
    const double2* gr = g + (int64_t)r * (m + 1);
    if (u > XU_LDS) {  // long cycles: the triangle from global memory
        if (threadIdx.x == 0) xupdate_triangle(Hr, m + 1, gr, u, yy);
        __syncthreads();
    }
    // The triangle's entries are loaded first (one per thread), then x and
    // the first XU_MAX directions (addresses clamped, unused columns load Z_0
    // again): waiting for the older entries leaves the data loads in flight
    // through the staging and the solve, which read LDS only.
    const bool staged = u <= XU_LDS;
    const int e = (staged && (int)threadIdx.x < u * u) ? (int)threadIdx.x : 0;
    const double2 hv = Hr[(int64_t)(e / u) * (m + 1) + e % u];
    const double2 gv = gr[(int)threadIdx.x < u ? threadIdx.x : 0];
    __builtin_amdgcn_sched_barrier(0);  // keep these loads ahead of the data loads
    double2 acc = x[(int64_t)r * n + pc];
    float2 z[XU_MAX];
#pragma unroll
    for (int k = 0; k < XU_MAX; ++k) z[k] = zr[(int64_t)(k < u ? k : 0) * vstride + pc];
    // every thread writes its slot (slots past u * u and u, and all of them
    // when the triangle is not staged, hold unused copies): a store behind a
    // branch would let the compiler sink the load to it, behind the data loads
    sh[threadIdx.x] = hv;
    sg[threadIdx.x] = gv;
    if (staged) {
        lds_barrier();
        if (threadIdx.x == 0) xupdate_triangle(sh, u, sg, u, yy);
        lds_barrier();
    }
    if (p >= n) return;
#pragma unroll
    for (int k = 0; k < XU_MAX; ++k)
        if (k < u) {
            const double2 c = yy[k];
            acc.x += c.x * z[k].x - c.y * z[k].y;
            acc.y += c.x * z[k].y + c.y * z[k].x;
        }
    for (int k = XU_MAX; k < u; ++k) {
        const float2 zk = zr[(int64_t)k * vstride + p];
        const double2 c = yy[k];
        acc.x += c.x * zk.x - c.y * zk.y;
        acc.y += c.x * zk.y + c.y * zk.x;
    }
    x[(int64_t)r * n + p] = acc;
}
__global__ __launch_bounds__(256) void k_xupdate(int64_t n, int nrhs, int m, const double2* __restrict__ H,
                                                 const double2* __restrict__ g, const float2* __restrict__ Z,
                                                 const int* __restrict__ used, double2* __restrict__ x) {
    k_xupdate_body(hw_block(), n, nrhs, m, H, g, Z, used, x);
}

__global__ void k_bi_izero_k(int count, int* p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// stop right-hand side r (stagnation): no further cycles run for it
__global__ void k_stop_rhs(int r, int* __restrict__ active, int* __restrict__ cyc) {
    active[r] = 0;
    cyc[r] = 0;
}

__global__ void k_zero_c128(int64_t count, double2* p) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = make_double2(0.0, 0.0);
}

// ---- segmented forms: one launch over the plans of a batched solve ---------
// Each workgroup finds its plan (segment) and local block indices, then runs
// the single-plan kernel body on that plan's arrays.  The batched solve takes
// the single-plan solve's lazily scaled, fused path: w = A z_j sits in
// V_{j+1}'s slot, the restart residual goes straight into V_0.
__device__ __forceinline__ int64_t seg_vstride(const SegKrylov& a) { return (int64_t)a.nrhs * a.n; }
__global__ __launch_bounds__(256, 3) void k_seg_dots(const SegDims* __restrict__ dims, int nseg,
                                                     const SegKrylov* __restrict__ seg, int j) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegKrylov& a = seg[s];
    k_dcgs_dots_body<8>(B, a.n, a.nrhs, a.nchd, a.kmax, j, a.V, a.V + (j + 1) * seg_vstride(a), a.cyc, a.partial,
                        a.partial + (int64_t)a.nrhs * a.kmax * a.nchd);
}
__global__ __launch_bounds__(256) void k_seg_reduce_update(const SegDims* __restrict__ dims, int nseg,
                                                           const SegKrylov* __restrict__ seg, int j, double tol) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegKrylov& a = seg[s];
    const int64_t pd = (int64_t)a.nrhs * a.kmax * a.nchd;
    k_dcgs_reduce_update_body(B, a.n, a.nrhs, a.nch, a.nchd, a.mm, a.kmax, j, tol, a.partial, a.partial + pd,
                              a.bnorm, a.H, a.cs, a.sn, a.g, a.est, a.cyc, a.used, a.V,
                              a.V + (j + 1) * seg_vstride(a), a.partial + 2 * pd, 1);
}
__global__ __launch_bounds__(256) void k_seg_norm_scale(const SegDims* __restrict__ dims, int nseg,
                                                        const SegKrylov* __restrict__ seg, int j, int last) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegKrylov& a = seg[s];
    float2* wj = a.V + (j + 1) * seg_vstride(a);
    k_dcgs_norm_scale_body(B, a.n, a.nch, a.mm, j, last, a.partial + (int64_t)2 * a.nrhs * a.kmax * a.nchd,
                           a.bnorm, a.H, a.cs, a.sn, a.g, a.est, a.inv, a.cyc, a.used, wj, wj, a.pscale);
}
__global__ __launch_bounds__(256) void k_seg_xupdate(const SegDims* __restrict__ dims, int nseg,
                                                     const SegKrylov* __restrict__ seg) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegKrylov& a = seg[s];
    k_xupdate_body(B, a.n, a.nrhs, a.mm, a.H, a.g, a.Z, a.used, a.x);
}
__global__ __launch_bounds__(256) void k_seg_residual(const SegDims* __restrict__ dims, int nseg,
                                                      const SegKrylov* __restrict__ seg, int conj) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegKrylov& a = seg[s];
    k_residual_block_body<RES_RB, RES_TZ>(B, a.nz, a.nx, a.n, a.nrhs, a.sx64, a.sz64, a.isxe64, a.isze64, a.eta64,
                                          a.kap64, a.k0sq, a.hh, conj, a.x, a.b, a.V, a.partial);
}
__global__ void k_seg_cycle_init(const SegDims* __restrict__ dims, int nseg, const SegKrylov* __restrict__ seg,
                                 int slot, double tol) {
    int s;
    VBlock B;
    if (!seg_block(dims, nseg, s, B)) return;
    const SegKrylov& a = seg[s];
    char* d = a.dstat + (int64_t)slot * a.stat_bytes;
    k_cycle_init_body(B, a.nchb, a.mm, tol, a.partial, a.bnorm, a.beta, a.rel, a.inv, a.active, a.cyc, a.used,
                      a.iters, a.g, a.pscale, (double*)d, (int*)(d + (size_t)a.ws_nrhs * sizeof(double)),
                      (int*)(d + (size_t)a.ws_nrhs * (sizeof(double) + sizeof(int))), 1);
}

void seg_dots(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_dots, dim3(blocks), dim3(256), 0, s, dims, nseg, k, j);
}
void seg_reduce_update(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, double tol,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_seg_reduce_update, dim3(blocks), dim3(256), 0, s, dims, nseg, k, j, tol);
}
void seg_norm_scale(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, int last, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_norm_scale, dim3(blocks), dim3(256), 0, s, dims, nseg, k, j, last);
}
void seg_xupdate(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_xupdate, dim3(blocks), dim3(256), 0, s, dims, nseg, k);
}
void seg_residual(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int conj, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_residual, dim3(blocks), dim3(256), 0, s, dims, nseg, k, conj);
}
void seg_cycle_init(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int slot, double tol,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_seg_cycle_init, dim3(blocks), dim3(64), 0, s, dims, nseg, k, slot, tol);
}

KrylovGrids krylov_grids(int nz, int nx, int nrhs) {
    KrylovGrids g{};
    const int64_t n = (int64_t)nz * nx;
    g.nch = nchunks(n);
    g.nchd = (int)((n + 2047) / 2048);  // k_dcgs_dots<8>: 2048-node chunks
    const int bx = (nx + 63) / 64, by = (nz + 4 * RES_TZ - 1) / (4 * RES_TZ);
    g.nchb = bx * by;
    auto set = [](int* d, int x, int y, int z) {
        d[0] = x;
        d[1] = y;
        d[2] = z;
    };
    set(g.dots, g.nchd, nrhs, 1);
    set(g.update, g.nch, nrhs, 1);
    set(g.norm, 1, nrhs, 1);
    set(g.xupdate, (int)((n + 255) / 256), nrhs, 1);
    set(g.residual, bx, by, (nrhs + RES_RB - 1) / RES_RB);
    set(g.init, nrhs, 1, 1);
    set(g.apply, (nx + 63) / 64, ((nz + 3) / 4 + 3) / 4, (nrhs + 3) / 4);  // k_apply_block<4, 4>
    return g;
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
    p->generation++;  // captured cycles refer to the old buffers
    const int nch = nchunks(n);
    const int kmax = restart + 1;
    w.nrhs = nrhs;
    w.restart = restart;
    w.n = n;
    HELM_TRY(alloc(&w.V, (size_t)(restart + 1) * nrhs * n));
    HELM_TRY(alloc(&w.Z, (size_t)restart * nrhs * n));
    HELM_TRY(alloc(&w.w, (size_t)nrhs * n));
    HELM_TRY(alloc(&w.r64, (size_t)nrhs * n));
    // <V_k,V_j> | <V_k,w> (per dot chunk) | |w|^2 (per chunk)
    w.partial_len = (int64_t)nrhs * (2 * kmax * ndotchunks_max(n) + nch);
    HELM_TRY(alloc(&w.partial, (size_t)w.partial_len));
    HELM_TRY(alloc(&w.H, (size_t)nrhs * restart * (restart + 1)));
    HELM_TRY(alloc(&w.cs, (size_t)nrhs * restart));
    HELM_TRY(alloc(&w.sn, (size_t)nrhs * restart));
    HELM_TRY(alloc(&w.g, (size_t)nrhs * (restart + 1)));
    HELM_TRY(alloc(&w.coef, (size_t)nrhs * (2 * kmax + 1)));
    HELM_TRY(alloc(&w.beta, nrhs));
    HELM_TRY(alloc(&w.bnorm, nrhs));
    HELM_TRY(alloc(&w.rel, nrhs));
    HELM_TRY(alloc(&w.inv, nrhs));
    HELM_TRY(alloc(&w.pscale, nrhs));
    HELM_TRY(alloc(&w.est, nrhs));
    HELM_TRY(alloc(&w.active, nrhs));
    HELM_TRY(alloc(&w.cyc, nrhs));
    HELM_TRY(alloc(&w.used, nrhs));
    HELM_TRY(alloc(&w.iters, nrhs));
    w.h_rel.assign(nrhs, 0.0);
    w.h_active.assign(nrhs, 0);
    w.h_iters.assign(nrhs, 0);
    if (w.hstat) (void)hipHostFree(w.hstat);
    w.hstat = nullptr;
    w.stat_bytes = (size_t)nrhs * (sizeof(double) + 2 * sizeof(int));
    HELM_HIP(hipHostMalloc(&w.hstat, 2 * w.stat_bytes, hipHostMallocMapped | hipHostMallocCoherent));
    for (hipEvent_t& e : w.stat_ev)
        if (!e) HELM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    {
        char* base = nullptr;
        HELM_HIP(hipHostGetDevicePointer((void**)&base, w.hstat, 0));
        w.dstat = base;
    }
    return HELM_OK;
}

SegKrylov krylov_seg(helm_plan* p, const double2* b, double2* x, int nrhs) {
    const Workspace& W = p->ws;
    const Level& L = p->levels[0];
    const KrylovGrids g = krylov_grids(L.nz, L.nx, nrhs);
    SegKrylov a{};
    a.n = L.n;
    a.nrhs = nrhs;
    a.nch = g.nch;
    a.nchd = g.nchd;
    a.nchb = g.nchb;
    a.nz = L.nz;
    a.nx = L.nx;
    a.ws_nrhs = W.nrhs;
    a.kmax = W.restart + 1;
    a.mm = W.restart;
    a.V = W.V;
    a.Z = W.Z;
    a.partial = W.partial;
    a.H = W.H;
    a.sn = W.sn;
    a.g = W.g;
    a.coef = W.coef;
    a.cs = W.cs;
    a.bnorm = W.bnorm;
    a.beta = W.beta;
    a.rel = W.rel;
    a.inv = W.inv;
    a.pscale = W.pscale;
    a.est = W.est;
    a.active = W.active;
    a.cyc = W.cyc;
    a.used = W.used;
    a.iters = W.iters;
    a.sx64 = L.sx64;
    a.sz64 = L.sz64;
    a.isxe64 = L.isxe64;
    a.isze64 = L.isze64;
    a.kap64 = L.kap64;
    a.eta64 = L.eta64;
    a.sx32 = L.sx32;
    a.sz32 = L.sz32;
    a.isxe32 = L.isxe32;
    a.isze32 = L.isze32;
    a.kap32 = L.kap32;
    a.eta32 = L.eta32;
    a.x = x;
    a.b = b;
    a.dstat = W.dstat;
    a.stat_bytes = (int64_t)W.stat_bytes;
    a.k0sq = p->k0 * p->k0;
    a.hh = 0.5 / (L.h_m * L.h_m);
    a.k0sq32 = (float)a.k0sq;
    a.hh32 = (float)a.hh;
    return a;
}

// start of a solve, as in solve_gmres: ||b|| per RHS, active flags, used = 0, x = 0 (unless x holds the start)
int32_t krylov_begin(helm_plan* p, const double2* b, double2* x, int nrhs, bool use_x0, hipStream_t s) {
    Workspace& W = p->ws;
    const int64_t n = p->levels[0].n;
    const int nch = nchunks(n);
    hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nch, nrhs), dim3(256), 0, s, n, nch, b, W.partial);
    hipLaunchKernelGGL(k_bnorm, dim3(nrhs), dim3(64), 0, s, nch, W.partial, W.bnorm, W.active, W.iters);
    HELM_HIP(hipMemsetAsync(W.used, 0, nrhs * sizeof(int), s));
    if (!use_x0) {
        const int64_t cnt = (int64_t)nrhs * n;
        hipLaunchKernelGGL(k_zero_c128, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, cnt, x);
    }
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

// ---- warm start (minimal-residual steps, wave_operator.py:129-135) --------
// inv[r] = 1 / ||r64_r|| (0 for a zero residual), cyc[r] = 1
__global__ void k_ws_scale(int nch, const double2* __restrict__ partial, double* __restrict__ inv,
                           int* __restrict__ cyc) {
    __shared__ double lds[8];
    const int r = blockIdx.x;
    double s = reduce_chunks(partial + (int64_t)r * nch, nch, lds);
    if (threadIdx.x == 0) {
        inv[r] = s > 0 ? 1.0 / sqrt(s) : 0.0;
        cyc[r] = 1;
    }
}

// part[(2r)nch + c] = sum conj(w) res, part[(2r+1)nch + c] = sum |w|^2 over chunk c
__global__ __launch_bounds__(256) void k_ws_partial(int64_t n, int nch, const float2* __restrict__ w,
                                                    const double2* __restrict__ res, double2* __restrict__ part) {
    __shared__ double lds[8];
    const int chunk = blockIdx.x, r = blockIdx.y;
    const int64_t beg = (int64_t)chunk * CHUNK, end = min(beg + CHUNK, n);
    double2 d = make_double2(0.0, 0.0);
    double q = 0.0;
    for (int64_t p = beg + threadIdx.x; p < end; p += 256) {
        const float2 a = w[(int64_t)r * n + p];
        const double2 b = res[(int64_t)r * n + p];
        d.x += (double)a.x * b.x + (double)a.y * b.y;
        d.y += (double)a.x * b.y - (double)a.y * b.x;
        q += (double)a.x * a.x + (double)a.y * a.y;
    }
    const double2 sd = block_sum(d, lds);
    const double2 sq = block_sum(make_double2(q, 0.0), lds);
    if (threadIdx.x == 0) {
        part[(int64_t)(2 * r) * nch + chunk] = sd;
        part[(int64_t)(2 * r + 1) * nch + chunk] = sq;
    }
}

// alpha_r = <A d, r> / |A d|^2 (0 when A d vanishes)
__global__ __launch_bounds__(256) void k_ws_alpha(int nch, const double2* __restrict__ part, double2* __restrict__ alpha) {
    __shared__ double lds[8];
    const int r = blockIdx.x;
    double2 d = make_double2(0.0, 0.0), q = make_double2(0.0, 0.0);
    for (int c = threadIdx.x; c < nch; c += blockDim.x) {
        d = cadd(d, part[(int64_t)(2 * r) * nch + c]);
        q = cadd(q, part[(int64_t)(2 * r + 1) * nch + c]);
    }
    d = block_sum(d, lds);
    q = block_sum(q, lds);
    if (threadIdx.x == 0) alpha[r] = q.x > 0 ? cscale(d, 1.0 / q.x) : make_double2(0.0, 0.0);
}

// x[r] += alpha_r z[r]
__global__ void k_ws_update(int64_t n, const float2* __restrict__ z, const double2* __restrict__ alpha,
                            double2* __restrict__ x) {
    const int r = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double2 a = alpha[r];
    const float2 v = z[(int64_t)r * n + p];
    double2& o = x[(int64_t)r * n + p];
    o.x += a.x * v.x - a.y * v.y;
    o.y += a.x * v.y + a.y * v.x;
}

int32_t warm_start(helm_plan* p, int pc, const double2* b, double2* x, int nrhs, int steps, bool adj,
                   hipStream_t s) {
    HELM_TRY(ensure_workspace(p, nrhs, std::max(1, p->ws.restart)));
    Workspace& W = p->ws;
    const Level& L = p->levels[0];
    const int64_t n = L.n;
    const int nch = nchunks(n);
    const double k0sq = p->k0 * p->k0;
    const dim3 blk(256), gvec((unsigned)((n + 255) / 256), nrhs);
    float2* r32 = W.V;
    float2* z = W.Z;
    for (int it = 0; it < steps; ++it) {
        HELM_TRY(launch_apply64(L, k0sq, x, W.r64, nrhs, adj, b, s));
        hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nch, nrhs), blk, 0, s, n, nch, W.r64, W.partial);
        hipLaunchKernelGGL(k_ws_scale, dim3(nrhs), dim3(64), 0, s, nch, W.partial, W.inv, W.cyc);
        hipLaunchKernelGGL(k_scale_to32, gvec, blk, 0, s, n, W.r64, r32, W.inv, W.cyc);
        if (pc != HELM_PC_NONE)
            HELM_TRY(precond_apply(p, pc, r32, z, nrhs, adj, s));
        else
            HELM_HIP(hipMemcpyAsync(z, r32, (size_t)nrhs * n * sizeof(float2), hipMemcpyDeviceToDevice, s));
        HELM_TRY(launch_apply32(L, k0sq, z, W.w, nrhs, adj, nullptr, 0.0f, s));
        hipLaunchKernelGGL(k_ws_partial, dim3(nch, nrhs), blk, 0, s, n, nch, W.w, W.r64, W.partial);
        hipLaunchKernelGGL(k_ws_alpha, dim3(nrhs), blk, 0, s, nch, W.partial, W.coef);
        hipLaunchKernelGGL(k_ws_update, gvec, blk, 0, s, n, z, W.coef, x);
        HELM_HIP(hipGetLastError());
    }
    return HELM_OK;
}

// Capture (when the key changed) and launch the restart cycle as a HIP graph.
// Capture runs on the workspace's own stream (the null stream cannot be
// captured); the graph is launched on the caller's stream.  A capture that
// fails (the runtime invalidated it, or a launch in it failed) has executed
// nothing: the runtime's last error is cleared and the cycle is enqueued
// directly on the caller's stream instead, which reports any genuine error.
template <typename F>
static int32_t cycle_graph(Workspace& W, const uint64_t key[8], F&& enqueue, int m, hipStream_t s) {
    if (!W.gexec || std::memcmp(W.gkey, key, sizeof(W.gkey)) != 0) {
        if (W.gexec) {
            (void)hipGraphExecDestroy(W.gexec);
            W.gexec = nullptr;
        }
        if (!W.gstream) HELM_HIP(hipStreamCreateWithFlags(&W.gstream, hipStreamNonBlocking));
        bool captured = hipStreamBeginCapture(W.gstream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (captured) {
            const int32_t st = enqueue(m, W.gstream);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(W.gstream, &g);
            captured = st == HELM_OK && e == hipSuccess && g &&
                       hipGraphInstantiate(&W.gexec, g, nullptr, nullptr, 0) == hipSuccess;
            if (g) (void)hipGraphDestroy(g);
        }
        if (!captured) {
            if (W.gexec) (void)hipGraphExecDestroy(W.gexec);
            W.gexec = nullptr;
            (void)hipGetLastError();
            W.graph_fallbacks++;
            return enqueue(m, s);
        }
        std::memcpy(W.gkey, key, sizeof(W.gkey));
    }
    HELM_HIP(hipGraphLaunch(W.gexec, s));
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
    int pc = o->precond;  // the learned preconditioner can fall back to the multigrid mid-solve
    const dim3 blk(256);
    const dim3 gvec((unsigned)((n + 255) / 256), nrhs);
    const int64_t vstride = (int64_t)nrhs * n;
    const int nchd = ndotchunks(n);
    double2* pd_part = W.partial + (int64_t)nrhs * kmax * nchd;      // <V_k, w>
    double2* pw2 = W.partial + (int64_t)2 * nrhs * kmax * nchd;      // |w|^2 after the update

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
    W.graph_fallbacks = 0;
    Profiler* pf = (p->prof && p->prof->on) ? p->prof : nullptr;
    const double nn = (double)n, rr = (double)nrhs;
    const double coef_bytes32 = L.op != HELM_OP_FLUX5 ? 16.0 : 12.0;  // per node, FP32 coefficients
    const double apply_bytes = 16.0 * nn * rr + coef_bytes32 * nn;
    const double resid_bytes = 48.0 * nn * rr + 2.0 * coef_bytes32 * nn;
    // lazily scaled basis: the multigrid's finest-level kernels normalise V_j as they load it
    // (conditioned plans: the prescale kernel applies it, or on adjoint solves the multigrid)
    // (the spectral preconditioner: its first FFT pass)
    const bool spc = pc == HELM_PC_SPECTRAL || pc == HELM_PC_SPECTRAL_PLAIN;
    const bool lazy = dcgs_fused() && lazy_basis() &&
                      ((pc == HELM_PC_MG && ((L.op == HELM_OP_COND8 && !adj) || mg_scaled_input_ok(p))) ||
                       (spc && spectral_scaled_input_ok(p)));
    // ... and for the flux operator the restart residual is written straight into V_0 (k_residual_v0)
    const dim3 rgrid((unsigned)((L.nx + 63) / 64), (unsigned)((L.nz + 3) / 4), 1);
    const int nchr = (int)(rgrid.x * rgrid.y);
    const int rgroups = (int)std::max<int64_t>(1, std::min<int64_t>(nrhs, (2048 + nchr - 1) / nchr));
    const int rrb = (nrhs + rgroups - 1) / rgroups;
    const bool v0raw = lazy && L.op == HELM_OP_FLUX5 && (int64_t)nrhs * nchr <= W.partial_len;
    // its block form: 64 columns x RES_TZ rows x RES_RB RHS per wavefront, 4 wavefronts per workgroup
    const dim3 bgrid((unsigned)((L.nx + 63) / 64), (unsigned)((L.nz + 4 * RES_TZ - 1) / (4 * RES_TZ)), 1);
    const int nchb = (int)(bgrid.x * bgrid.y);
    const int nchr_used = resid_block() ? nchb : nchr;
    // restart status into pinned slot `slot`: the FP64 true residual, its norm
    // and k_cycle_init, then the slot's event
    auto enqueue_status = [&](int slot) -> int32_t {
        hipEvent_t ev = pf ? prof_begin(pf, PK_RESIDUAL, s) : nullptr;
        if (v0raw && resid_block()) {
            hipLaunchKernelGGL((k_residual_block<RES_RB, RES_TZ>), dim3(bgrid.x, bgrid.y, (nrhs + RES_RB - 1) / RES_RB),
                               dim3(256), 0, s, L.nz, L.nx, n, nrhs, L.sx64, L.sz64, L.isxe64, L.isze64, L.eta64,
                               L.kap64, k0sq, 0.5 / (L.h_m * L.h_m), (int)adj, (const double2*)x, b, W.V, W.partial);
        } else if (v0raw) {
            hipLaunchKernelGGL(k_residual_v0, dim3(rgrid.x, rgrid.y, (nrhs + rrb - 1) / rrb), dim3(256), 0, s, L.nz,
                               L.nx, n, nrhs, rrb, L.sx64, L.sz64, L.isxe64, L.isze64, L.eta64, L.kap64, k0sq,
                               0.5 / (L.h_m * L.h_m), (int)adj, (const double2*)x, b, W.V, W.partial);
        } else {
            HELM_TRY(launch_apply64(L, k0sq, x, W.r64, nrhs, adj, b, s));
        }
        if (pf) prof_end(pf, PK_RESIDUAL, ev, resid_bytes, s);
        if (!v0raw)
            hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nch, nrhs), blk, 0, s, n, nch, W.r64, W.partial);
        char* d = W.dstat + (size_t)slot * W.stat_bytes;
        hipLaunchKernelGGL(k_cycle_init, dim3(nrhs), dim3(64), 0, s, v0raw ? nchr_used : nch, mm, o->tolerance, W.partial,
                           W.bnorm, W.beta, W.rel, W.inv, W.active, W.cyc, W.used, W.iters, W.g, W.pscale, (double*)d,
                           (int*)(d + (size_t)W.nrhs * sizeof(double)),
                           (int*)(d + (size_t)W.nrhs * (sizeof(double) + sizeof(int))), (int)v0raw);
        HELM_HIP(hipGetLastError());
        HELM_HIP(hipEventRecord(W.stat_ev[slot], s));
        return HELM_OK;
    };
    double vec_bytes = 8.0 * nn * nrhs;  // one c64 vector over the RHS active at the last status read
    auto enqueue_cycle = [&](int nsteps, hipStream_t cs) -> int32_t {
        if (!v0raw) hipLaunchKernelGGL(k_scale_to32, gvec, blk, 0, cs, n, W.r64, W.V, W.inv, W.cyc);
        for (int j = 0; j < nsteps; ++j) {
            float2* vj = W.V + (int64_t)j * vstride;
            // the direction must be the V_j the operator sees (DCGS2 rewrites V_j afterwards)
            float2* zj = W.Z + (int64_t)j * vstride;
            // w = A z_j: in V_{j+1}'s slot when the basis is scaled lazily
            float2* wj = lazy ? W.V + (int64_t)(j + 1) * vstride : W.w;
            hipEvent_t ev;
            if (pc != HELM_PC_NONE) {
                ev = pf ? prof_begin(pf, PK_PRECOND, cs) : nullptr;
                if (lazy && spc)
                    HELM_TRY(spectral_precond(p, vj, zj, nrhs, adj, cs, pc == HELM_PC_SPECTRAL, W.pscale));
                else if (lazy && L.op == HELM_OP_COND8)
                    HELM_TRY(cond_precond(p, vj, zj, nrhs, adj, cs, W.pscale));
                else if (lazy)
                    HELM_TRY(mg_vcycle(p, vj, zj, nrhs, adj, cs, W.pscale, W.cyc));
                else
                    HELM_TRY(precond_apply(p, pc, vj, zj, nrhs, adj, cs));
                if (pf) prof_end(pf, PK_PRECOND, ev, 0.0, cs);
            } else {
                HELM_HIP(hipMemcpyAsync(zj, vj, (size_t)vstride * sizeof(float2), hipMemcpyDeviceToDevice, cs));
            }
            ev = pf ? prof_begin(pf, PK_APPLY, cs) : nullptr;
            HELM_TRY(launch_apply32(L, k0sq, zj, wj, nrhs, adj, nullptr, 0.0f, cs));
            if (pf) prof_end(pf, PK_APPLY, ev, apply_bytes, cs);
            hipEvent_t ea, eb;
            if (prof_pair(pf, PK_DOTS, &ea, &eb)) {
                hipExtLaunchKernelGGL(dot_npt() == 4 ? k_dcgs_dots<4> : k_dcgs_dots<8>, dim3(nchd, nrhs), blk, 0, cs,
                                      ea, eb, 0, n, nrhs, nchd, kmax, j,
                                      (const float2*)W.V, (const float2*)wj, (const int*)W.cyc, W.partial,
                                      pd_part);
                prof_push(pf, PK_DOTS, ea, eb, (j + 2) * vec_bytes);
            } else {
                hipLaunchKernelGGL(dot_npt() == 4 ? k_dcgs_dots<4> : k_dcgs_dots<8>, dim3(nchd, nrhs), blk, 0, cs, n,
                                   nrhs, nchd, kmax, j, W.V, wj,
                                   W.cyc, W.partial, pd_part);
            }
            if (dcgs_fused()) {
                if (prof_pair(pf, PK_UPDATE, &ea, &eb)) {
                    hipExtLaunchKernelGGL(k_dcgs_reduce_update, dim3(nch, nrhs), blk, 0, cs, ea, eb, 0, n, nrhs,
                                          nch, nchd, mm, kmax, j, o->tolerance, (const double2*)W.partial,
                                          (const double2*)pd_part, (const double*)W.bnorm, W.H, W.cs, W.sn, W.g,
                                          W.est, W.cyc, W.used, W.V, wj, pw2, (int)v0raw);
                    prof_push(pf, PK_UPDATE, ea, eb, (j + 4) * vec_bytes);
                } else {
                    hipLaunchKernelGGL(k_dcgs_reduce_update, dim3(nch, nrhs), blk, 0, cs, n, nrhs, nch, nchd, mm,
                                       kmax, j, o->tolerance, W.partial, pd_part, W.bnorm, W.H, W.cs, W.sn, W.g,
                                       W.est, W.cyc, W.used, W.V, wj, pw2, (int)v0raw);
                }
                const bool last = j + 1 >= nsteps;
                hipLaunchKernelGGL(k_dcgs_norm_scale, dim3(last || lazy ? 1 : nch, nrhs), blk, 0, cs, n, nch, mm,
                                   j, (int)last, pw2, W.bnorm, W.H, W.cs, W.sn, W.g, W.est, W.inv, W.cyc, W.used,
                                   wj, lazy ? wj : W.V + (int64_t)(last ? j : j + 1) * vstride,
                                   lazy ? W.pscale : nullptr);
                HELM_HIP(hipGetLastError());
                continue;
            }
            hipLaunchKernelGGL(k_dcgs_reduce, dim3(nrhs), dim3(256), 0, cs, nchd, mm, kmax, j, o->tolerance,
                               W.partial, pd_part, W.bnorm, W.H, W.cs, W.sn, W.g, W.est, W.cyc, W.used, W.coef);
            if (prof_pair(pf, PK_UPDATE, &ea, &eb)) {
                hipExtLaunchKernelGGL(k_dcgs_update, dim3(nch, nrhs), blk, 0, cs, ea, eb, 0, n, nrhs, nch, kmax, j,
                                      W.V, W.w, (const double2*)W.coef, (const int*)W.cyc, pw2);
                prof_push(pf, PK_UPDATE, ea, eb, (j + 4) * vec_bytes);
            } else {
                hipLaunchKernelGGL(k_dcgs_update, dim3(nch, nrhs), blk, 0, cs, n, nrhs, nch, kmax, j, W.V, W.w,
                                   W.coef, W.cyc, pw2);
            }
            hipLaunchKernelGGL(k_dcgs_norm, dim3(nrhs), dim3(64), 0, cs, nch, mm, j, pw2, W.bnorm, W.H, W.cs, W.sn,
                               W.g, W.est, W.inv, W.cyc, W.used);
            if (j + 1 < nsteps)
                hipLaunchKernelGGL(k_scale32, gvec, blk, 0, cs, n, W.w, W.V + (int64_t)(j + 1) * vstride, W.inv,
                                   W.cyc);
            else
                hipLaunchKernelGGL(k_dcgs_finish, dim3((nrhs + 63) / 64), dim3(64), 0, cs, nrhs, mm, j, W.H,
                                   W.cs, W.sn, W.g, W.est, W.bnorm, W.cyc, W.used);
            HELM_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_xupdate, gvec, blk, 0, cs, n, nrhs, mm, W.H, W.g, W.Z, W.used, x);
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    };
    // one restart cycle; full cycles after the first (which performs any lazy
    // allocation) replay a captured graph
    auto run_cycle = [&](int msteps, int index) -> int32_t {
        if (o->use_graph && !pf && msteps == m && index > 1) {
            const uint64_t key[8] = {(uint64_t)m, (uint64_t)nrhs, (uint64_t)adj, (uint64_t)pc, (uint64_t)(uintptr_t)x,
                                     (uint64_t)(uintptr_t)p->net, p->generation, (uint64_t)(o->tolerance * 1e18)};
            return cycle_graph(W, key, enqueue_cycle, m, s);
        }
        return enqueue_cycle(msteps, s);
    };
    // The next cycle and its status are queued before the host reads the
    // current status (HELM_RESTART_AHEAD=0: after), so the device does not
    // idle while the host decides.  When the status says every RHS has
    // stopped, the cycle queued ahead changes nothing: each kernel skips RHS
    // with cyc = 0 (the multigrid whole groups of them), used stays 0 and x is
    // updated only over used > 0.  The iteration budget is known on the host,
    // so no cycle beyond it is queued.
    const bool ahead = restart_ahead();
    std::vector<int> prev_iters(nrhs, 0);
    // stagnation: a RHS whose true residual is above ratio x its value at least
    // `window` Krylov steps earlier (checked for the learned preconditioner, or
    // whenever the caller sets stagnation_steps)
    const bool watch = pc == HELM_PC_LEARNED || o->stagnation_steps > 0;
    const int window = o->stagnation_steps > 0 ? o->stagnation_steps : 60;
    const double ratio = o->stagnation_ratio > 0 ? o->stagnation_ratio : 0.5;
    std::vector<std::vector<std::pair<int, double>>> hist(nrhs);
    std::vector<int> stopped(nrhs, 0), fell(nrhs, 0);
    int slot = 0, last_steps = 0;
    HELM_TRY(enqueue_status(slot));
    applications = 1;
    for (;;) {
        const int msteps = std::min(m, o->max_iterations - steps);
        const size_t done = pf ? pf->pending.size() : 0;  // profiled launches up to this status
        const bool queued = ahead && msteps > 0;
        if (queued) {
            HELM_TRY(run_cycle(msteps, cycles + 1));
            HELM_TRY(enqueue_status(slot ^ 1));
        }
        HELM_HIP(hipEventSynchronize(W.stat_ev[slot]));
        {
            const char* hs = (const char*)W.hstat + (size_t)slot * W.stat_bytes;
            std::memcpy(W.h_rel.data(), hs, nrhs * sizeof(double));
            std::memcpy(W.h_active.data(), hs + (size_t)W.nrhs * sizeof(double), nrhs * sizeof(int));
            std::memcpy(W.h_iters.data(), hs + (size_t)W.nrhs * (sizeof(double) + sizeof(int)), nrhs * sizeof(int));
        }
        if (pf) {
            // the launches of the cycle that just ended enter the profile only
            // if every RHS ran all its steps (no launch skipped work)
            bool full = true;
            for (int r = 0; r < nrhs; ++r) full = full && W.h_iters[r] - prev_iters[r] == last_steps;
            prof_flush_first(pf, queued ? done : pf->pending.size(), full);
        }
        prev_iters = W.h_iters;
        if (o->callback) o->callback(o->user, applications, W.h_rel.data(), nrhs);
        if (watch) {
            bool stagnant = false;
            std::vector<int> which;
            for (int r = 0; r < nrhs; ++r) {
                if (!W.h_active[r] || stopped[r]) continue;
                auto& h = hist[r];
                h.emplace_back(W.h_iters[r], W.h_rel[r]);
                // the newest status at least `window` steps old
                const std::pair<int, double>* old_st = nullptr;
                for (const auto& e : h)
                    if (e.first <= W.h_iters[r] - window) old_st = &e;
                if (old_st && !(W.h_rel[r] <= ratio * old_st->second)) {
                    stagnant = true;
                    which.push_back(r);
                }
            }
            if (stagnant) {
                if (o->fallback && pc == HELM_PC_LEARNED && p->mg.levels > 0) {
                    // go on with the multigrid preconditioner for every RHS still running
                    pc = HELM_PC_MG;
                    for (int r = 0; r < nrhs; ++r)
                        if (W.h_active[r] && !stopped[r]) fell[r] = 1;
                    for (auto& h : hist) h.clear();
                } else {
                    for (int r : which) {
                        stopped[r] = 1;
                        hipLaunchKernelGGL(k_stop_rhs, dim3(1), dim3(1), 0, s, r, W.active, W.cyc);
                    }
                    HELM_HIP(hipGetLastError());
                }
            }
        }
        int n_active = 0;
        for (int r = 0; r < nrhs; ++r) n_active += W.h_active[r] && !stopped[r];
        if (n_active == 0 || steps >= o->max_iterations) {
            budget = n_active > 0;
            if (queued) {
                HELM_HIP(hipStreamSynchronize(s));
                if (pf) prof_flush_first(pf, pf->pending.size(), false);  // the no-op cycle
            }
            break;
        }
        vec_bytes = 8.0 * nn * n_active;
        if (!queued) {
            HELM_TRY(run_cycle(msteps, cycles + 1));
            HELM_TRY(enqueue_status(slot ^ 1));
        }
        cycles++;
        steps += msteps;
        last_steps = msteps;
        applications += msteps + 1;
        slot ^= 1;
    }
    for (int r = 0; r < nrhs; ++r) {
        info[r].reason = W.h_active[r] ? HELM_ITERATION_BUDGET : HELM_CONVERGED;
        if (!std::isfinite(W.h_rel[r])) info[r].reason = HELM_BREAKDOWN;
        if (stopped[r]) info[r].reason = HELM_STAGNATION;
        info[r].fallback = fell[r];
        info[r].first_reason = fell[r] ? HELM_STAGNATION : info[r].reason;
        info[r].graph_fallbacks = W.graph_fallbacks;
        info[r].iterations = W.h_iters[r];
        info[r].applications = W.h_iters[r] + cycles + 1;
        info[r].cycles = cycles;
        info[r].relative_residual = W.h_rel[r];
    }
    (void)budget;
    return HELM_OK;
}

// ---- reverse-communication FGMRES (helm_krylov_*) ---------------------------
// The Krylov state of the fused solver without a plan: the caller applies the
// operator and the preconditioner itself (any code, any stream order) and
// the library keeps the Hessenberg columns, the Givens rotations and the
// reduction partials.  V (m + 1 vectors), Z (m) and W (1) are caller-owned
// c64 arrays [k][nrhs][n]; x and residuals are c128 [nrhs][n].
struct KrylovState {
    int64_t n = 0;
    int nrhs = 0, m = 0;
    Workspace W;  // only the small per-RHS arrays and the partials are used
    std::vector<double> h_rel;
    std::vector<int> h_act;
};

static int32_t krylov_alloc(KrylovState& k) {
    Workspace& w = k.W;
    const int64_t n = k.n;
    const int nrhs = k.nrhs, restart = k.m, kmax = restart + 1;
    w.nrhs = nrhs;
    w.restart = restart;
    w.n = n;
    w.partial_len = (int64_t)nrhs * (2 * kmax * ndotchunks_max(n) + nchunks(n));
    HELM_TRY(alloc(&w.partial, (size_t)w.partial_len));
    HELM_TRY(alloc(&w.H, (size_t)nrhs * restart * (restart + 1)));
    HELM_TRY(alloc(&w.cs, (size_t)nrhs * restart));
    HELM_TRY(alloc(&w.sn, (size_t)nrhs * restart));
    HELM_TRY(alloc(&w.g, (size_t)nrhs * (restart + 1)));
    HELM_TRY(alloc(&w.coef, (size_t)nrhs * (2 * kmax + 1)));
    HELM_TRY(alloc(&w.beta, nrhs));
    HELM_TRY(alloc(&w.bnorm, nrhs));
    HELM_TRY(alloc(&w.rel, nrhs));
    HELM_TRY(alloc(&w.inv, nrhs));
    HELM_TRY(alloc(&w.est, nrhs));
    HELM_TRY(alloc(&w.active, nrhs));
    HELM_TRY(alloc(&w.cyc, nrhs));
    HELM_TRY(alloc(&w.used, nrhs));
    HELM_TRY(alloc(&w.iters, nrhs));
    k.h_rel.assign(nrhs, 0.0);
    k.h_act.assign(nrhs, 0);
    return HELM_OK;
}

static void krylov_free(KrylovState& k) {
    Workspace& w = k.W;
    void* bufs[] = {w.partial, w.H, w.cs, w.sn, w.g, w.coef, w.beta, w.bnorm, w.rel, w.inv, w.est, w.active,
                    w.cyc, w.used, w.iters};
    for (void* q : bufs)
        if (q) (void)hipFree(q);
}

}  // namespace helm

struct helm_krylov {
    helm::KrylovState st;
};

using namespace helm;

extern "C" int32_t helm_krylov_create(int64_t n, int32_t nrhs, int32_t restart, helm_krylov** out) {
    HELM_CHECK(out, HELM_ERR_ARG, "helm_krylov_create: null argument");
    *out = nullptr;
    HELM_CHECK(n >= 1 && nrhs >= 1 && nrhs <= 65535 && n * nrhs < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_krylov_create: n and nrhs out of range");
    HELM_CHECK(restart >= 1 && restart <= 63, HELM_ERR_ARG, "helm_krylov_create: restart must be in [1, 63]");
    helm_krylov* k = new helm_krylov();
    k->st.n = n;
    k->st.nrhs = nrhs;
    k->st.m = restart;
    const int32_t st = krylov_alloc(k->st);
    if (st != HELM_OK) {
        krylov_free(k->st);
        delete k;
        return st;
    }
    *out = k;
    return HELM_OK;
}

extern "C" int32_t helm_krylov_destroy(helm_krylov* k) {
    if (!k) return HELM_OK;
    krylov_free(k->st);
    delete k;
    return HELM_OK;
}

extern "C" int32_t helm_krylov_begin(helm_krylov* k, const void* b, const void* r, void* V, double tolerance,
                                     int32_t first, double* rel, int32_t* active, void* stream) {
    HELM_CHECK(k && r && V && rel && active && (b || !first), HELM_ERR_ARG, "helm_krylov_begin: null argument");
    HELM_CHECK(std::isfinite(tolerance) && tolerance > 0, HELM_ERR_ARG, "helm_krylov_begin: bad tolerance");
    KrylovState& K = k->st;
    Workspace& W = K.W;
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = K.n;
    const int nrhs = K.nrhs, nch = nchunks(n);
    const dim3 blk(256);
    if (first) {
        hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nch, nrhs), blk, 0, s, n, nch, (const double2*)b, W.partial);
        hipLaunchKernelGGL(k_bnorm, dim3(nrhs), dim3(64), 0, s, nch, W.partial, W.bnorm, W.active, W.iters);
        hipLaunchKernelGGL(k_bi_izero_k, dim3((nrhs + 255) / 256), blk, 0, s, nrhs, W.used);
    }
    hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nch, nrhs), blk, 0, s, n, nch, (const double2*)r, W.partial);
    hipLaunchKernelGGL(k_cycle_init, dim3(nrhs), dim3(64), 0, s, nch, K.m, tolerance, W.partial, W.bnorm, W.beta,
                       W.rel, W.inv, W.active, W.cyc, W.used, W.iters, W.g, (double*)nullptr, (double*)nullptr,
                       (int*)nullptr, (int*)nullptr, 0);
    hipLaunchKernelGGL(k_scale_to32, dim3((unsigned)((n + 255) / 256), nrhs), blk, 0, s, n, (const double2*)r,
                       (float2*)V, W.inv, W.cyc);
    HELM_HIP(hipGetLastError());
    HELM_HIP(hipMemcpyAsync(K.h_rel.data(), W.rel, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
    HELM_HIP(hipMemcpyAsync(K.h_act.data(), W.active, nrhs * sizeof(int), hipMemcpyDeviceToHost, s));
    HELM_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nrhs; ++i) {
        rel[i] = K.h_rel[i];
        active[i] = K.h_act[i];
    }
    return HELM_OK;
}

extern "C" int32_t helm_krylov_step(helm_krylov* k, int32_t j, int32_t last, void* V, void* w, double tolerance,
                                    void* stream) {
    HELM_CHECK(k && V && w, HELM_ERR_ARG, "helm_krylov_step: null argument");
    KrylovState& K = k->st;
    HELM_CHECK(j >= 0 && j < K.m, HELM_ERR_ARG, "helm_krylov_step: step index out of range");
    Workspace& W = K.W;
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = K.n;
    const int nrhs = K.nrhs, nch = nchunks(n), nchd = ndotchunks(n), kmax = K.m + 1, mm = K.m;
    const dim3 blk(256);
    const int64_t vstride = (int64_t)nrhs * n;
    double2* pd_part = W.partial + (int64_t)nrhs * kmax * nchd;
    double2* pw2 = W.partial + (int64_t)2 * nrhs * kmax * nchd;
    float2* Vb = (float2*)V;
    float2* wj = (float2*)w;
    const bool fin = last || j + 1 >= K.m;
    hipLaunchKernelGGL(dot_npt() == 4 ? k_dcgs_dots<4> : k_dcgs_dots<8>, dim3(nchd, nrhs), blk, 0, s, n, nrhs, nchd,
                       kmax, j, (const float2*)Vb, (const float2*)wj, (const int*)W.cyc, W.partial, pd_part);
    hipLaunchKernelGGL(k_dcgs_reduce_update, dim3(nch, nrhs), blk, 0, s, n, nrhs, nch, nchd, mm, kmax, j, tolerance,
                       (const double2*)W.partial, (const double2*)pd_part, (const double*)W.bnorm, W.H, W.cs, W.sn,
                       W.g, W.est, W.cyc, W.used, Vb, wj, pw2, 0);
    hipLaunchKernelGGL(k_dcgs_norm_scale, dim3(fin ? 1 : nch, nrhs), blk, 0, s, n, nch, mm, j, (int)fin,
                       (const double2*)pw2, (const double*)W.bnorm, W.H, W.cs, W.sn, W.g, W.est, W.inv, W.cyc,
                       W.used, (const float2*)wj, Vb + (int64_t)(fin ? j : j + 1) * vstride, (double*)nullptr);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

extern "C" int32_t helm_krylov_end(helm_krylov* k, const void* Z, void* x, int32_t* steps, void* stream) {
    HELM_CHECK(k && Z && x, HELM_ERR_ARG, "helm_krylov_end: null argument");
    KrylovState& K = k->st;
    Workspace& W = K.W;
    hipStream_t s = (hipStream_t)stream;
    const int nrhs = K.nrhs, kmax = K.m + 1;
    hipLaunchKernelGGL(k_xupdate, dim3((unsigned)((K.n + 255) / 256), nrhs), dim3(256), 0, s, K.n, nrhs, K.m, W.H,
                       W.g, (const float2*)Z, W.used, (double2*)x);
    HELM_HIP(hipGetLastError());
    if (steps) {
        std::vector<int> u(nrhs);
        HELM_HIP(hipMemcpyAsync(u.data(), W.used, nrhs * sizeof(int), hipMemcpyDeviceToHost, s));
        HELM_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < nrhs; ++i) steps[i] = u[i];
    }
    return HELM_OK;
}

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
    HELM_CHECK(o->precond >= HELM_PC_NONE && o->precond <= HELM_PC_SPECTRAL_PLAIN, HELM_ERR_ARG,
               "helm_solve: unknown preconditioner");
    HELM_CHECK((o->precond != HELM_PC_SPECTRAL && o->precond != HELM_PC_SPECTRAL_PLAIN) || p->fft_mult, HELM_ERR_STATE,
               "helm_solve: call helm_spectral_set_preconditioner before using the spectral preconditioner");
    HELM_CHECK(o->precond != HELM_PC_MG || p->mg.levels > 0, HELM_ERR_STATE,
               "helm_solve: call helm_mg_setup before using the multigrid preconditioner");
    hipStream_t s = (hipStream_t)stream;
    if (o->method == HELM_GMRES) return solve_gmres(p, (const double2*)rhs, (double2*)x, nrhs, o, info, s);
    if (o->method != HELM_BICGSTAB) {
        set_error("helm_solve: unknown method");
        return HELM_ERR_ARG;
    }
    HELM_TRY(solve_bicgstab(p, (const double2*)rhs, (double2*)x, nrhs, o, info, s));
    if (!o->fallback) return HELM_OK;
    // dense_wave.py:212-216: right-hand sides BiCGSTAB left unconverged go on
    // with GMRES and the fixed (here: multigrid) preconditioner
    bool any = false;
    for (int r = 0; r < nrhs; ++r) any = any || info[r].reason != HELM_CONVERGED;
    if (!any) return HELM_OK;
    const int64_t n = p->levels[0].n;
    for (int r = 0; r < nrhs; ++r)
        if (!std::isfinite(info[r].relative_residual))  // a non-finite iterate restarts from zero
            hipLaunchKernelGGL(k_zero_c128, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                               (double2*)x + (int64_t)r * n);
    HELM_HIP(hipGetLastError());
    helm_solve_opts g = *o;
    g.method = HELM_GMRES;
    g.precond = p->mg.levels > 0 ? HELM_PC_MG : HELM_PC_NONE;
    g.restart = std::min(24, std::max(1, p->desc.restart));
    g.use_x0 = 1;
    g.fallback = 0;
    g.stagnation_steps = 0;
    std::vector<helm_solve_info> gi(nrhs);
    HELM_TRY(solve_gmres(p, (const double2*)rhs, (double2*)x, nrhs, &g, gi.data(), s));
    for (int r = 0; r < nrhs; ++r) {
        if (info[r].reason == HELM_CONVERGED) continue;
        info[r].first_reason = info[r].reason;
        info[r].fallback = 1;
        info[r].reason = gi[r].reason;
        info[r].iterations += gi[r].iterations;
        info[r].applications += gi[r].applications;
        info[r].cycles += gi[r].cycles;
        info[r].relative_residual = gi[r].relative_residual;
        info[r].graph_fallbacks = gi[r].graph_fallbacks;
    }
    return HELM_OK;
}

extern "C" int32_t helm_warm_start(helm_plan* p, int32_t precond, const void* b, void* x, int32_t nrhs,
                                   int32_t steps, int32_t adjoint, void* stream) {
    HELM_CHECK(p && b && x, HELM_ERR_ARG, "helm_warm_start: null argument");
    HELM_CHECK(b != x, HELM_ERR_ARG, "helm_warm_start: b and x must not alias");
    HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_warm_start: set the medium first");
    HELM_CHECK(nrhs >= 1 && nrhs <= 65535 && (int64_t)nrhs * p->levels[0].n < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_warm_start: nrhs out of range");
    HELM_CHECK(steps >= 0 && steps <= 1000, HELM_ERR_ARG, "helm_warm_start: steps must be in [0, 1000]");
    HELM_CHECK(precond >= HELM_PC_NONE && precond <= HELM_PC_SPECTRAL_PLAIN, HELM_ERR_ARG,
               "helm_warm_start: unknown preconditioner");
    HELM_CHECK(precond == HELM_PC_NONE || precond == HELM_PC_SPECTRAL || precond == HELM_PC_SPECTRAL_PLAIN ||
                   p->mg.levels > 0, HELM_ERR_STATE,
               "helm_warm_start: call helm_mg_setup before using the multigrid preconditioner");
    HELM_CHECK((precond != HELM_PC_SPECTRAL && precond != HELM_PC_SPECTRAL_PLAIN) || p->fft_mult, HELM_ERR_STATE,
               "helm_warm_start: call helm_spectral_set_preconditioner first");
    return warm_start(p, precond, (const double2*)b, (double2*)x, nrhs, steps, adjoint != 0, (hipStream_t)stream);
}
