// Kernels around the solve: the Born Jacobian contraction, the running DFT
// that turns time snapshots into frequency fields, and band-limited RF
// synthesis from solved spectra.
//
// * helm_born_jacobian: nsrno/theranostic/wave.py:331-389.  At zero contrast
//   the derivative of the port data d[tx, rx] with respect to basis
//   coefficient b is a reciprocal product of background fields without
//   conjugation,
//     J_kappa[p, b] = sum_q wn[q, b] F[q, tx_p] F[q, rx_p],
//     J_eta[p, b]   = sum_e we[e, b] G[e, tx_p] G[e, rx_p],
//     G[e, s]       = F[j_e, s] - F[i_e, s]   (a ghost end j_e < 0 is zero),
//   with the node weights wn = h^2 k0^2 sz sx B and the edge weights
//   we = -stretch_e (avg B)_e (the h^2 of the sum and the 1/h^2 of the two
//   differences cancel).  One thread per (pair, basis function): it walks the
//   basis function's support (CSC taps) and gathers the two fields of its pair
//   from the node's row F[q, :] (the rows of the background fields are
//   contiguous, so the pairs of a wavefront hit the same cache lines).
// * helm_capture_accumulate: nsrno/native_capture.cu:4-11, the running DFT
//   fields[f][i] += p_k[i] phase[f][k] over captured pressure snapshots p_k,
//   with K snapshots per launch: each thread holds its node's K pressures in
//   registers and the launch's F x K phase table sits in LDS, so the fields
//   are read and written once per K snapshots (HBM bytes n (4 K + 32 F)
//   instead of K n (4 + 32 F)).
// * helm_band_rf: nsrno/wave_operator.py:169-180, irfft(..., norm='forward')
//   of a spectrum that is zero except at the given bins:
//     rf[r][t] = sum_b c_b Re(X[r][b] e^{2 pi i k_b t / N}),  c_b = 2, or 1 at
//   k = 0 and at k = N/2; the phase (k t mod N) is exact integer arithmetic
//   into an FP64 twiddle table in LDS, the sum in FP64.
#include <cmath>

#include "helm_internal.h"

namespace helm {

__global__ __launch_bounds__(256) void k_born_jacobian(int n_src, int n_pairs, int n_basis,
                                                      const double2* __restrict__ F,
                                                      const int* __restrict__ tx, const int* __restrict__ rx,
                                                      const int* __restrict__ node_ptr,
                                                      const int* __restrict__ node_idx,
                                                      const double2* __restrict__ node_w,
                                                      const int* __restrict__ edge_ptr,
                                                      const int* __restrict__ edge_i, const int* __restrict__ edge_j,
                                                      const double2* __restrict__ edge_w,
                                                      double2* __restrict__ J) {
    const int b = blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int t = tx[p], r = rx[p];
    double2 ak = make_double2(0.0, 0.0), ae = make_double2(0.0, 0.0);
    for (int k = node_ptr[b]; k < node_ptr[b + 1]; ++k) {
        const double2* row = F + (int64_t)node_idx[k] * n_src;
        cfma(ak, node_w[k], cmul(row[t], row[r]));
    }
    for (int k = edge_ptr[b]; k < edge_ptr[b + 1]; ++k) {
        const double2* ri = F + (int64_t)edge_i[k] * n_src;
        const int j = edge_j[k];
        double2 gt = make_double2(-ri[t].x, -ri[t].y), gr = make_double2(-ri[r].x, -ri[r].y);
        if (j >= 0) {
            const double2* rj = F + (int64_t)j * n_src;
            gt = cadd(gt, rj[t]);
            gr = cadd(gr, rj[r]);
        }
        cfma(ae, edge_w[k], cmul(gt, gr));
    }
    J[(int64_t)p * 2 * n_basis + b] = ak;
    J[(int64_t)p * 2 * n_basis + n_basis + b] = ae;
}

constexpr int CAP_K = 32;     // snapshots per launch
constexpr int CAP_F = 64;     // frequencies per pass over the phase table in LDS

__global__ __launch_bounds__(256) void k_capture(int64_t n, int K, int F, int captures, int first,
                                                 const float* __restrict__ pressure,
                                                 const double2* __restrict__ phase, double2* __restrict__ fields) {
    __shared__ double2 ph[CAP_F * CAP_K];
    float pk[CAP_K];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < n;
#pragma unroll
    for (int k = 0; k < CAP_K; ++k) pk[k] = (in && k < K) ? pressure[(int64_t)k * n + i] : 0.f;
    for (int f0 = 0; f0 < F; f0 += CAP_F) {
        const int nf = min(CAP_F, F - f0);
        __syncthreads();
        for (int q = threadIdx.x; q < nf * K; q += blockDim.x) {
            const int f = q / K, k = q % K;
            ph[f * CAP_K + k] = phase[(int64_t)(f0 + f) * captures + first + k];
        }
        __syncthreads();
        if (!in) continue;
        // four frequencies per pass: their four loads are in flight together
        for (int f = 0; f < nf; f += 4) {
            double2 acc[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[q] = f + q < nf ? fields[(int64_t)(f0 + f + q) * n + i] : make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < CAP_K; ++k) {
                if (k < K) {
                    const double pv = (double)pk[k];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double2 w = ph[min(f + q, nf - 1) * CAP_K + k];
                        acc[q].x = fma(pv, w.x, acc[q].x);
                        acc[q].y = fma(pv, w.y, acc[q].y);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (f + q < nf) fields[(int64_t)(f0 + f + q) * n + i] = acc[q];
        }
    }
}

constexpr int RF_MAXN = 8192;  // samples the twiddle table (LDS, 128 KiB) holds

__global__ __launch_bounds__(256) void k_band_rf(int R, int nb, int nt, const float2* __restrict__ spec,
                                                 const int* __restrict__ bins, float* __restrict__ rf) {
    extern __shared__ double2 tw[];  // tw[m] = e^{2 pi i m / nt}
    for (int m = threadIdx.x; m < nt; m += blockDim.x) {
        double s, c;
        sincospi(2.0 * m / nt, &s, &c);
        tw[m] = make_double2(c, s);
    }
    __syncthreads();
    const int r = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const float2* xr = spec + (int64_t)r * nb;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) {
        const int k = bins[b];
        const int64_t kk = ((int64_t)k % nt + nt) % nt;  // in [0, nt) whatever the caller passed
        const int m = (int)((kk * t) % nt);
        const double2 w = tw[m];
        const float2 x = xr[b];
        const double re = (double)x.x * w.x - (double)x.y * w.y;
        acc += (k == 0 || 2 * k == nt) ? re : 2.0 * re;
    }
    rf[(int64_t)r * nt + t] = (float)acc;
}

}  // namespace helm

using namespace helm;

extern "C" {

int32_t helm_born_jacobian(int64_t n_nodes, int32_t n_src, const void* fields, int32_t n_pairs, const int32_t* tx,
                           const int32_t* rx, int32_t n_basis, const int32_t* node_ptr, const int32_t* node_idx,
                           const void* node_w, const int32_t* edge_ptr, const int32_t* edge_i,
                           const int32_t* edge_j, const void* edge_w, void* J, void* stream) {
    HELM_CHECK(fields && tx && rx && node_ptr && edge_ptr && J, HELM_ERR_ARG, "helm_born_jacobian: null argument");
    HELM_CHECK(n_nodes >= 1 && n_src >= 1 && n_nodes * n_src < ((int64_t)1 << 40), HELM_ERR_ARG,
               "helm_born_jacobian: field sizes out of range");
    HELM_CHECK(n_pairs >= 1 && n_basis >= 1 && n_basis <= 65535 && (int64_t)n_pairs * 2 * n_basis < ((int64_t)1 << 40),
               HELM_ERR_ARG, "helm_born_jacobian: pairs or basis out of range");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_born_jacobian, dim3((n_pairs + 255) / 256, n_basis), dim3(256), 0, s, n_src, n_pairs,
                       n_basis, (const double2*)fields, tx, rx, node_ptr, node_idx, (const double2*)node_w, edge_ptr,
                       edge_i, edge_j, (const double2*)edge_w, (double2*)J);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

int32_t helm_capture_accumulate(const float* pressure, int32_t snapshots, const void* phase, int32_t first,
                                int32_t captures, int32_t frequencies, int64_t n, void* fields, void* stream) {
    HELM_CHECK(pressure && phase && fields, HELM_ERR_ARG, "helm_capture_accumulate: null argument");
    HELM_CHECK(snapshots >= 1 && snapshots <= CAP_K, HELM_ERR_ARG,
               "helm_capture_accumulate: 1 to 32 snapshots per call");
    HELM_CHECK(frequencies >= 1 && n >= 1 && n * frequencies < ((int64_t)1 << 40), HELM_ERR_ARG,
               "helm_capture_accumulate: sizes out of range");
    HELM_CHECK(first >= 0 && captures >= 1 && first + snapshots <= captures, HELM_ERR_ARG,
               "helm_capture_accumulate: capture index out of range");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_capture, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, snapshots, frequencies,
                       captures, first, pressure, (const double2*)phase, (double2*)fields);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

int32_t helm_band_rf(const void* spectra, const int32_t* bins, int32_t n_bins, int32_t traces, int32_t samples,
                     float* rf, void* stream) {
    HELM_CHECK(spectra && bins && rf, HELM_ERR_ARG, "helm_band_rf: null argument");
    HELM_CHECK(samples >= 1 && samples <= RF_MAXN, HELM_ERR_ARG, "helm_band_rf: 1 to 8192 samples");
    HELM_CHECK(n_bins >= 1 && traces >= 1 && traces <= 65535, HELM_ERR_ARG, "helm_band_rf: sizes out of range");
    hipStream_t s = (hipStream_t)stream;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)k_band_rf, hipFuncAttributeMaxDynamicSharedMemorySize, RF_MAXN * 16);
    HELM_HIP(attr);
    hipLaunchKernelGGL(k_band_rf, dim3((samples + 255) / 256, traces), dim3(256), (size_t)samples * sizeof(double2),
                       s, traces, n_bins, samples, (const float2*)spectra, bins, rf);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

}  // extern "C"
