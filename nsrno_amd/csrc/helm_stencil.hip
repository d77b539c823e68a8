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
    return launch_flux32(L, k0sq, u, out, nrhs, adjoint, b, shift, s);
}

int32_t launch_apply64(const Level& L, double k0sq, const double2* u, double2* out, int nrhs, bool adjoint,
                       const double2* b, hipStream_t s) {
    if (L.op == HELM_OP_STAG8) return launch_stag8_64(L, k0sq, u, out, nrhs, adjoint, b, b ? 1 : 0, s);
    return launch_flux64(L, k0sq, u, out, nrhs, adjoint, b, s);
}

int32_t launch_flux32(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                      const float2* b, float shift, hipStream_t s) {
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
