// Per-node coefficients of the flux-form stencil (shared by the operator,
// the multigrid smoother and the coarse-level builder).
#pragma once
#include "helm_internal.h"

namespace helm {

template <typename C>
struct Coef {
    C dg, fw, fe, fn, fs;
};

// The stencil of node (i, j) from its operands: eta at the node and its four
// neighbours (1 beyond the grid edge), the row and column stretch factors,
// the inverse edge stretches on both sides, and kappa.
template <typename R, typename C>
__device__ __forceinline__ void node_coef_form(R ec, R ew, R ee, R en, R es, C szi, C sxj, C ixw, C ixe, C izn,
                                               C izs, C kp, R k0sq, R hh, R shift, bool conj, Coef<C>& c) {
    c.fw = cscale(cmul(szi, ixw), hh * (ec + ew));
    c.fe = cscale(cmul(szi, ixe), hh * (ec + ee));
    c.fn = cscale(cmul(sxj, izn), hh * (ec + en));
    c.fs = cscale(cmul(sxj, izs), hh * (ec + es));
    C mass = cscale(cmul(cmul(szi, sxj), kp), k0sq);
    if (shift != R(0)) {  // complex shifted operator: kappa_c -> kappa_c (1 + i shift)
        C m2 = mass;
        m2.x = mass.x - shift * mass.y;
        m2.y = mass.y + shift * mass.x;
        mass = m2;
    }
    c.dg = csub(mass, cadd(cadd(c.fw, c.fe), cadd(c.fn, c.fs)));
    if (conj) {
        c.dg = cconj(c.dg);
        c.fw = cconj(c.fw);
        c.fe = cconj(c.fe);
        c.fn = cconj(c.fn);
        c.fs = cconj(c.fs);
    }
}

template <typename R, typename C>
__device__ __forceinline__ void node_coef(int i, int j, int nz, int nx, const C* __restrict__ sx,
                                          const C* __restrict__ sz, const C* __restrict__ isxe,
                                          const C* __restrict__ isze, const R* __restrict__ eta,
                                          const C* __restrict__ kap, R k0sq, R hh, R shift, bool conj,
                                          Coef<C>& c) {
    const int64_t idx = (int64_t)i * nx + j;
    const R ec = eta[idx];
    const R ew = j > 0 ? eta[idx - 1] : R(1);
    const R ee = j < nx - 1 ? eta[idx + 1] : R(1);
    const R en = i > 0 ? eta[idx - nx] : R(1);
    const R es = i < nz - 1 ? eta[idx + nx] : R(1);
    node_coef_form<R, C>(ec, ew, ee, en, es, sz[i], sx[j], isxe[j], isxe[j + 1], isze[i], isze[i + 1], kap[idx],
                         k0sq, hh, shift, conj, c);
}

// Operands of one node's FP32 stencil, loaded with every address clamped into
// its array so that no load sits behind a branch.  A kernel that forms the
// stencils of several nodes per thread loads all of them first
// (node_raw_load) and forms them after (node_raw_form): the loads are then in
// flight together, one memory latency in all, where loads behind bounds
// branches are waited for node by node.
struct NodeRaw {
    float ec, ew, ee, en, es;
    float2 szi, sxj, ixw, ixe, izn, izs, kap;
};
// ok false: any in-range node is loaded, and the result must not be used
__device__ __forceinline__ void node_raw_load(int i, int j, bool ok, int nz, int nx, const float2* __restrict__ sx,
                                              const float2* __restrict__ sz, const float2* __restrict__ isxe,
                                              const float2* __restrict__ isze, const float* __restrict__ eta,
                                              const float2* __restrict__ kap, NodeRaw& r) {
    i = ok ? i : 0;
    j = ok ? j : 0;
    const int64_t idx = (int64_t)i * nx + j;
    r.ec = eta[idx];
    r.ew = eta[j > 0 ? idx - 1 : idx];
    r.ee = eta[j < nx - 1 ? idx + 1 : idx];
    r.en = eta[i > 0 ? idx - nx : idx];
    r.es = eta[i < nz - 1 ? idx + nx : idx];
    r.szi = sz[i];
    r.sxj = sx[j];
    r.ixw = isxe[j];
    r.ixe = isxe[j + 1];
    r.izn = isze[i];
    r.izs = isze[i + 1];
    r.kap = kap[idx];
}
// the stencil of node (i, j) from node_raw_load's operands: the same
// arithmetic as node_coef
__device__ __forceinline__ void node_raw_form(const NodeRaw& r, int i, int j, int nz, int nx, float k0sq, float hh,
                                              float shift, bool conj, Coef<float2>& c) {
    const float ew = j > 0 ? r.ew : 1.f;
    const float ee = j < nx - 1 ? r.ee : 1.f;
    const float en = i > 0 ? r.en : 1.f;
    const float es = i < nz - 1 ? r.es : 1.f;
    node_coef_form<float, float2>(r.ec, ew, ee, en, es, r.szi, r.sxj, r.ixw, r.ixe, r.izn, r.izs, r.kap, k0sq, hh,
                                  shift, conj, c);
}
}  // namespace helm
