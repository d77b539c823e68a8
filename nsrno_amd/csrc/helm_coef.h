// Per-node coefficients of the flux-form stencil (shared by the operator,
// the multigrid smoother and the coarse-level builder).
#pragma once
#include "helm_internal.h"

namespace helm {

template <typename C>
struct Coef {
    C dg, fw, fe, fn, fs;
};

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
    const C szi = sz[i], sxj = sx[j];
    c.fw = cscale(cmul(szi, isxe[j]), hh * (ec + ew));
    c.fe = cscale(cmul(szi, isxe[j + 1]), hh * (ec + ee));
    c.fn = cscale(cmul(sxj, isze[i]), hh * (ec + en));
    c.fs = cscale(cmul(sxj, isze[i + 1]), hh * (ec + es));
    C mass = cscale(cmul(cmul(szi, sxj), kap[idx]), k0sq);
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
}  // namespace helm
