// Two-dimensional FFTs of periodic boxes and the Fourier-multiplier
// preconditioner built on them.
//
// The spectral operator's reference preconditions with a multiplier in
// Fourier space, d = gain * ifft2(M * fft2(r)) (nsrno/wave_operator.py:93-102,
// 129-135; nsrno/conditioned_wave.py:90-92 with gain = the medium mask).
// Box sides are multiples of 32 with arbitrary cofactors (H = 32 ceil((h +
// 64)/32)), so the transforms are mixed-radix Stockham passes: for a length N
// = R_1 R_2 ... and the product Ns of the radices already applied, pass R maps
//   v[r] = in[j + r N/R] e^{s 2 pi i k r/(Ns R)},  k = j mod Ns,
//   out[(j div Ns) Ns R + k + q Ns] = sum_r v[r] e^{s 2 pi i r q/R},
// which leaves the result in natural order after the last pass (the
// autosorting formulation of Govindaraju et al., SC 2008).  Radices 2-31
// (2, 3, 4, 5, 7, 8 and the primes 11-31) are register kernels; a larger
// prime factor runs a generic kernel that forms one output per thread (and
// so reads each input R times).  Rows (x, unit stride) and columns (z, stride
// W) have their own kernels so that a wavefront always reads consecutive
// complex numbers: along x it walks j, along z it walks x.  FP32 data with
// twiddle phases formed from exact integer products.
#include <algorithm>
#include <cmath>
#include <utility>

#include "helm_internal.h"

namespace helm {

// e^{2 pi i m / span} (sign folded into m's sign) in the precision of C
template <typename C> __device__ __forceinline__ C cis(int m, int span);
template <> __device__ __forceinline__ float2 cis<float2>(int m, int span) {
    float s, c;
    sincospif(2.f * (float)m / (float)span, &s, &c);
    return make_float2(c, s);
}
template <> __device__ __forceinline__ double2 cis<double2>(int m, int span) {
    double s, c;
    sincospi(2.0 * (double)m / (double)span, &s, &c);
    return make_double2(c, s);
}

// e^{2 pi i q / R} for compile-time R and q < R, evaluated at compile time:
// octant reduction in integers (angle (pi/4)(o + x) with o = 8q div R), then
// Taylor series on [0, pi/4] in double (error below 1e-16)
constexpr double CE_PI = 3.14159265358979323846264338327950288;
constexpr double ce_sin_r(double x) {
    double t = x, s = x;
    for (int n = 1; n < 12; ++n) {
        t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        s += t;
    }
    return s;
}
constexpr double ce_cos_r(double x) {
    double t = 1.0, s = 1.0;
    for (int n = 1; n < 12; ++n) {
        t *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        s += t;
    }
    return s;
}
struct CeCis {
    double c, s;
};
constexpr CeCis ce_cis(int q, int R) {
    const int m = 8 * (q % R), o = m / R, rr = m - o * R;
    double c = 0.0, s = 0.0;
    if (o % 2 == 0) {
        const double x = CE_PI / 4.0 * rr / R;
        c = ce_cos_r(x);
        s = ce_sin_r(x);
    } else {
        const double y = CE_PI / 4.0 * (R - rr) / R;
        c = ce_sin_r(y);
        s = ce_cos_r(y);
    }
    switch ((o / 2) % 4) {
        case 1: return CeCis{-s, c};
        case 2: return CeCis{-c, -s};
        case 3: return CeCis{s, -c};
        default: return CeCis{c, s};
    }
}
template <int R, typename Seq>
struct RootTab;
template <int R, int... Q>
struct RootTab<R, std::integer_sequence<int, Q...>> {
    static constexpr double re[sizeof...(Q)] = {ce_cis(Q, R).c...};
    static constexpr double im[sizeof...(Q)] = {ce_cis(Q, R).s...};
};
// e^{sign 2 pi i q / R} (q < R) from the compile-time table, in C's precision
template <typename C, int R>
__device__ __forceinline__ C root(int q, int sign) {
    using T = RootTab<R, std::make_integer_sequence<int, R>>;
    C w;
    w.x = (decltype(w.x))T::re[q];
    w.y = (decltype(w.y))(sign < 0 ? -T::im[q] : T::im[q]);
    return w;
}

// v * e^{sign 2 pi i q / R} for compile-time q and R: the multiples of a
// quarter turn as sign flips and swaps, the others as one complex product
template <typename C, int R, int Q>
__device__ __forceinline__ C cmul_root(C v, int sign) {
    constexpr int q = Q % R;
    if constexpr (q == 0) {
        return v;
    } else if constexpr (2 * q == R) {
        v.x = -v.x;
        v.y = -v.y;
        return v;
    } else if constexpr (4 * q == R || 4 * q == 3 * R) {
        // times +-i: i (x + iy) = -y + ix
        const auto s = (4 * q == R) ? sign : -sign;
        C o;
        o.x = s > 0 ? -v.y : v.y;
        o.y = s > 0 ? v.x : -v.x;
        return o;
    } else {
        return cmul(v, root<C, R>(q, sign));
    }
}

template <typename C, int R, int... Q>
__device__ __forceinline__ void roots_times(C* o, int sign, std::integer_sequence<int, Q...>) {
    ((o[Q] = cmul_root<C, R, Q>(o[Q], sign)), ...);
}
// a[r1][q2] *= w_{R1 R2}^(r1 q2)
template <typename C, int R1, int R2, int... I>
__device__ __forceinline__ void inner_twiddles(C (&a)[R1][R2], int sign, std::integer_sequence<int, I...>) {
    ((a[I / R2][I % R2] = cmul_root<C, R1 * R2, (I / R2) * (I % R2)>(a[I / R2][I % R2], sign)), ...);
}

// In-place R-point DFT X[q] = sum_r x[r] e^{sign 2 pi i r q / R} for a
// compile-time R.  Even R: one radix-2 split (two R/2-point DFTs of the even
// and odd samples, X[q] = E[q] + w^q O[q], X[q + R/2] = E[q] - w^q O[q]).
// Odd R: the symmetric pairs a_r = x_r + x_{R-r}, b_r = x_r - x_{R-r}
// (r = 1..(R-1)/2) give
//   X[0] = x_0 + sum a_r,  X[q], X[R-q] = P_q +- i sign Q_q,
//   P_q = x_0 + sum_r a_r cos(2 pi r q / R),  Q_q = sum_r b_r sin(2 pi r q / R)
// with real coefficients: (R-1)^2 FMAs instead of 4 (R-1) R.
template <typename C, int R>
__device__ __forceinline__ void dft_small(C (&x)[R], int sign);
template <typename C, int R>
__device__ __forceinline__ void dft_small(C (&x)[R], int sign) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        const C a = x[0], b = x[1];
        x[0].x = a.x + b.x;
        x[0].y = a.y + b.y;
        x[1].x = a.x - b.x;
        x[1].y = a.y - b.y;
    } else if constexpr (R % 2 == 0) {
        constexpr int H = R / 2;
        C e[H], o[H];
#pragma unroll
        for (int r = 0; r < H; ++r) {
            e[r] = x[2 * r];
            o[r] = x[2 * r + 1];
        }
        dft_small<C, H>(e, sign);
        dft_small<C, H>(o, sign);
        roots_times<C, R>(o, sign, std::make_integer_sequence<int, H>{});
#pragma unroll
        for (int q = 0; q < H; ++q) {
            x[q].x = e[q].x + o[q].x;
            x[q].y = e[q].y + o[q].y;
            x[q + H].x = e[q].x - o[q].x;
            x[q + H].y = e[q].y - o[q].y;
        }
    } else {
        constexpr int H = (R - 1) / 2;
        using T = RootTab<R, std::make_integer_sequence<int, R>>;
        using F = decltype(x[0].x);
        C a[H], b[H];
        C x0 = x[0], s0 = x[0];
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            a[r - 1].x = x[r].x + x[R - r].x;
            a[r - 1].y = x[r].y + x[R - r].y;
            b[r - 1].x = x[r].x - x[R - r].x;
            b[r - 1].y = x[r].y - x[R - r].y;
            s0.x += a[r - 1].x;
            s0.y += a[r - 1].y;
        }
        x[0] = s0;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            C P = x0, Q;
            Q.x = 0;
            Q.y = 0;
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const F c = (F)T::re[(r * q) % R], sn = (F)T::im[(r * q) % R];
                P.x += a[r - 1].x * c;
                P.y += a[r - 1].y * c;
                Q.x += b[r - 1].x * sn;
                Q.y += b[r - 1].y * sn;
            }
            // i sign Q = sign (-Q.y, Q.x)
            const F iqx = sign > 0 ? -Q.y : Q.y, iqy = sign > 0 ? Q.x : -Q.x;
            x[q].x = P.x + iqx;
            x[q].y = P.y + iqy;
            x[R - q].x = P.x - iqx;
            x[R - q].y = P.y - iqy;
        }
    }
}

// pass twiddles: v[r] *= e^{sign 2 pi i k r / (Ns R)} for r = r1 + R1 r2,
// formed as the product of the r1 and R1 r2 factors (R1 + R2 - 2 sine-cosine
// evaluations instead of R - 1)
template <typename C, int R1, int R2>
__device__ __forceinline__ void fft_twiddle(C (&v)[R1 * R2], int k, int Ns, int sign) {
    constexpr int R = R1 * R2;
    const int span = Ns * R;
    if (Ns == 1) return;  // k = 0
    C t1[R1], t2[R2];
#pragma unroll
    for (int r1 = 1; r1 < R1; ++r1) t1[r1] = cis<C>(sign * (int)(((int64_t)k * r1) % span), span);
#pragma unroll
    for (int r2 = 1; r2 < R2; ++r2) t2[r2] = cis<C>(sign * (int)(((int64_t)k * R1 * r2) % span), span);
#pragma unroll
    for (int r2 = 0; r2 < R2; ++r2)
#pragma unroll
        for (int r1 = 0; r1 < R1; ++r1) {
            if (r1 == 0 && r2 == 0) continue;
            const C w = r2 == 0 ? t1[r1] : (r1 == 0 ? t2[r2] : cmul(t1[r1], t2[r2]));
            v[r1 + R1 * r2] = cmul(v[r1 + R1 * r2], w);
        }
}

template <typename C, int R>
__device__ __forceinline__ void fft_butterfly(C (&v)[R], int k, int Ns, int sign) {
    const int span = Ns * R;
    if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const int m = (int)(((int64_t)k * r) % span);
            v[r] = cmul(v[r], cis<C>(sign * m, span));
        }
    }
    dft_small<C, R>(v, sign);
}

// Composite radix R = R1 R2 in one pass (fewer passes over the data): the
// pass twiddles, then the R-point DFT by Cooley-Tukey in registers, with
// r = r1 + R1 r2 and q = q2 + R2 q1:
//   A[r1][q2] = sum_r2 v[r1 + R1 r2] w_R2^(r2 q2)       (R1 DFTs of size R2)
//   out[q2 + R2 q1] = sum_r1 (w_R^(r1 q2) A[r1][q2]) w_R1^(r1 q1)
template <typename C, int R1, int R2>
__device__ __forceinline__ void fft_butterfly2(C (&v)[R1 * R2], int k, int Ns, int sign) {
    fft_twiddle<C, R1, R2>(v, k, Ns, sign);
    C a[R1][R2];
#pragma unroll
    for (int r1 = 0; r1 < R1; ++r1) {
#pragma unroll
        for (int r2 = 0; r2 < R2; ++r2) a[r1][r2] = v[r1 + R1 * r2];
        dft_small<C, R2>(a[r1], sign);
    }
    inner_twiddles<C, R1, R2>(a, sign, std::make_integer_sequence<int, R1 * R2>{});
#pragma unroll
    for (int q2 = 0; q2 < R2; ++q2) {
        C b[R1];
#pragma unroll
        for (int r1 = 0; r1 < R1; ++r1) b[r1] = a[r1][q2];
        dft_small<C, R1>(b, sign);
#pragma unroll
        for (int q1 = 0; q1 < R1; ++q1) v[q2 + R2 * q1] = b[q1];
    }
}
template <typename C, int R1, int R2>
__device__ __forceinline__ void fft_pass_butterfly(C (&v)[R1 * R2], int k, int Ns, int sign) {
    if constexpr (R2 == 1) fft_butterfly<C, R1>(v, k, Ns, sign);
    else fft_butterfly2<C, R1, R2>(v, k, Ns, sign);
}

// transform along x: lines = rows of all batches (H * B of them), length W
// Pass options the preconditioner folds into its first and last passes: the
// conjugate of the input on load (conj_in), and on store
// out = conj?(mul[node] (scale v)) (scale 1 and no mul: the plain output;
// nodes of the [H][W] box, the same for every batch)
template <typename C>
struct FftEpi {
    int conj_in = 0;
    const double* bscale = nullptr;  // first pass: input of batch b scaled by bscale[b]
    const C* mul = nullptr;
    float scale = 1.f;
    int conj_out = 0;
};
template <typename C>
__device__ __forceinline__ C fft_epi(C v, const FftEpi<C>& e, int64_t node) {
    if (e.scale != 1.f) {
        v.x *= e.scale;
        v.y *= e.scale;
    }
    if (e.mul) v = cmul(e.mul[node], v);
    if (e.conj_out) v.y = -v.y;
    return v;
}

// STAGE (passes with Ns < 16): the outputs o + q Ns of a thread are Ns apart,
// so lanes j and j + 1 store R elements apart and every store touches 64
// lines.  The block (256 threads = 256 / M whole lines) then writes its lines
// into LDS and copies them out contiguously.
extern __shared__ __align__(16) unsigned char fft_rows_lds[];
template <typename C, int R1, int R2 = 1, bool STAGE = false>
__global__ __launch_bounds__(256) void k_fft_rows(int H, int W, int64_t lines, int Ns, int sign,
                                                 const C* __restrict__ in, C* __restrict__ out, FftEpi<C> ep) {
    constexpr int R = R1 * R2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int M = W / R;
    const bool live = t < (int64_t)M * lines;
    if (!STAGE && !live) return;
    const int j = (int)(t % M);
    const int64_t line = t / M, base = line * W;
    const int k = j % Ns;
    const int o = (j / Ns) * Ns * R + k;
    C v[R];
    if (live) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[base + j + r * M];
        if (ep.bscale) {
            const float bs = (float)ep.bscale[line / H];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v[r].x *= bs;
                v[r].y *= bs;
            }
        }
        if (ep.conj_in) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r].y = -v[r].y;
        }
        fft_pass_butterfly<C, R1, R2>(v, k, Ns, sign);
    }
    if constexpr (!STAGE) {
        const int64_t nrow = (line % H) * W;
#pragma unroll
        for (int q = 0; q < R; ++q) out[base + o + q * Ns] = fft_epi(v[q], ep, nrow + o + q * Ns);
    } else {
        C* sb = reinterpret_cast<C*>(fft_rows_lds);
        const int L = (int)blockDim.x / M;
        const int64_t line0 = (int64_t)blockIdx.x * L;
        if (live) {
            const int lo = (int)(line - line0) * W + o;
#pragma unroll
            for (int q = 0; q < R; ++q) sb[lo + q * Ns] = v[q];
        }
        __syncthreads();
        const int n = (int)min((int64_t)L, lines - line0) * W;
        int pos = (int)threadIdx.x, ln = 0;
        while (pos >= W) {
            pos -= W;
            ++ln;
        }
        for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) {
            out[line0 * W + i] = fft_epi(sb[i], ep, ((line0 + ln) % H) * W + pos);
            pos += (int)blockDim.x;
            while (pos >= W) {
                pos -= W;
                ++ln;
            }
        }
    }
}

// Both passes of a two-pass row plan (W = RA RB, passes (A1 x A2) with Ns = 1
// then (B1 x B2) with Ns = RA) in one kernel: each block holds FUSE_L whole
// lines in LDS between the passes, so the rows go through HBM once instead of
// twice.  Every block reads only its own lines before it writes them, so in
// may equal out.
template <int FUSE_L, int RA, int RB>
constexpr int fuse_threads() {
    return (FUSE_L * (RA > RB ? RA : RB) + 63) / 64 * 64;
}
template <typename C, int A1, int A2, int B1, int B2, int FUSE_L>
__global__ __launch_bounds__((fuse_threads<FUSE_L, A1 * A2, B1 * B2>())) void k_fft_rows_2p(
    int H, int64_t lines, int sign, const C* in, C* out, FftEpi<C> ep_first, FftEpi<C> ep_last) {
    constexpr int RA = A1 * A2, RB = B1 * B2, W = RA * RB;
    C* sb = reinterpret_cast<C*>(fft_rows_lds);
    const int64_t line0 = (int64_t)blockIdx.x * FUSE_L;
    const int nl = (int)min((int64_t)FUSE_L, lines - line0);
    const int tid = (int)threadIdx.x;
    if (tid < nl * RB) {
        const int li = tid / RB, j = tid % RB;
        const int64_t base = (line0 + li) * W;
        C v[RA];
#pragma unroll
        for (int r = 0; r < RA; ++r) v[r] = in[base + j + r * RB];
        if (ep_first.bscale) {
            const float bs = (float)ep_first.bscale[(line0 + li) / H];
#pragma unroll
            for (int r = 0; r < RA; ++r) {
                v[r].x *= bs;
                v[r].y *= bs;
            }
        }
        if (ep_first.conj_in) {
#pragma unroll
            for (int r = 0; r < RA; ++r) v[r].y = -v[r].y;
        }
        fft_pass_butterfly<C, A1, A2>(v, 0, 1, sign);
#pragma unroll
        for (int q = 0; q < RA; ++q) sb[li * W + j * RA + q] = v[q];
    }
    __syncthreads();
    if (tid < nl * RA) {
        const int li = tid / RA, k = tid % RA;
        C v[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) v[r] = sb[li * W + k + r * RA];
        fft_pass_butterfly<C, B1, B2>(v, k, RA, sign);
        const int64_t base = (line0 + li) * W, nrow = ((line0 + li) % H) * W;
#pragma unroll
        for (int q = 0; q < RB; ++q) out[base + k + q * RA] = fft_epi(v[q], ep_last, nrow + k + q * RA);
    }
}

// transform along z: for every batch and column x, length H with stride W
template <typename C, int R1, int R2 = 1>
__global__ __launch_bounds__(256) void k_fft_cols(int H, int W, int B, int Ns, int sign, const C* __restrict__ in,
                                                 C* __restrict__ out, FftEpi<C> ep) {
    constexpr int R = R1 * R2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int M = H / R;
    if (t >= (int64_t)W * M * B) return;
    const int x = (int)(t % W);
    const int j = (int)((t / W) % M);
    const int64_t base = (t / ((int64_t)W * M)) * H * W + x;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[base + (int64_t)(j + r * M) * W];
    if (ep.bscale) {
        const float bs = (float)ep.bscale[t / ((int64_t)W * M)];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r].x *= bs;
            v[r].y *= bs;
        }
    }
    if (ep.conj_in) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r].y = -v[r].y;
    }
    const int k = j % Ns;
    fft_pass_butterfly<C, R1, R2>(v, k, Ns, sign);
    const int o = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q)
        out[base + (int64_t)(o + q * Ns) * W] = fft_epi(v[q], ep, (int64_t)(o + q * Ns) * W + x);
}

// any radix: one output (line, j, q) per thread; stride = element stride
// of the transformed axis, lines addressed as base = (l / per) * bstride +
// (l % per) * lstride
template <typename C>
__global__ __launch_bounds__(256) void k_fft_generic(int N, int R, int64_t lines, int per, int64_t bstride,
                                                     int64_t lstride, int64_t stride, int Ns, int sign,
                                                     const C* __restrict__ in, C* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int M = N / R;
    if (t >= (int64_t)N * lines) return;
    const int q = (int)(t % R);
    const int j = (int)((t / R) % M);
    const int64_t l = t / N;
    const int64_t base = (l / per) * bstride + (l % per) * lstride;
    const int k = j % Ns;
    const int span = Ns * R;
    C acc;
    acc.x = 0;
    acc.y = 0;
    for (int r = 0; r < R; ++r) {
        const int m1 = (int)(((int64_t)k * r) % span);
        const int m2 = (r * q) % R;
        // e^{s 2 pi i (k r/(Ns R) + r q / R)} = e^{s 2 pi i (k r + r q Ns) / (Ns R)}
        const int m = (int)(((int64_t)m1 + (int64_t)m2 * Ns) % span);
        cfma(acc, in[base + (int64_t)(j + r * M) * stride], cis<C>(sign * m, span));
    }
    const int o = (j / Ns) * Ns * R + k;
    out[base + (int64_t)(o + q * Ns) * stride] = acc;
}

static std::vector<int> fft_radices(int n) {
    std::vector<int> r;
    while (n % 8 == 0) { r.push_back(8); n /= 8; }
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    for (int p = 3; n > 1; p += 2)
        while (n % p == 0) { r.push_back(p); n /= p; }
    return r;
}

template <typename C, int R1, int R2 = 1>
static void launch_rows(int H, int W, int64_t lines, int Ns, int sign, const C* in, C* out, const FftEpi<C>& ep,
                        hipStream_t s) {
    constexpr int R = R1 * R2;
    const int M = W / R;
    const int64_t cnt = (int64_t)M * lines;
    const dim3 grid((unsigned)((cnt + 255) / 256));
    // staged stores: whole lines per block, at most 80 KB of LDS (256 R elements)
    constexpr size_t stage_bytes = (size_t)256 * R * sizeof(C);
    if (Ns * (int)sizeof(C) < 128 && 256 % M == 0 && stage_bytes <= 80 * 1024) {
        static bool attr = false;
        if (!attr) {
            attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_rows<C, R1, R2, true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)stage_bytes) == hipSuccess;
        }
        if (attr) {
            hipLaunchKernelGGL((k_fft_rows<C, R1, R2, true>), grid, dim3(256), stage_bytes, s, H, W, lines, Ns, sign,
                               in, out, ep);
            return;
        }
    }
    hipLaunchKernelGGL((k_fft_rows<C, R1, R2>), grid, dim3(256), 0, s, H, W, lines, Ns, sign, in, out, ep);
}
template <typename C, int R1, int R2 = 1>
static void launch_cols(int H, int W, int B, int Ns, int sign, const C* in, C* out, const FftEpi<C>& ep,
                        hipStream_t s) {
    const int64_t cnt = (int64_t)W * (H / (R1 * R2)) * B;
    hipLaunchKernelGGL((k_fft_cols<C, R1, R2>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, H, W, B, Ns,
                       sign, in, out, ep);
}

// Composite passes (FP32 transforms): (R1, R2) pairs with kernels, and the
// plan that packs a side's factors into them: odd primes paired (product up
// to 40), powers of two attached to single odd primes, the rest of the power
// of two in passes of 32 (8 x 4).  1216 = (2 x 19)(8 x 4) and 1056 =
// (3 x 11)(8 x 4): two passes each instead of three and four.
#define HELM_FFT_PAIRS(X) \
    X(8, 4) X(8, 2) X(2, 3) X(2, 5) X(2, 7) X(2, 11) X(2, 13) X(2, 17) X(2, 19) X(4, 3) X(4, 5) X(4, 7) X(8, 3) \
    X(8, 5) X(3, 5) X(3, 7) X(3, 11) X(3, 13) X(5, 7)
static bool fft_pair_ok(int r1, int r2) {
#define HELM_FFT_PAIR_EQ(a, b) if (r1 == a && r2 == b) return true;
    HELM_FFT_PAIRS(HELM_FFT_PAIR_EQ)
#undef HELM_FFT_PAIR_EQ
    return false;
}
// HELM_FFT_COMPOSITE=0: one pass per prime-or-power radix in FP32 too
// lines per workgroup of the fused row kernel: 4 (39 KB of LDS for 1216-wide
// lines, four workgroups per CU, so one's loads overlap another's stores)
constexpr int FFT_FUSE_L = 4;
// fused two-pass row kernel (HELM_FFT_ROWS_FUSED=0: one kernel per pass)
static bool fft_rows_fused_on() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_FFT_ROWS_FUSED");
        return !(v && std::atoi(v) == 0);
    }();
    return t;
}
static bool fft_composite_on() {
    static const bool t = [] {
        const char* v = std::getenv("HELM_FFT_COMPOSITE");
        return !(v && std::atoi(v) == 0);
    }();
    return t;
}
static std::vector<std::pair<int, int>> fft_plan(int n, bool composite) {
    std::vector<std::pair<int, int>> passes;
    if (composite) {
        int m = n, a = 0;
        while (m % 2 == 0) {
            m /= 2;
            ++a;
        }
        std::vector<int> odd;
        for (int p = 3; m > 1; p += 2)
            while (m % p == 0) {
                odd.push_back(p);
                m /= p;
            }
        bool ok = true;
        for (int p : odd) ok = ok && p <= 31;
        if (ok) {
            std::sort(odd.rbegin(), odd.rend());
            std::vector<bool> used(odd.size(), false);
            for (size_t i = 0; i < odd.size(); ++i) {
                if (used[i]) continue;
                used[i] = true;
                int mate = -1;
                for (size_t j = odd.size(); j-- > i + 1;)
                    if (!used[j] && fft_pair_ok(odd[j], odd[i])) {
                        mate = (int)j;
                        break;
                    }
                if (mate >= 0) {
                    used[mate] = true;
                    passes.push_back({odd[mate], odd[i]});
                } else {
                    passes.push_back({odd[i], 1});
                }
            }
            for (auto& pr : passes) {
                if (pr.second != 1) continue;
                for (int b = 8, lb = 3; b >= 2; b /= 2, --lb)
                    if (a >= lb && fft_pair_ok(b, pr.first)) {
                        pr = {b, pr.first};
                        a -= lb;
                        break;
                    }
            }
            while (a >= 5) {
                passes.push_back({8, 4});
                a -= 5;
            }
            if (a == 4) passes.push_back({8, 2});
            else if (a == 3) passes.push_back({8, 1});
            else if (a == 2) passes.push_back({4, 1});
            else if (a == 1) passes.push_back({2, 1});
            return passes;
        }
    }
    for (int r : fft_radices(n)) passes.push_back({r, 1});
    return passes;
}

// One axis of a 2-D transform of B boxes [H][W]: src -> (passes ping-pong
// between a and b) -> returns the buffer holding the result.  The first pass
// takes first.conj_in, the last pass last's store options and writes to
// last_out when given (the generic kernel of a radix above 31 takes neither).
template <typename C>
static C* fft_axis(bool rows, int H, int W, int B, int sign, const C* src, C* a, C* b, hipStream_t s,
                   const FftEpi<C>& first = FftEpi<C>(), const FftEpi<C>& last = FftEpi<C>(),
                   C* last_out = nullptr) {
    const int N = rows ? W : H;
    const std::vector<std::pair<int, int>> rad = fft_plan(N, sizeof(C) == sizeof(float2) && fft_composite_on());
    const C* in = src;
    C* out = a;
    if constexpr (sizeof(C) == sizeof(float2)) {
        // two-pass row plans with a fused kernel (1216 = (2 x 19)(8 x 4))
        if (rows && rad.size() == 2 && rad[0] == std::make_pair(2, 19) && rad[1] == std::make_pair(8, 4) &&
            fft_rows_fused_on()) {
            constexpr int L = FFT_FUSE_L;
            constexpr size_t lds = (size_t)L * 1216 * sizeof(C);
            static bool attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_rows_2p<C, 2, 19, 8, 4, L>),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)lds) == hipSuccess;
            if (attr) {
                const int64_t lines = (int64_t)H * B;
                C* dst = last_out ? last_out : b;
                hipLaunchKernelGGL((k_fft_rows_2p<C, 2, 19, 8, 4, L>), dim3((unsigned)((lines + L - 1) / L)),
                                   dim3(fuse_threads<L, 38, 32>()), lds, s, H, lines, sign, src, dst, first, last);
                return dst;
            }
        }
    }
    int Ns = 1;
    for (size_t p = 0; p < rad.size(); ++p) {
        const int R = rad[p].first * rad[p].second;
        FftEpi<C> ep;
        if (p == 0) {
            ep.conj_in = first.conj_in;
            ep.bscale = first.bscale;
        }
        if (p + 1 == rad.size()) {
            ep.mul = last.mul;
            ep.scale = last.scale;
            ep.conj_out = last.conj_out;
            if (last_out) out = last_out;
        }
        if (rad[p].second != 1) {
            if constexpr (sizeof(C) == sizeof(float2)) {  // composite plans are FP32 only
                const int r1 = rad[p].first, r2 = rad[p].second;
                const int64_t lines = (int64_t)H * B;
#define HELM_FFT_PAIR_LAUNCH(A, Bq)                                                           \
    if (r1 == A && r2 == Bq) {                                                                \
        if (rows) launch_rows<C, A, Bq>(H, W, lines, Ns, sign, in, out, ep, s);               \
        else launch_cols<C, A, Bq>(H, W, B, Ns, sign, in, out, ep, s);                        \
    }
                HELM_FFT_PAIRS(HELM_FFT_PAIR_LAUNCH)
#undef HELM_FFT_PAIR_LAUNCH
            }
        } else if (rows) {
            const int64_t lines = (int64_t)H * B;
            switch (R) {
                case 2: launch_rows<C, 2>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 3: launch_rows<C, 3>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 4: launch_rows<C, 4>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 5: launch_rows<C, 5>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 7: launch_rows<C, 7>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 8: launch_rows<C, 8>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 11: launch_rows<C, 11>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 13: launch_rows<C, 13>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 17: launch_rows<C, 17>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 19: launch_rows<C, 19>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 23: launch_rows<C, 23>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 29: launch_rows<C, 29>(H, W, lines, Ns, sign, in, out, ep, s); break;
                case 31: launch_rows<C, 31>(H, W, lines, Ns, sign, in, out, ep, s); break;
                default: {
                    const int64_t cnt = (int64_t)W * lines;
                    hipLaunchKernelGGL(k_fft_generic<C>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, W, R,
                                       lines, (int)std::min<int64_t>(lines, 1 << 30), (int64_t)0, (int64_t)W,
                                       (int64_t)1, Ns, sign, in, out);
                }
            }
        } else {
            switch (R) {
                case 2: launch_cols<C, 2>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 3: launch_cols<C, 3>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 4: launch_cols<C, 4>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 5: launch_cols<C, 5>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 7: launch_cols<C, 7>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 8: launch_cols<C, 8>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 11: launch_cols<C, 11>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 13: launch_cols<C, 13>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 17: launch_cols<C, 17>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 19: launch_cols<C, 19>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 23: launch_cols<C, 23>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 29: launch_cols<C, 29>(H, W, B, Ns, sign, in, out, ep, s); break;
                case 31: launch_cols<C, 31>(H, W, B, Ns, sign, in, out, ep, s); break;
                default: {
                    const int64_t lines = (int64_t)W * B;
                    const int64_t cnt = (int64_t)H * lines;
                    hipLaunchKernelGGL(k_fft_generic<C>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, H, R,
                                       lines, W, (int64_t)H * W, (int64_t)1, (int64_t)W, Ns, sign, in, out);
                }
            }
        }
        Ns *= R;
        in = out;
        out = (out == a) ? b : a;
    }
    return (C*)in;
}

// 2-D transform of B boxes (sign -1 forward, +1 inverse without 1/N);
// returns the buffer (a or b) that holds the result.  src may be a or b
// only if it is not overwritten before it is read (the first pass reads src
// and writes a).
template <typename C>
C* fft2(int H, int W, int B, int sign, const C* src, C* a, C* b, hipStream_t s, int conj_in = 0,
        const FftEpi<C>& last = FftEpi<C>(), C* last_out = nullptr, const double* bscale = nullptr) {
    FftEpi<C> first;
    first.conj_in = conj_in;
    first.bscale = bscale;
    C* r = fft_axis<C>(true, H, W, B, sign, src, a, b, s, first);
    C* other = (r == a) ? b : a;
    return fft_axis<C>(false, H, W, B, sign, r, other, r, s, FftEpi<C>(), last, last_out);
}
// largest radix of a side (the pass-per-radix kernels take options up to 31)
static int fft_max_radix(int n) {
    int m = 1;
    for (int r : fft_radices(n)) m = std::max(m, r);
    return m;
}

// z[r] = conj?(gain (.) buf[r] / N) ; buf = M (.) fft2(conj?(v))  (two kernels)
__global__ void k_sp_conj_copy(int64_t cnt, const float2* __restrict__ v, float2* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) out[e] = cconj(v[e]);
}
__global__ void k_sp_mult(int64_t n, int64_t cnt, const float2* __restrict__ M, float2* __restrict__ x) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) x[e] = cmul(M[e % n], x[e]);
}
__global__ void k_sp_gain(int64_t n, int64_t cnt, float scale, int conj, const float2* __restrict__ gain,
                          const float2* __restrict__ x, float2* __restrict__ z) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cnt) return;
    float2 v = cscale(x[e], scale);
    if (gain) v = cmul(gain[e % n], v);
    z[e] = conj ? cconj(v) : v;
}

bool spectral_scaled_input_ok(const helm_plan* p) {
    const Level& L = p->levels[0];
    return p->fft_mult && fft_max_radix(L.nz) <= 31 && fft_max_radix(L.nx) <= 31;
}

int32_t spectral_precond(helm_plan* p, const float2* v, float2* z, int nrhs, bool adjoint, hipStream_t s,
                         bool use_gain, const double* fscale) {
    HELM_CHECK(p->fft_mult, HELM_ERR_STATE, "spectral preconditioner: call helm_spectral_set_preconditioner first");
    HELM_CHECK(nrhs <= p->fft_cap, HELM_ERR_ARG, "spectral preconditioner: more right-hand sides than its buffers");
    const Level& L = p->levels[0];
    const int H = L.nz, W = L.nx;
    const int64_t n = L.n, cnt = (int64_t)nrhs * n;
    const dim3 g((unsigned)((cnt + 255) / 256)), blk(256);
    float2 *a = p->fft_a, *b = p->fft_b;
    // The map as given, d = gain ifft2(M fft2(v)), preconditions the operator
    // the multiplier was designed for: the plan's own operator (forward
    // solves) except on spectral plans, which store A = k0^2 conj(L) and whose
    // reference operator L is reached by adjoint solves.  The other direction
    // takes the conjugate map conj(P(conj v)) (for a multiplier even in k and a
    // real gain this is the reference's adjoint preconditioner gain ifft2(conj(M)
    // fft2(v)), dense_wave.py:195-196).
    const bool as_given = L.op == HELM_OP_STAG8 ? adjoint : !adjoint;
    HELM_CHECK(!fscale || spectral_scaled_input_ok(p), HELM_ERR_STATE,
               "spectral preconditioner: scaled input needs the fused passes");
    if (fft_max_radix(H) <= 31 && fft_max_radix(W) <= 31) {
        // the per-RHS input scale and the conjugate of the input in the first
        // pass, M in the last forward pass, 1/N, the gain and the conjugate in
        // the last inverse pass, which writes z
        FftEpi<float2> m;
        m.mul = p->fft_mult;
        float2* f = fft2<float2>(H, W, nrhs, -1, v, a, b, s, (int)!as_given, m, nullptr, fscale);
        float2* fo = (f == a) ? b : a;
        FftEpi<float2> g;
        g.scale = (float)(1.0 / (double)n);
        g.mul = use_gain ? p->fft_gain : nullptr;
        g.conj_out = (int)!as_given;
        fft2<float2>(H, W, nrhs, +1, f, fo, f, s, 0, g, z);
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }
    const float2* src = v;
    if (!as_given) {
        hipLaunchKernelGGL(k_sp_conj_copy, g, blk, 0, s, cnt, v, b);
        src = b;
    }
    float2* f = fft2<float2>(H, W, nrhs, -1, src, a, b, s);
    hipLaunchKernelGGL(k_sp_mult, g, blk, 0, s, n, cnt, (const float2*)p->fft_mult, f);
    float2* fo = (f == a) ? b : a;
    float2* r = fft2<float2>(H, W, nrhs, +1, f, fo, f, s);
    hipLaunchKernelGGL(k_sp_gain, g, blk, 0, s, n, cnt, (float)(1.0 / (double)n), (int)!as_given,
                       use_gain ? (const float2*)p->fft_gain : nullptr, (const float2*)r, z);
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

__global__ void k_c128_to_c64(int64_t n, const double2* __restrict__ src, float2* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = make_float2((float)src[e].x, (float)src[e].y);
}

}  // namespace helm

using namespace helm;

extern "C" {

int32_t helm_spectral_set_preconditioner(helm_plan* p, const void* multiplier, const void* gain, void* stream) {
    HELM_CHECK(p && multiplier, HELM_ERR_ARG, "helm_spectral_set_preconditioner: null argument");
    const Level& L = p->levels[0];
    HELM_CHECK(L.nz >= 2 && L.nx >= 2, HELM_ERR_ARG, "helm_spectral_set_preconditioner: box too small");
    const int64_t n = L.n;
    const int cap = std::max(1, p->desc.nrhs);
    hipStream_t s = (hipStream_t)stream;
    auto take = [&](float2** q, int64_t count) -> int32_t {
        if (*q) return HELM_OK;
        hipError_t e = hipMalloc((void**)q, (size_t)count * sizeof(float2) + 256);
        if (e != hipSuccess) {
            *q = nullptr;
            set_error(std::string("hipMalloc (spectral preconditioner): ") + hipGetErrorString(e));
            return HELM_ERR_ALLOC;
        }
        return HELM_OK;
    };
    HELM_TRY(take(&p->fft_mult, n));
    HELM_TRY(take(&p->fft_a, (int64_t)cap * n));
    HELM_TRY(take(&p->fft_b, (int64_t)cap * n));
    p->fft_cap = cap;
    hipLaunchKernelGGL(k_c128_to_c64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                       (const double2*)multiplier, p->fft_mult);
    if (gain) {
        HELM_TRY(take(&p->fft_gain, n));
        hipLaunchKernelGGL(k_c128_to_c64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                           (const double2*)gain, p->fft_gain);
    } else if (p->fft_gain) {
        (void)hipFree(p->fft_gain);
        p->fft_gain = nullptr;
    }
    HELM_HIP(hipGetLastError());
    p->generation++;
    return HELM_OK;
}

int32_t helm_fft2(int32_t H, int32_t W, int32_t batch, int32_t inverse, int32_t precision, const void* in, void* out,
                  void* work, void* stream) {
    HELM_CHECK(in && out && work && in != out && in != work && out != work, HELM_ERR_ARG,
               "helm_fft2: null or aliased argument");
    HELM_CHECK(H >= 2 && W >= 2 && batch >= 1 && (int64_t)H * W * batch < ((int64_t)1 << 31), HELM_ERR_ARG,
               "helm_fft2: sizes out of range");
    HELM_CHECK(precision == 32 || precision == 64, HELM_ERR_ARG, "helm_fft2: precision must be 32 or 64");
    hipStream_t s = (hipStream_t)stream;
    const int sign = inverse ? +1 : -1;
    void* r;
    size_t el;
    if (precision == 32) {
        r = fft2<float2>(H, W, batch, sign, (const float2*)in, (float2*)out, (float2*)work, s);
        el = sizeof(float2);
    } else {
        r = fft2<double2>(H, W, batch, sign, (const double2*)in, (double2*)out, (double2*)work, s);
        el = sizeof(double2);
    }
    if (r != out) HELM_HIP(hipMemcpyAsync(out, r, (size_t)H * W * batch * el, hipMemcpyDeviceToDevice, s));
    HELM_HIP(hipGetLastError());
    return HELM_OK;
}

}  // extern "C"
