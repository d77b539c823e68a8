// Internal declarations shared by the translation units of libhelm.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/helm.h"

namespace helm {

// ---------------------------------------------------------------- complex
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// a += b * c
__host__ __device__ inline void cfma(float2& a, float2 b, float2 c) {
    a.x = fmaf(b.x, c.x, fmaf(-b.y, c.y, a.x));
    a.y = fmaf(b.x, c.y, fmaf(b.y, c.x, a.y));
}
__host__ __device__ inline void cfma(double2& a, double2 b, double2 c) {
    a.x = fma(b.x, c.x, fma(-b.y, c.y, a.x));
    a.y = fma(b.x, c.y, fma(b.y, c.x, a.y));
}
// Block indices of a launch as values: the hardware's, or one segment's
// share of a segmented launch (one launch over the grids of several plans,
// the batched solve).  Kernel bodies take a VBlock instead of reading
// blockIdx / gridDim, so that the same body serves both launches.
struct VBlock {
    int x, y, z, gx, gy, gz;
};
__device__ __forceinline__ VBlock hw_block() {
    return VBlock{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z,
                  (int)gridDim.x,  (int)gridDim.y,  (int)gridDim.z};
}

// Segmented launch: a 1-D grid in which segment s owns the workgroups
// [start_s, start_s + gx gy gz); starts increase and are multiples of 8, so
// that a workgroup's XCD (its linear id mod 8) is that of its local id.
struct SegDims {
    int start, gx, gy, gz;
};
constexpr int SEG_MAX = 256;  // segments per launch (four ballots)
// Segment and local block of this workgroup; false for the padding
// workgroups at a segment's end.  Each wave finds the segment with one
// ballot per 64 segments (lane l tests segment base + l).
__device__ __forceinline__ bool seg_block(const SegDims* __restrict__ dims, int nseg, int& seg, VBlock& B) {
    const int b = blockIdx.x;
    const int lane = __lane_id();
    int cnt = 0;
    for (int base = 0; base < nseg; base += 64) {
        const int i = base + lane;
        cnt += __popcll(__ballot(i < nseg && dims[i].start <= b));
    }
    const int s = __builtin_amdgcn_readfirstlane(cnt - 1);
    const SegDims d = dims[s];
    const int lb = b - d.start, tot = d.gx * d.gy * d.gz;
    seg = s;
    if (lb >= tot) return false;
    B = VBlock{lb % d.gx, (lb / d.gx) % d.gy, lb / (d.gx * d.gy), d.gx, d.gy, d.gz};
    return true;
}

// XCD-aware tile order: workgroups are handed to the 8 XCDs of the device in
// turn (workgroup b to XCD b mod 8), each XCD with its own L2.  Remapping the
// linear workgroup id so that each XCD works on one contiguous run of tiles
// keeps the halos that neighbouring tiles share in one L2.  A bijection of
// [0, total): XCD k takes ids [start_k, start_k + count_k).  With rhs_fast
// the z index (the right-hand-side group) runs fastest in that order, so the
// groups of one tile share an XCD and read the tile's stencil coefficients
// and halos through one L2; otherwise it runs slowest.
__device__ __forceinline__ void xcd_tile(const VBlock& B, int& bx, int& by, int& bz, bool on = true,
                                         bool rhs_fast = false) {
    if (!on) {
        bx = B.x;
        by = B.y;
        bz = B.z;
        return;
    }
    const int gx = B.gx, gy = B.gy, gz = B.gz;
    const int total = gx * gy * gz;
    const int b = B.x + gx * (B.y + gy * B.z);
    const int q = total / 8, r = total % 8;
    const int k = b % 8, slot = b / 8;
    const int t = k < r ? k * (q + 1) + slot : r * (q + 1) + (k - r) * q + slot;
    if (rhs_fast) {
        bz = t % gz;
        bx = (t / gz) % gx;
        by = t / (gz * gx);
    } else {
        bx = t % gx;
        by = (t / gx) % gy;
        bz = t / (gx * gy);
    }
}
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz, bool on = true, bool rhs_fast = false) {
    xcd_tile(hw_block(), bx, by, bz, on, rhs_fast);
}

// Workgroup barrier that orders LDS only.  __syncthreads() fences global
// memory too, which on this target waits for every outstanding vector-memory
// load (vmcnt(0)) and so drains a prefetch issued before the barrier; where a
// barrier only separates LDS writes from LDS reads, this one keeps the
// prefetch in flight.  Not for kernels whose threads communicate through
// global memory.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// keep ? v : 0 and keep ? v : alt formed with bit masks.  A select on a
// loaded value that the compiler turns into a branch also sinks the load into
// that branch, where it is then waited for as soon as it is issued; a mask
// keeps the load unconditional, so it can be issued ahead with the others.
__device__ __forceinline__ float2 keep_or_zero(float2 v, bool keep) {
    const int m = -(int)keep;
    return make_float2(__int_as_float(__float_as_int(v.x) & m), __int_as_float(__float_as_int(v.y) & m));
}
__device__ __forceinline__ float keep_or(float v, bool keep, float alt) {
    const int m = -(int)keep;
    return __int_as_float((__float_as_int(v) & m) | (__float_as_int(alt) & ~m));
}

// Load through a pointer known to address global memory.  A pointer read
// from memory (an argument table) is a generic pointer to the compiler, and
// generic loads count against the LDS counter as well, so waits for them also
// wait for LDS operations; this emits a global load.  (The host pass of the
// same source never executes device code and keeps a plain load.)
__device__ __forceinline__ float2 ld_global(const float2* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    using g64 = const __attribute__((address_space(1))) unsigned long long*;
    const unsigned long long v = *(g64)(const unsigned long long*)p;  // C-style: an address-space cast
    float2 r;
    __builtin_memcpy(&r, &v, sizeof(r));
    return r;
#else
    return *p;
#endif
}

template <typename R> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
// host-to-device copy, complete on return, that avoids the legacy null stream
int32_t upload_h2d(void* dst, const void* src, size_t bytes);
#define HELM_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            helm::set_error(std::string(#call) + ": " + hipGetErrorString(e_));           \
            return HELM_ERR_HIP;                                                          \
        }                                                                                 \
    } while (0)
#define HELM_CHECK(cond, code, msg)                                                       \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            helm::set_error(msg);                                                         \
            return code;                                                                  \
        }                                                                                 \
    } while (0)
#define HELM_TRY(call)                                                                    \
    do {                                                                                  \
        int32_t s_ = (call);                                                              \
        if (s_ != HELM_OK) return s_;                                                     \
    } while (0)

// ---------------------------------------------------------------- geometry
// One grid level: 1-D stretch profiles and per-node coefficients.
struct Level {
    int nz = 0, nx = 0;
    int64_t n = 0;
    double h_m = 0;
    // FP64 and FP32 copies; profile arrays: sx[nx], sz[nz], 1/sxe[nx+1], 1/sze[nz+1]
    double2 *sx64 = nullptr, *sz64 = nullptr, *isxe64 = nullptr, *isze64 = nullptr;
    float2 *sx32 = nullptr, *sz32 = nullptr, *isxe32 = nullptr, *isze32 = nullptr;
    double* eta64 = nullptr;   // eta_rel [n]
    double2* kap64 = nullptr;  // kappa_c [n]
    float* eta32 = nullptr;
    float2* kap32 = nullptr;
    // multigrid work vectors (c64, nrhs * n each)
    float2 *mg_u = nullptr, *mg_f = nullptr, *mg_r = nullptr, *mg_t = nullptr;
    int mg_cap = 0;  // right-hand sides the work vectors hold
    float2* wgt = nullptr;  // per-node complex Jacobi weight
    float2* wid = nullptr;  // per-node Jacobi factor w / dg of the shifted operator
    float2* tcf = nullptr;  // [n][6] precomputed shifted stencil for the LDS tail
    // HELM_OP_STAG8 (level 0 of a spectral plan only): periodic 8-tap operator
    // with face coefficients ax[i][j] (face j+1/2), az[i][j] (face i+1/2); the
    // eta/kap arrays above then hold the 5-point approximation the multigrid uses
    int op = HELM_OP_FLUX5;
    float *ax32 = nullptr, *az32 = nullptr;
    double *ax64 = nullptr, *az64 = nullptr;
    int sp_cells = 0;      // sponge width
    double sp_amp = 0.0;   // sponge peak
    // HELM_OP_COND8 (level 0 only): masked split-PML operator; ax/az above hold
    // the masked face means, cmass the node mass, ctab the per-column [4][nx]
    // then per-row [4][nz] factors (sv, g - 1, cA, cB), cbox the native box
    // (z0, x0, h, w); eta/kap/sx/.. hold the multigrid's 5-point approximation
    double2* cmass64 = nullptr;
    float2* cmass32 = nullptr;
    // conjugate mass (FP32), set where the adjoint differs from the operator
    // only in it (the S operator: real faces): its adjoint is then the forward
    // kernel with this mass
    float2* cmassh32 = nullptr;
    double2* ctab64 = nullptr;
    float2* ctab32 = nullptr;
    int cbox[4] = {0, 0, 0, 0};
    // streaming-stencil launch shape chosen by helm_autotune (0: default)
    int st_tz = 0, st_rb = 0;
    bool st_block = false;  // FP32 flux stencil: block form (whole run loaded up front)
    int st_xcd = 0;         // workgroup order: 0 hardware, 1 XCD-contiguous, 2 the same with RHS groups fastest
    bool st_pair = false;   // paired block form (two columns per lane)
};

struct Workspace {
    int nrhs = 0, restart = 0;
    int64_t n = 0;
    float2* V = nullptr;     // (m+1) * nrhs * n
    float2* Z = nullptr;     // m * nrhs * n (flexible directions)
    float2* w = nullptr;     // nrhs * n
    double2* r64 = nullptr;  // nrhs * n (true residual)
    double2* partial = nullptr;  // reduction partials
    int64_t partial_len = 0;
    // per-RHS small state (device)
    double2* H = nullptr;     // nrhs * m * (m+1): rotated Hessenberg columns
    double* cs = nullptr;     // nrhs * m
    double2* sn = nullptr;    // nrhs * m
    double2* g = nullptr;     // nrhs * (m+1)
    double2* coef = nullptr;  // nrhs * (m+1)
    double* beta = nullptr;   // nrhs
    double* bnorm = nullptr;  // nrhs
    double* rel = nullptr;    // nrhs
    double* inv = nullptr;    // nrhs (1/h or 1/beta)
    double* pscale = nullptr; // nrhs: scale the preconditioner applies to an unnormalised basis vector
    double* est = nullptr;    // nrhs
    int* active = nullptr;    // nrhs: not yet converged
    int* cyc = nullptr;       // nrhs: active in this cycle
    int* used = nullptr;      // nrhs: steps used in this cycle
    int* iters = nullptr;     // nrhs: Krylov steps so far
    // BiCGSTAB vectors (c64, nrhs * n each) and scalars
    float2 *bp = nullptr, *bv = nullptr, *bs = nullptr, *bt = nullptr, *bphat = nullptr, *bshat = nullptr,
           *bshadow = nullptr, *br = nullptr;
    double2* bsc = nullptr;   // nrhs * 8 scalars
    size_t bi_cap = 0;        // elements the BiCGSTAB vectors hold
    // captured restart cycle (helm_solve_opts.use_graph) and what it was captured for
    hipGraphExec_t gexec = nullptr;
    uint64_t gkey[8] = {};
    hipStream_t gstream = nullptr;  // capture stream (capture is not allowed on the null stream)
    int graph_fallbacks = 0;        // cycles of the current solve enqueued directly after a failed capture
    // host mirrors
    std::vector<double> h_rel;
    std::vector<int> h_active;
    std::vector<int> h_iters;
    // restart status written by the device straight into pinned, mapped host
    // memory (rel, active, iters per RHS): no copies per restart.  Two slots of
    // stat_bytes, used in turn, so the status of the cycle queued ahead does
    // not overwrite the one the host is reading
    void* hstat = nullptr;      // host block (hipHostMalloc)
    char* dstat = nullptr;      // its device-side address; per slot: rel [nrhs], active [nrhs], iters [nrhs]
    size_t stat_bytes = 0;
    hipEvent_t stat_ev[2] = {nullptr, nullptr};  // recorded after each slot's k_cycle_init
};

// ---- segmented launches of the batched solve (helm_batch_solve) -----------
// One plan's share of a segmented Krylov / stencil launch: its workspace,
// grid and operator.  Filled on the host, read by the kernels (uniform per
// workgroup, so the fields come in as scalar loads).
struct SegKrylov {
    int64_t n;
    int nrhs, nch, nchd, nchb;  // RHS; chunks of the norm/update, dots and residual-block reductions
    int nz, nx, ws_nrhs;        // grid; RHS capacity of the workspace (status slot layout)
    int kmax, mm;               // basis capacity + 1 and the Hessenberg row stride of the workspace
    float2 *V, *Z;
    double2 *partial, *H, *sn, *g, *coef;
    double *cs, *bnorm, *beta, *rel, *inv, *pscale, *est;
    int *active, *cyc, *used, *iters;
    double2* x;
    const double2* b;
    char* dstat;        // status slots (device view of the mapped host block)
    int64_t stat_bytes;
    // FP64 operator (restart residual) and FP32 operator (Krylov steps)
    const double2 *sx64, *sz64, *isxe64, *isze64, *kap64;
    const double* eta64;
    double k0sq, hh;
    const float2 *sx32, *sz32, *isxe32, *isze32, *kap32;
    const float* eta32;
    float k0sq32, hh32;
};
// One plan's grid level l of the fused V-cycle (pre = post = 2): the
// operands of k_mgt_pre2_rr and k_mgt_post2 on that level.  On the finest
// level f and u are the step's basis and direction vectors: f + j f_step.
struct SegLevel {
    int nz, nx, nzc, nxc, nrhs, rb_pre, rb_post, xcd_pre, xcd_post;
    int64_t n;
    const float2 *sx, *sz, *isxe, *isze, *kap, *wid, *tcf;
    const float* eta;
    float k0sq, hh, shift;
    const float2* f;
    int64_t f_step;
    float2* u;
    int64_t u_step;
    const double* fscale;
    float2 *t, *fc;     // smoothed field between the two launches; coarse right-hand side
    const float2* uc;   // coarse correction
    const int* act;
};
struct LevelView {
    int nz, nx;
    int64_t n;
    const float2* cf;       // [n][6]: w/dg, dg, fw, fe, fn, fs
    float2 *f, *r;          // [nrhs][n] global work arrays
};
// One plan's LDS-resident tail (levels t..last), one workgroup per RHS
struct SegTail {
    const LevelView* views;
    int t, last, pre, post, coarse;
    const float2* f;
    float2* u;
    const int* act;
};

// segmented kernel launches (grid: sum of the segments' padded block counts)
// level: the grid level of every plan in the launch (level 0 forms its stencils from the medium)
void seg_mg_pre(const SegDims* dims, int nseg, int blocks, const SegLevel* lv, int j, int conj, int level,
                hipStream_t s);
void seg_mg_post(const SegDims* dims, int nseg, int blocks, const SegLevel* lv, int j, int conj, int level,
                 hipStream_t s);
void seg_mg_tail(const SegDims* dims, int nseg, int blocks, size_t lds, const SegTail* tv, int conj, hipStream_t s);
void seg_apply32(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, int conj, hipStream_t s);
void seg_dots(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, hipStream_t s);
void seg_reduce_update(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, double tol,
                       hipStream_t s);
void seg_norm_scale(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int j, int last, hipStream_t s);
void seg_xupdate(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, hipStream_t s);
void seg_residual(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int conj, hipStream_t s);
void seg_cycle_init(const SegDims* dims, int nseg, int blocks, const SegKrylov* k, int slot, double tol,
                    hipStream_t s);
// grids of the Krylov launches of one plan (x, y, z workgroups), as the
// single-plan solve launches them on its lazily scaled, fused path
struct KrylovGrids {
    int dots[3], update[3], norm[3], xupdate[3], residual[3], init[3], apply[3];
    int nch, nchd, nchb;
};
KrylovGrids krylov_grids(int nz, int nx, int nrhs);
// start of a solve: ||b|| per RHS, active flags, used = 0, x = 0 unless x0 is used
int32_t krylov_begin(helm_plan* p, const double2* b, double2* x, int nrhs, bool use_x0, hipStream_t s);
// plan p's Krylov table entry (workspace, grid, operator) for b, x and nrhs RHS
SegKrylov krylov_seg(helm_plan* p, const double2* b, double2* x, int nrhs);
// multigrid: work arrays for nrhs RHS; first tail level (-1: the V-cycle
// does not end in the LDS tail); the tail's dynamic LDS bytes
int32_t mg_prepare(helm_plan* p, int nrhs);
int mg_tail_level(const helm_plan* p);
size_t mg_tail_lds(const helm_plan* p, int t);
// level l of plan p in a batched V-cycle, with the grids of its pre- and
// post-smoothing launches; level 0 reads f + j f_step and writes u + j u_step;
// wg_target: workgroups one plan's launch aims for (sets the RHS per group)
SegLevel mg_seg_level(helm_plan* p, int l, int nrhs, const float2* f, int64_t f_step, float2* u, int64_t u_step,
                      const double* fscale, const int* act, int64_t wg_target, int pre_grid[3], int post_grid[3]);
SegTail mg_seg_tail(helm_plan* p, int t, const int* act);
bool mg_fused_ok(const helm_plan* p);  // pre = post = 2 on the fused level kernels

// HIP-event timing of the kernels inside a solve (helm_profile)
enum ProfKernel {
    PK_APPLY = 0, PK_PRECOND, PK_DOTS, PK_UPDATE, PK_RESIDUAL,
    PK_MG_PRE, PK_MG_POST,    // fused multigrid kernels, every level
    PK_MG_PRE0, PK_MG_POST0,  // the same kernels, finest level only
    PK_COUNT
};
struct Profiler {
    bool on = false;
    int mask = 0;  // bit k: time kernel k
    struct Rec {
        int id;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Rec> pending;
    int64_t launches[PK_COUNT] = {};
    double ms[PK_COUNT] = {};
    double bytes[PK_COUNT] = {};
};
hipEvent_t prof_begin(Profiler* pf, int id, hipStream_t s);  // nullptr when id is not timed
void prof_end(Profiler* pf, int id, hipEvent_t a, double bytes, hipStream_t s);
void prof_flush(Profiler* pf);  // after the stream has been synchronised
// the first `count` pending records only (work known to have completed);
// keep = false drops them instead of adding them to the totals
void prof_flush_first(Profiler* pf, size_t count, bool keep = true);
// event pair for a kernel launched with hipExtLaunchKernelGGL (the dispatch
// itself records the events, so no marker packets are added); false when id
// is not timed
bool prof_pair(Profiler* pf, int id, hipEvent_t* a, hipEvent_t* b);
void prof_push(Profiler* pf, int id, hipEvent_t a, hipEvent_t b, double bytes);

struct MGConfig {
    int levels = 0;
    double shift = 0.5;
    int pre = 1, post = 1, coarse = 10;
    double omega = 0.0;  // 0: per-node complex weights; > 0: this real weight everywhere
    int requested = 0;   // levels asked for by the last helm_mg_setup (the grid may allow fewer)
};

}  // namespace helm

struct helm_net;

struct helm_plan {
    helm_desc desc;
    int device = 0;
    double omega = 0, k0 = 0;
    bool medium_set = false;
    std::vector<helm::Level> levels;  // levels[0] = fine grid
    helm::MGConfig mg;
    struct helm_net* net = nullptr;  // learned preconditioner (not owned)
    float* net_scale = nullptr;      // per-RHS residual scale for the net
    double* net_part = nullptr;      // partial sums of that scale (4096 x 128)
    void* mg_views = nullptr;      // device copy of the per-level views used by k_tail
    bool mg_views_valid = false;
    helm::Workspace ws;
    helm::Profiler* prof = nullptr;   // kernel timing inside solves (helm_profile)
    uint64_t generation = 0;          // bumped whenever buffers a captured cycle uses are reallocated
    int64_t mg_wg_target = 0;         // multigrid workgroup target chosen by helm_autotune (0: default)
    helm_spectral_medium spectral{};  // geometry of the last helm_set_medium_spectral
    double2* sp_buf = nullptr;        // incident-field scratch: 4 * nx c128
    helm_conditioned_medium cond{};   // geometry of the last helm_set_medium_conditioned
    float2* cond_scr = nullptr;       // scaled preconditioner input (HELM_OP_COND8)
    int64_t cond_cap = 0;             // elements cond_scr holds
    double2* cond_grad = nullptr;     // material-gradient scratch: 3 * n c128
    // Fourier-multiplier preconditioner (helm_spectral_set_preconditioner)
    float2* fft_mult = nullptr;       // multiplier M [n]
    float2* fft_gain = nullptr;       // gain [n] (nullptr: 1)
    float2 *fft_a = nullptr, *fft_b = nullptr;  // transform buffers, fft_cap * n each
    int fft_cap = 0;
};

namespace helm {
// level construction (helm_plan.hip)
int32_t build_level_profiles(helm_plan* p, Level& L, double x0_mm, double z0_mm, double h_mm, int nz, int nx);
// stencil (helm_stencil.hip): the level's operator (5-point or, for STAG8, 8-tap)
int32_t launch_apply32(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                       const float2* b, float shift_im, hipStream_t s);
int32_t launch_apply64(const Level& L, double k0sq, const double2* u, double2* out, int nrhs, bool adjoint,
                       const double2* b, hipStream_t s);
// the 5-point flux operator of the level's eta/kap (the multigrid's operator)
int32_t launch_flux32(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                      const float2* b, float shift_im, hipStream_t s);
int32_t launch_flux64(const Level& L, double k0sq, const double2* u, double2* out, int nrhs, bool adjoint,
                      const double2* b, hipStream_t s);
// 8-tap periodic operator (helm_spectral.hip); mode 0: A u, 1: b - A u, 2: -(A - A0) u
int32_t launch_stag8_32(const Level& L, double k0sq, const float2* u, float2* out, int nrhs, bool adjoint,
                        const float2* b, hipStream_t s);
int32_t launch_stag8_64(const Level& L, double k0sq, const double2* u, double2* out, int nrhs, bool adjoint,
                        const double2* b, int mode, hipStream_t s);
// conditioned split-PML operator (helm_conditioned.hip): A u, or b - A u when b is given
int32_t launch_cond8_32(const Level& L, const float2* u, float2* out, int nrhs, bool adjoint, const float2* b,
                        hipStream_t s);
int32_t launch_cond8_64(const Level& L, const double2* u, double2* out, int nrhs, bool adjoint, const double2* b,
                        hipStream_t s);
// multigrid stretch profiles of level `level` from the split-PML factors
int32_t cond_level_profiles(const helm_plan* p, int level, std::vector<double2>& sx, std::vector<double2>& sz,
                            std::vector<double2>& isxe, std::vector<double2>& isze);
int32_t upload_cond_profiles(helm_plan* p, int level);
// the multigrid preconditioner of a conditioned plan (node scaling + V-cycle)
// fscale (nullable, per RHS): precondition fscale[r] * v[r]; with adjoint = true it
// needs mg_scaled_input_ok (the multigrid then applies it)
int32_t cond_precond(helm_plan* p, const float2* v, float2* z, int nrhs, bool adjoint, hipStream_t s,
                     const double* fscale = nullptr);
// multigrid (helm_mg.hip)
// fscale (nullable, per RHS): the V-cycle solves for fscale[r] * v[r] (applied as the
// finest level's kernels load v); only where mg_scaled_input_ok holds.
// active (nullable, per RHS, device): workgroups of the fused level kernels and
// the tail whose RHS are all 0 there return at once, leaving z of those RHS
// unspecified (for RHS whose Krylov cycle has stopped)
int32_t mg_vcycle(helm_plan* p, const float2* v, float2* z, int nrhs, bool adjoint, hipStream_t s,
                  const double* fscale = nullptr, const int* active = nullptr);
bool mg_scaled_input_ok(const helm_plan* p);  // the finest level runs the fused kernels
int32_t mg_build(helm_plan* p, hipStream_t s);
// krylov (helm_krylov.hip)
int32_t ensure_workspace(helm_plan* p, int nrhs, int restart);
int32_t solve_gmres(helm_plan* p, const double2* b, double2* x, int nrhs, const helm_solve_opts* o,
                    helm_solve_info* info, hipStream_t s);
int32_t solve_bicgstab(helm_plan* p, const double2* b, double2* x, int nrhs, const helm_solve_opts* o,
                       helm_solve_info* info, hipStream_t s);
int32_t precond_apply(helm_plan* p, int kind, const float2* v, float2* z, int nrhs, bool adjoint,
                      hipStream_t s);
// launch-shape tuning (helm_stencil.hip, helm_mg.hip)
int32_t tune_stencil(helm_plan* p, int nrhs, hipStream_t s, double* best_us);
int32_t tune_mg(helm_plan* p, int nrhs, hipStream_t s, double* best_us);
// Fourier-multiplier preconditioner z = gain ifft2(M fft2(v)) (helm_fft.hip);
// adjoint = false applies its conjugate map
// fscale (nullable, per RHS, device): input of RHS r scaled by fscale[r] (the
// lazily scaled Krylov basis), when spectral_scaled_input_ok
int32_t spectral_precond(helm_plan* p, const float2* v, float2* z, int nrhs, bool adjoint, hipStream_t s,
                         bool use_gain = true, const double* fscale = nullptr);
bool spectral_scaled_input_ok(const helm_plan* p);
// learned correction z += N(v) (helm_net.hip)
int32_t net_correction(helm_plan* p, struct helm_net* n, const float2* v, float2* z, int nrhs, hipStream_t s);
}  // namespace helm
