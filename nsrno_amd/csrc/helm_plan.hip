// Plan life cycle, stretch profiles, medium coefficients, library info.
//
// The stretch s(x) = 1 + i (sigma_max/omega) clip(frac, 0, 1)^2 with
// sigma_max = pml_strength * c0 / pml_width follows
// nsrno/theranostic/wave.py:127-135; node, edge-midpoint and exterior-ghost
// evaluation points follow wave.py:143-156 (the ghost edges sit at
// x_0 - h/2 and x_end + h/2).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "helm_internal.h"

namespace helm {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

static double2 stretch_at(double coord, double lower, double upper, double width, double omega, double c0,
                          double strength) {
    if (width == 0.0) return make_double2(1.0, 0.0);
    double frac = fmax((lower + width - coord) / width, (coord - upper + width) / width);
    double prof = fmin(fmax(frac, 0.0), 1.0);
    prof *= prof;
    double sigma_max = strength * c0 / (width * 1e-3);
    return make_double2(1.0, sigma_max / omega * prof);
}

static double2 crecip(double2 a) {
    double d = a.x * a.x + a.y * a.y;
    return make_double2(a.x / d, -a.y / d);
}

template <typename T>
static int32_t dev_alloc(T** ptr, size_t count) {
    if (*ptr) return HELM_OK;
    hipError_t e = hipMalloc((void**)ptr, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        *ptr = nullptr;
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return HELM_ERR_ALLOC;
    }
    return HELM_OK;
}

// Host-to-device copy that completes before returning, on a non-blocking
// stream of the current device (one per device, created on first use) rather
// than the legacy null stream: the runtime refuses a legacy-stream copy while
// another host thread captures a graph (a solve on another plan).
int32_t upload_h2d(void* dst, const void* src, size_t bytes) {
    static std::mutex mu;
    static hipStream_t streams[64] = {};
    int dev = 0;
    HELM_HIP(hipGetDevice(&dev));
    HELM_CHECK(dev >= 0 && dev < 64, HELM_ERR_STATE, "upload: device index out of range");
    hipStream_t st;
    {
        std::lock_guard<std::mutex> g(mu);
        if (!streams[dev]) HELM_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
        st = streams[dev];
    }
    HELM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    HELM_HIP(hipStreamSynchronize(st));
    return HELM_OK;
}

static int32_t upload_profile(const std::vector<double2>& h, double2** d64, float2** d32) {
    HELM_TRY(dev_alloc(d64, h.size()));
    HELM_TRY(dev_alloc(d32, h.size()));
    std::vector<float2> f(h.size());
    for (size_t i = 0; i < h.size(); ++i) f[i] = make_float2((float)h[i].x, (float)h[i].y);
    HELM_TRY(upload_h2d(*d64, h.data(), h.size() * sizeof(double2)));
    HELM_TRY(upload_h2d(*d32, f.data(), f.size() * sizeof(float2)));
    return HELM_OK;
}

int32_t build_level_profiles(helm_plan* p, Level& L, double x0_mm, double z0_mm, double h_mm, int nz, int nx) {
    const helm_desc& d = p->desc;
    const double lower_x = d.x0_mm, upper_x = d.x0_mm + (d.nx - 1) * d.h_mm;
    const double lower_z = d.z0_mm, upper_z = d.z0_mm + (d.nz - 1) * d.h_mm;
    L.nz = nz;
    L.nx = nx;
    L.n = (int64_t)nz * nx;
    L.h_m = h_mm * 1e-3;
    auto s = [&](double c, double lo, double hi) {
        return stretch_at(c, lo, hi, d.pml_mm, p->omega, d.c0_m_s, d.pml_strength);
    };
    std::vector<double2> sx(nx), sz(nz), isxe(nx + 1), isze(nz + 1);
    if (d.op_kind == HELM_OP_COND8 && p->cond.h > 0) {  // split-PML factors of the conditioned medium
        int level = 0;
        while (level < 30 && std::fabs(h_mm - d.h_mm * (double)(1 << level)) > 1e-9 * d.h_mm) ++level;
        HELM_TRY(cond_level_profiles(p, level, sx, sz, isxe, isze));
    } else {
        for (int j = 0; j < nx; ++j) sx[j] = s(x0_mm + j * h_mm, lower_x, upper_x);
        for (int i = 0; i < nz; ++i) sz[i] = s(z0_mm + i * h_mm, lower_z, upper_z);
        for (int k = 0; k <= nx; ++k) isxe[k] = crecip(s(x0_mm + (k - 0.5) * h_mm, lower_x, upper_x));
        for (int k = 0; k <= nz; ++k) isze[k] = crecip(s(z0_mm + (k - 0.5) * h_mm, lower_z, upper_z));
    }
    HELM_TRY(upload_profile(sx, &L.sx64, &L.sx32));
    HELM_TRY(upload_profile(sz, &L.sz64, &L.sz32));
    HELM_TRY(upload_profile(isxe, &L.isxe64, &L.isxe32));
    HELM_TRY(upload_profile(isze, &L.isze64, &L.isze32));
    HELM_TRY(dev_alloc(&L.eta64, L.n));
    HELM_TRY(dev_alloc(&L.kap64, L.n));
    HELM_TRY(dev_alloc(&L.eta32, L.n));
    HELM_TRY(dev_alloc(&L.kap32, L.n));
    return HELM_OK;
}

int32_t upload_cond_profiles(helm_plan* p, int level) {
    Level& L = p->levels[level];
    std::vector<double2> sx(L.nx), sz(L.nz), isxe(L.nx + 1), isze(L.nz + 1);
    HELM_TRY(cond_level_profiles(p, level, sx, sz, isxe, isze));
    HELM_TRY(upload_profile(sx, &L.sx64, &L.sx32));
    HELM_TRY(upload_profile(sz, &L.sz64, &L.sz32));
    HELM_TRY(upload_profile(isxe, &L.isxe64, &L.isxe32));
    HELM_TRY(upload_profile(isze, &L.isze64, &L.isze32));
    return HELM_OK;
}

static void free_level(Level& L) {
    void* ptrs[] = {L.sx64, L.sz64, L.isxe64, L.isze64, L.sx32, L.sz32, L.isxe32, L.isze32,
                    L.eta64, L.kap64, L.eta32, L.kap32, L.mg_u, L.mg_f, L.mg_r, L.mg_t, L.wgt, L.tcf,
                    L.wid, L.ax32, L.az32, L.ax64, L.az64, L.cmass64, L.cmass32, L.ctab64, L.ctab32, L.cmassh32};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    L = Level();
}

void free_workspace(Workspace& w) {
    if (w.gexec) (void)hipGraphExecDestroy(w.gexec);
    if (w.gstream) (void)hipStreamDestroy(w.gstream);
    void* ptrs[] = {w.V, w.Z, w.w, w.r64, w.partial, w.H, w.cs, w.sn, w.g, w.coef, w.beta, w.bnorm, w.rel,
                    w.inv, w.pscale, w.est, w.active, w.cyc, w.used, w.iters, w.bp, w.bv, w.bs, w.bt, w.bphat, w.bshat,
                    w.bshadow, w.br, w.bsc};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (w.hstat) (void)hipHostFree(w.hstat);
    for (hipEvent_t e : w.stat_ev)
        if (e) (void)hipEventDestroy(e);
    w = Workspace();
}

// ---- medium kernels -----------------------------------------------------
__global__ void k_medium_copy(int64_t n, const double* __restrict__ eta, const double2* __restrict__ kap,
                              double* eta64, double2* kap64, float* eta32, float2* kap32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double e = eta[i];
    double2 k = kap[i];
    eta64[i] = e;
    kap64[i] = k;
    eta32[i] = (float)e;
    kap32[i] = make_float2((float)k.x, (float)k.y);
}

__global__ void k_medium_physical(int64_t n, const double* __restrict__ c, const double* __restrict__ rho,
                                  const double* __restrict__ alpha, double rho0, double c0, double omega,
                                  double* eta64, double2* kap64, float* eta32, float2* kap32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ci = c[i], ri = rho[i];
    double e = rho0 / ri;
    double kr = rho0 * c0 * c0 / (ri * ci * ci);
    double2 att = make_double2(1.0, alpha ? alpha[i] * ci / omega : 0.0);
    double2 k = cscale(cmul(att, att), kr);
    eta64[i] = e;
    kap64[i] = k;
    eta32[i] = (float)e;
    kap32[i] = make_float2((float)k.x, (float)k.y);
}

// ---- kernel timing --------------------------------------------------------
hipEvent_t prof_begin(Profiler* pf, int id, hipStream_t s) {
    if (!(pf->mask >> id & 1)) return nullptr;
    hipEvent_t e = nullptr;
    if (!pf->pool.empty()) {
        e = pf->pool.back();
        pf->pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    (void)hipEventRecord(e, s);
    return e;
}

void prof_end(Profiler* pf, int id, hipEvent_t a, double bytes, hipStream_t s) {
    if (!a) return;
    hipEvent_t b = nullptr;
    if (!pf->pool.empty()) {
        b = pf->pool.back();
        pf->pool.pop_back();
    } else if (hipEventCreate(&b) != hipSuccess) {
        pf->pool.push_back(a);
        return;
    }
    (void)hipEventRecord(b, s);
    pf->pending.push_back({id, a, b, bytes});
}

static hipEvent_t prof_event(Profiler* pf) {
    hipEvent_t e = nullptr;
    if (!pf->pool.empty()) {
        e = pf->pool.back();
        pf->pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

bool prof_pair(Profiler* pf, int id, hipEvent_t* a, hipEvent_t* b) {
    *a = *b = nullptr;
    if (!pf || !(pf->mask >> id & 1)) return false;
    *a = prof_event(pf);
    *b = prof_event(pf);
    if (!*a || !*b) {
        if (*a) pf->pool.push_back(*a);
        if (*b) pf->pool.push_back(*b);
        *a = *b = nullptr;
        return false;
    }
    return true;
}

void prof_push(Profiler* pf, int id, hipEvent_t a, hipEvent_t b, double bytes) {
    if (a && b) pf->pending.push_back({id, a, b, bytes});
}

void prof_flush_first(Profiler* pf, size_t count, bool keep) {
    count = std::min(count, pf->pending.size());
    for (size_t i = 0; i < count; ++i) {
        auto& r = pf->pending[i];
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && keep && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            pf->launches[r.id]++;
            pf->ms[r.id] += ms;
            pf->bytes[r.id] += r.bytes;
        }
        pf->pool.push_back(r.a);
        pf->pool.push_back(r.b);
    }
    pf->pending.erase(pf->pending.begin(), pf->pending.begin() + count);
}

void prof_flush(Profiler* pf) { prof_flush_first(pf, pf->pending.size()); }

static void prof_destroy(Profiler* pf) {
    if (!pf) return;
    prof_flush(pf);
    for (hipEvent_t e : pf->pool) (void)hipEventDestroy(e);
    delete pf;
}

}  // namespace helm

using namespace helm;

extern "C" {

int32_t helm_profile(helm_plan* p, int32_t mask) {
    HELM_CHECK(p, HELM_ERR_ARG, "helm_profile: null plan");
    if (!p->prof) p->prof = new Profiler();
    Profiler* pf = p->prof;
    prof_flush(pf);
    pf->on = mask != 0;
    pf->mask = mask;
    if (pf->on) {
        for (int k = 0; k < PK_COUNT; ++k) {
            pf->launches[k] = 0;
            pf->ms[k] = 0.0;
            pf->bytes[k] = 0.0;
        }
    }
    return HELM_OK;
}

int32_t helm_profile_read(helm_plan* p, helm_kernel_stat* stats, int32_t capacity, int32_t* count) {
    HELM_CHECK(p && count && (stats || capacity == 0), HELM_ERR_ARG, "helm_profile_read: null argument");
    static const char* names[PK_COUNT] = {"operator_apply", "preconditioner", "dcgs_dots", "dcgs_update",
                                          "true_residual", "mg_pre2_restrict", "mg_prolong_post2",
                                          "mg_pre2_restrict_finest", "mg_prolong_post2_finest"};
    *count = PK_COUNT;
    Profiler* pf = p->prof;
    for (int k = 0; k < PK_COUNT && k < capacity; ++k) {
        std::memset(&stats[k], 0, sizeof(helm_kernel_stat));
        std::strncpy(stats[k].name, names[k], sizeof(stats[k].name) - 1);
        if (pf) {
            stats[k].launches = pf->launches[k];
            stats[k].total_ms = pf->ms[k];
            stats[k].bytes = pf->bytes[k];
        }
    }
    return HELM_OK;
}

const char* helm_version(void) { return "helm 0.1 (gfx950)"; }

const char* helm_error_string(int32_t code) {
    switch (code) {
        case HELM_OK: return "success";
        case HELM_ERR_ARG: return "invalid argument";
        case HELM_ERR_HIP: return "HIP runtime error";
        case HELM_ERR_ALLOC: return "device allocation failed";
        case HELM_ERR_STATE: return "invalid call order";
        default: return "unknown error";
    }
}

const char* helm_last_error(void) { return g_last_error.c_str(); }

int32_t helm_plan_create(const helm_desc* desc, helm_plan** out) {
    HELM_CHECK(desc && out, HELM_ERR_ARG, "helm_plan_create: null argument");
    *out = nullptr;
    const helm_desc& d = *desc;
    HELM_CHECK(d.nz >= 3 && d.nx >= 3, HELM_ERR_ARG, "grid needs >= 3 nodes per axis");
    HELM_CHECK(d.nrhs >= 1 && d.nrhs <= 4096, HELM_ERR_ARG, "nrhs must be in [1, 4096]");
    HELM_CHECK(d.restart >= 0 && d.restart <= 256, HELM_ERR_ARG, "restart must be in [0, 256]");
    HELM_CHECK(std::isfinite(d.h_mm) && d.h_mm > 0, HELM_ERR_ARG, "h_mm must be finite and positive");
    HELM_CHECK(std::isfinite(d.frequency_mhz) && d.frequency_mhz > 0, HELM_ERR_ARG,
               "frequency_mhz must be finite and positive");
    HELM_CHECK(std::isfinite(d.c0_m_s) && d.c0_m_s > 0, HELM_ERR_ARG, "c0_m_s must be finite and positive");
    HELM_CHECK(std::isfinite(d.pml_strength) && d.pml_strength > 0, HELM_ERR_ARG,
               "pml_strength must be finite and positive");
    HELM_CHECK(std::isfinite(d.pml_mm) && d.pml_mm >= 0, HELM_ERR_ARG, "pml_mm must be finite and nonnegative");
    HELM_CHECK(2 * d.pml_mm < fmin((d.nx - 1) * d.h_mm, (d.nz - 1) * d.h_mm), HELM_ERR_ARG,
               "PML leaves no physical domain");
    HELM_CHECK((int64_t)d.nz * d.nx * d.nrhs < (int64_t)1 << 31, HELM_ERR_ARG,
               "nz*nx*nrhs must stay below 2^31");
    HELM_CHECK(d.op_kind == HELM_OP_FLUX5 || d.op_kind == HELM_OP_STAG8 || d.op_kind == HELM_OP_COND8,
               HELM_ERR_ARG, "unknown op_kind");
    HELM_CHECK(d.op_kind != HELM_OP_STAG8 || (d.pml_mm == 0.0 && d.nz >= 16 && d.nx >= 16), HELM_ERR_ARG,
               "op_kind HELM_OP_STAG8 needs pml_mm = 0 (the sponge absorbs) and a box of at least 16 x 16");
    HELM_CHECK(d.op_kind != HELM_OP_COND8 || (d.pml_mm == 0.0 && d.nz >= 24 && d.nx >= 24), HELM_ERR_ARG,
               "op_kind HELM_OP_COND8 needs pml_mm = 0 (the split PML lies inside the medium) and a box of at "
               "least 24 x 24");
    helm_plan* p = new helm_plan();
    p->desc = d;
    (void)hipGetDevice(&p->device);
    p->omega = 2.0 * M_PI * d.frequency_mhz * 1e6;
    p->k0 = p->omega / d.c0_m_s;
    p->levels.resize(1);
    int32_t st = build_level_profiles(p, p->levels[0], d.x0_mm, d.z0_mm, d.h_mm, d.nz, d.nx);
    if (st != HELM_OK) {
        helm_plan_destroy(p);
        return st;
    }
    *out = p;
    return HELM_OK;
}

int32_t helm_plan_destroy(helm_plan* p) {
    if (!p) return HELM_OK;
    (void)hipDeviceSynchronize();
    for (auto& L : p->levels) free_level(L);
    free_workspace(p->ws);
    if (p->mg_views) (void)hipFree(p->mg_views);
    if (p->net_scale) (void)hipFree(p->net_scale);
    if (p->net_part) (void)hipFree(p->net_part);
    if (p->sp_buf) (void)hipFree(p->sp_buf);
    if (p->cond_scr) (void)hipFree(p->cond_scr);
    if (p->cond_grad) (void)hipFree(p->cond_grad);
    for (void* q : {(void*)p->fft_mult, (void*)p->fft_gain, (void*)p->fft_a, (void*)p->fft_b})
        if (q) (void)hipFree(q);
    prof_destroy(p->prof);
    delete p;
    return HELM_OK;
}

int32_t helm_set_medium(helm_plan* p, const double* eta_rel, const double* kappa_c, void* stream) {
    HELM_CHECK(p && eta_rel && kappa_c, HELM_ERR_ARG, "helm_set_medium: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_FLUX5, HELM_ERR_STATE,
               "helm_set_medium: spectral and conditioned plans take their own medium call");
    hipStream_t s = (hipStream_t)stream;
    Level& L = p->levels[0];
    int64_t n = L.n;
    hipLaunchKernelGGL(k_medium_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, eta_rel,
                       (const double2*)kappa_c, L.eta64, L.kap64, L.eta32, L.kap32);
    HELM_HIP(hipGetLastError());
    p->medium_set = true;
    if (p->mg.levels > 0) HELM_TRY(mg_build(p, s));
    return HELM_OK;
}

int32_t helm_set_medium_physical(helm_plan* p, const double* c, const double* rho, const double* alpha,
                                 double rho0, void* stream) {
    HELM_CHECK(p && c && rho, HELM_ERR_ARG, "helm_set_medium_physical: null argument");
    HELM_CHECK(p->desc.op_kind == HELM_OP_FLUX5, HELM_ERR_STATE,
               "helm_set_medium_physical: spectral and conditioned plans take their own medium call");
    HELM_CHECK(std::isfinite(rho0) && rho0 > 0, HELM_ERR_ARG, "rho0 must be finite and positive");
    hipStream_t s = (hipStream_t)stream;
    Level& L = p->levels[0];
    int64_t n = L.n;
    hipLaunchKernelGGL(k_medium_physical, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, c, rho, alpha,
                       rho0, p->desc.c0_m_s, p->omega, L.eta64, L.kap64, L.eta32, L.kap32);
    HELM_HIP(hipGetLastError());
    p->medium_set = true;
    if (p->mg.levels > 0) HELM_TRY(mg_build(p, s));
    return HELM_OK;
}

}  // extern "C"
