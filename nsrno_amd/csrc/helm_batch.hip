// Batched solve over plans of different grid sizes (helm_batch_solve).
//
// The reference solves one frequency at a time: reconstruct_dense.py:91
// hands frequencies to processes, and solve_system (dense_wave.py:167-225)
// batches the angles of one frequency.  Here every plan is one frequency
// with its angles as right-hand sides, and all plans advance together: each
// kernel of a restart cycle is ONE segmented launch over every plan still
// running (multigrid level l over the plans that have a level l above their
// tail, the LDS tail over all, the FP32 stencil, DCGS2, the x update, the
// FP64 restart residual), so a cycle is the same short launch sequence
// whatever the number of plans, and the launches of small grids and coarse
// levels fill the device together instead of one plan at a time.  The cycle
// is captured once as a HIP graph and replayed; the restart status of every
// plan lands in pinned host memory behind one event per cycle.  When a plan
// has converged its entries leave the tables (rebuilt between cycles, after
// the stream has drained) and the graph is captured again.
//
// Each plan keeps the numerics of helm_solve's multigrid FGMRES on its
// lazily scaled, fused path: the same kernel bodies on the same arrays, in
// the same order per plan, so that a plan solved here and by helm_solve
// (without autotuning, default launch shapes) gives the same fields bit for
// bit.  The stencil takes helm_solve's default shape on every plan (block
// form, 4 RHS per wavefront, 4-row runs): one launch serves every plan, so
// there is no per-plan autotuned shape.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "helm_internal.h"

namespace helm {
namespace {

// one segmented launch: dims / table offsets into the device block, then pointers
struct Launch {
    size_t dims_off = 0, tab_off = 0;
    const SegDims* dims = nullptr;
    const void* tab = nullptr;
    int nseg = 0, blocks = 0;
};

struct Packer {
    std::vector<char> buf;
    size_t add(const void* src, size_t bytes) {
        const size_t o = (buf.size() + 255) & ~(size_t)255;
        buf.resize(o + bytes);
        std::memcpy(buf.data() + o, src, bytes);
        return o;
    }
};

struct DimsBuilder {
    std::vector<SegDims> d;
    int cur = 0;
    void add(const int g[3]) {
        d.push_back(SegDims{cur, g[0], g[1], g[2]});
        cur += ((g[0] * g[1] * g[2] + 7) / 8) * 8;  // starts stay multiples of 8 (XCD of the local id)
    }
};

struct PlanState {
    helm_plan* p = nullptr;
    const double2* b = nullptr;
    double2* x = nullptr;
    int tail = 0;  // first tail level
    bool running = true;
    int cycles = 0;
    std::vector<double> rel;
    std::vector<int> act, iters;
};

struct Batch {
    std::vector<PlanState> plans;
    int nrhs = 0, m = 0, conj = 0, max_levels = 0;
    double tol = 0;
    // device tables of the running plans
    void* dev = nullptr;
    size_t dev_cap = 0;
    Launch dots, update, norm, xupdate, residual, init, apply, tail;
    std::vector<Launch> pre, post;  // per level
    size_t tail_lds = 0;
    // captured cycle
    hipGraphExec_t gexec = nullptr;
    hipStream_t gstream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int graph_fallbacks = 0;

    hipStream_t stream = nullptr;  // the caller's: drained before the tables are freed
    ~Batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (gstream) (void)hipStreamDestroy(gstream);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (dev) (void)hipFree(dev);
    }

    // rebuild every table for the plans still running (the stream must be idle)
    int32_t build() {
        Packer pk;
        std::vector<SegKrylov> ks;
        DimsBuilder dd, du, dn, dx, dr, di, da, dt;
        std::vector<std::vector<SegLevel>> lv(max_levels);
        std::vector<DimsBuilder> dpre(max_levels), dpost(max_levels);
        std::vector<SegTail> tv;
        tail_lds = 0;
        // workgroups per plan and grid-level launch: a single plan's 512,
        // shared among the plans running, so that with many plans each
        // workgroup takes every RHS of its tile (the tile's stencil is read
        // once, not once per RHS group) and the launch is filled by plans
        int running = 0;
        for (const PlanState& ps : plans) running += ps.running;
        const int64_t wg_target = std::max<int64_t>(1, 512 / std::max(1, running));
        for (PlanState& ps : plans) {
            if (!ps.running) continue;
            helm_plan* p = ps.p;
            const Level& L = p->levels[0];
            SegKrylov k = krylov_seg(p, ps.b, ps.x, nrhs);
            const KrylovGrids g = krylov_grids(L.nz, L.nx, nrhs);
            ks.push_back(k);
            dd.add(g.dots);
            du.add(g.update);
            dn.add(g.norm);
            dx.add(g.xupdate);
            dr.add(g.residual);
            di.add(g.init);
            da.add(g.apply);
            const int64_t vs = (int64_t)nrhs * L.n;
            for (int l = 0; l < ps.tail; ++l) {
                int gp[3], gq[3];
                const bool fine = l == 0;
                lv[l].push_back(mg_seg_level(p, l, nrhs, fine ? k.V : p->levels[l].mg_f, fine ? vs : 0,
                                             fine ? k.Z : p->levels[l].mg_u, fine ? vs : 0,
                                             fine ? k.pscale : nullptr, k.cyc, wg_target, gp, gq));
                dpre[l].add(gp);
                dpost[l].add(gq);
            }
            tv.push_back(mg_seg_tail(p, ps.tail, k.cyc));
            const int gt[3] = {nrhs, 1, 1};
            dt.add(gt);
            tail_lds = std::max(tail_lds, mg_tail_lds(p, ps.tail));
        }
        auto put = [&](Launch& ln, const DimsBuilder& d, const void* tab, size_t bytes) {
            ln.nseg = (int)d.d.size();
            ln.blocks = d.cur;
            ln.dims_off = ln.nseg ? pk.add(d.d.data(), d.d.size() * sizeof(SegDims)) : 0;
            ln.tab_off = ln.nseg ? pk.add(tab, bytes) : 0;
        };
        const size_t kb = ks.size() * sizeof(SegKrylov);
        put(dots, dd, ks.data(), kb);
        // the other Krylov launches share the first copy of the table
        auto share = [&](Launch& ln, const DimsBuilder& d) {
            ln.nseg = (int)d.d.size();
            ln.blocks = d.cur;
            ln.dims_off = ln.nseg ? pk.add(d.d.data(), d.d.size() * sizeof(SegDims)) : 0;
            ln.tab_off = dots.tab_off;
        };
        share(update, du);
        share(norm, dn);
        share(xupdate, dx);
        share(residual, dr);
        share(init, di);
        share(apply, da);
        put(tail, dt, tv.data(), tv.size() * sizeof(SegTail));
        pre.assign(max_levels, Launch{});
        post.assign(max_levels, Launch{});
        for (int l = 0; l < max_levels; ++l) {
            put(pre[l], dpre[l], lv[l].data(), lv[l].size() * sizeof(SegLevel));
            share_level(pk, post[l], dpost[l], pre[l].tab_off);
        }
        if (pk.buf.size() > dev_cap) {
            if (dev) HELM_HIP(hipFree(dev));
            dev = nullptr;
            dev_cap = 0;
            HELM_HIP(hipMalloc(&dev, pk.buf.size()));
            dev_cap = pk.buf.size();
        }
        if (!pk.buf.empty()) HELM_TRY(upload_h2d(dev, pk.buf.data(), pk.buf.size()));
        auto fix = [&](Launch& ln) {
            ln.dims = (const SegDims*)((char*)dev + ln.dims_off);
            ln.tab = (const void*)((char*)dev + ln.tab_off);
        };
        for (Launch* ln : {&dots, &update, &norm, &xupdate, &residual, &init, &apply, &tail}) fix(*ln);
        for (int l = 0; l < max_levels; ++l) {
            fix(pre[l]);
            fix(post[l]);
        }
        if (gexec) {
            (void)hipGraphExecDestroy(gexec);
            gexec = nullptr;
        }
        return HELM_OK;
    }
    static void share_level(Packer& pk, Launch& ln, const DimsBuilder& d, size_t tab_off) {
        ln.nseg = (int)d.d.size();
        ln.blocks = d.cur;
        ln.dims_off = ln.nseg ? pk.add(d.d.data(), d.d.size() * sizeof(SegDims)) : 0;
        ln.tab_off = tab_off;
    }

    const SegKrylov* ktab() const { return (const SegKrylov*)dots.tab; }

    // the restart residual of every running plan into V_0 and its status into slot `slot`
    int32_t status(int slot, hipStream_t s) {
        if (residual.nseg) {
            seg_residual(residual.dims, residual.nseg, residual.blocks, ktab(), conj, s);
            seg_cycle_init(init.dims, init.nseg, init.blocks, ktab(), slot, tol, s);
        }
        HELM_HIP(hipGetLastError());
        HELM_HIP(hipEventRecord(ev[slot], s));
        return HELM_OK;
    }

    // z = M v on the finest level of every running plan, V-cycle as in vcycle()
    void vcycle(int j, hipStream_t s) {
        for (int l = 0; l < max_levels; ++l)
            if (pre[l].nseg)
                seg_mg_pre(pre[l].dims, pre[l].nseg, pre[l].blocks, (const SegLevel*)pre[l].tab, j, conj, l, s);
        seg_mg_tail(tail.dims, tail.nseg, tail.blocks, tail_lds, (const SegTail*)tail.tab, conj, s);
        for (int l = max_levels - 1; l >= 0; --l)
            if (post[l].nseg)
                seg_mg_post(post[l].dims, post[l].nseg, post[l].blocks, (const SegLevel*)post[l].tab, j, conj, l, s);
    }

    int32_t cycle(int nsteps, hipStream_t s) {
        if (!dots.nseg) return HELM_OK;
        for (int j = 0; j < nsteps; ++j) {
            vcycle(j, s);
            seg_apply32(apply.dims, apply.nseg, apply.blocks, ktab(), j, conj, s);
            seg_dots(dots.dims, dots.nseg, dots.blocks, ktab(), j, s);
            seg_reduce_update(update.dims, update.nseg, update.blocks, ktab(), j, tol, s);
            seg_norm_scale(norm.dims, norm.nseg, norm.blocks, ktab(), j, j + 1 >= nsteps ? 1 : 0, s);
        }
        seg_xupdate(xupdate.dims, xupdate.nseg, xupdate.blocks, ktab(), s);
        HELM_HIP(hipGetLastError());
        return HELM_OK;
    }

    // a full cycle: replayed from a captured graph once the tables are stable
    int32_t run_cycle(int nsteps, bool graph, hipStream_t s) {
        if (!graph || nsteps != m) return cycle(nsteps, s);
        if (!gexec) {
            if (!gstream) HELM_HIP(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
            bool ok = hipStreamBeginCapture(gstream, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                const int32_t st = cycle(m, gstream);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(gstream, &g);
                ok = st == HELM_OK && e == hipSuccess && g &&
                     hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0) == hipSuccess;
                if (g) (void)hipGraphDestroy(g);
            }
            if (!ok) {
                if (gexec) (void)hipGraphExecDestroy(gexec);
                gexec = nullptr;
                (void)hipGetLastError();
                graph_fallbacks++;
                return cycle(nsteps, s);
            }
        }
        HELM_HIP(hipGraphLaunch(gexec, s));
        return HELM_OK;
    }
};

int32_t batch_solve(helm_plan* const* plans, int np, const double2* const* b, double2* const* x, int nrhs,
                    const helm_solve_opts* o, helm_solve_info* info, hipStream_t s) {
    Batch B;
    B.stream = s;
    B.nrhs = nrhs;
    B.m = o->restart;
    B.conj = o->adjoint != 0;
    B.tol = o->tolerance;
    B.plans.resize(np);
    for (int i = 0; i < np; ++i) {
        helm_plan* p = plans[i];
        PlanState& ps = B.plans[i];
        ps.p = p;
        ps.b = b[i];
        ps.x = x[i];
        HELM_TRY(ensure_workspace(p, nrhs, B.m));
        HELM_TRY(mg_prepare(p, nrhs));
        ps.tail = mg_tail_level(p);
        HELM_CHECK(ps.tail >= 1 && mg_scaled_input_ok(p), HELM_ERR_ARG,
                   "helm_batch_solve: every plan needs a V-cycle of fused grid levels ending in the LDS tail "
                   "(pre = post = 2, a finest level above the tail size)");
        const KrylovGrids g = krylov_grids(p->levels[0].nz, p->levels[0].nx, nrhs);
        HELM_CHECK((int64_t)nrhs * g.nchb <= p->ws.partial_len, HELM_ERR_ARG,
                   "helm_batch_solve: plan grid too narrow for the block-form restart residual");
        B.max_levels = std::max(B.max_levels, ps.tail);
        ps.rel.assign(nrhs, 0.0);
        ps.act.assign(nrhs, 1);
        ps.iters.assign(nrhs, 0);
    }
    for (hipEvent_t& e : B.ev) HELM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (PlanState& ps : B.plans) HELM_TRY(krylov_begin(ps.p, ps.b, ps.x, nrhs, o->use_x0 != 0, s));
    HELM_TRY(B.build());

    // restart loop of solve_gmres, over all running plans at once: the next
    // cycle and its status are queued before the host reads the current one
    int steps = 0, slot = 0;
    const bool graph = o->use_graph != 0;
    HELM_TRY(B.status(slot, s));
    for (;;) {
        const int msteps = std::min(B.m, o->max_iterations - steps);
        const bool queued = msteps > 0;  // no cycle beyond the iteration budget
        if (queued) {
            HELM_TRY(B.run_cycle(msteps, graph && steps > 0, s));
            HELM_TRY(B.status(slot ^ 1, s));
        }
        HELM_HIP(hipEventSynchronize(B.ev[slot]));
        bool changed = false;
        int running = 0;
        for (PlanState& ps : B.plans) {
            if (!ps.running) continue;
            const Workspace& W = ps.p->ws;
            const char* hs = (const char*)W.hstat + (size_t)slot * W.stat_bytes;
            std::memcpy(ps.rel.data(), hs, nrhs * sizeof(double));
            std::memcpy(ps.act.data(), hs + (size_t)W.nrhs * sizeof(double), nrhs * sizeof(int));
            std::memcpy(ps.iters.data(), hs + (size_t)W.nrhs * (sizeof(double) + sizeof(int)), nrhs * sizeof(int));
            bool any = false;
            for (int r = 0; r < nrhs; ++r) any = any || ps.act[r] != 0;
            if (!any) {
                ps.running = false;
                changed = true;
            } else {
                running++;
            }
        }
        if (running == 0 || steps >= o->max_iterations) {
            if (queued) HELM_HIP(hipStreamSynchronize(s));
            break;
        }
        for (PlanState& ps : B.plans)
            if (ps.running) ps.cycles++;
        if (changed) {
            // the cycle queued ahead still reads the old tables: drain, then rebuild
            HELM_HIP(hipStreamSynchronize(s));
            HELM_TRY(B.build());
        }
        steps += msteps;
        slot ^= 1;
    }
    for (int i = 0; i < np; ++i) {
        const PlanState& ps = B.plans[i];
        for (int r = 0; r < nrhs; ++r) {
            helm_solve_info& f = info[(int64_t)i * nrhs + r];
            f.reason = ps.act[r] ? HELM_ITERATION_BUDGET : HELM_CONVERGED;
            if (!std::isfinite(ps.rel[r])) f.reason = HELM_BREAKDOWN;
            f.iterations = ps.iters[r];
            f.applications = ps.iters[r] + ps.cycles + 1;
            f.cycles = ps.cycles;
            f.relative_residual = ps.rel[r];
            f.fallback = 0;
            f.first_reason = f.reason;
            f.graph_fallbacks = B.graph_fallbacks;
        }
    }
    return HELM_OK;
}

}  // namespace
}  // namespace helm

using namespace helm;

extern "C" int32_t helm_batch_solve(int32_t nplans, helm_plan* const* plans, const void* const* rhs,
                                    void* const* x, int32_t nrhs, const helm_solve_opts* o, helm_solve_info* info,
                                    void* stream) {
    HELM_CHECK(plans && rhs && x && o && info, HELM_ERR_ARG, "helm_batch_solve: null argument");
    HELM_CHECK(nplans >= 1 && nplans <= SEG_MAX, HELM_ERR_ARG, "helm_batch_solve: 1 to 256 plans");
    HELM_CHECK(nrhs >= 1 && nrhs <= 64, HELM_ERR_ARG, "helm_batch_solve: nrhs must be in [1, 64]");
    HELM_CHECK(o->method == HELM_GMRES && o->precond == HELM_PC_MG, HELM_ERR_ARG,
               "helm_batch_solve: multigrid-preconditioned GMRES only");
    HELM_CHECK(o->restart >= 1 && o->restart <= 63, HELM_ERR_ARG, "helm_batch_solve: restart must be in [1, 63]");
    HELM_CHECK(o->max_iterations >= 1, HELM_ERR_ARG, "helm_batch_solve: max_iterations must be positive");
    HELM_CHECK(std::isfinite(o->tolerance) && o->tolerance > 0, HELM_ERR_ARG,
               "helm_batch_solve: tolerance must be finite and positive");
    for (int i = 0; i < nplans; ++i) {
        helm_plan* p = plans[i];
        HELM_CHECK(p && rhs[i] && x[i], HELM_ERR_ARG, "helm_batch_solve: null plan or vector");
        HELM_CHECK(p->medium_set, HELM_ERR_STATE, "helm_batch_solve: call helm_set_medium on every plan first");
        HELM_CHECK(p->levels[0].op == HELM_OP_FLUX5, HELM_ERR_ARG, "helm_batch_solve: flux-form plans only");
        HELM_CHECK(p->mg.levels > 0, HELM_ERR_STATE, "helm_batch_solve: call helm_mg_setup on every plan first");
        HELM_CHECK((int64_t)nrhs * p->levels[0].n < ((int64_t)1 << 31), HELM_ERR_ARG,
                   "helm_batch_solve: nrhs x nodes out of range");
        for (int k = 0; k < i; ++k)
            HELM_CHECK(plans[k] != p, HELM_ERR_ARG, "helm_batch_solve: a plan appears twice");
    }
    return batch_solve(plans, nplans, (const double2* const*)rhs, (double2* const*)x, nrhs, o, info,
                       (hipStream_t)stream);
}
