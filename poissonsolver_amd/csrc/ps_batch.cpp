// Batches of independent solves of uneven size (C ABI in include/ps_capi.h).
//
// Each problem is a full solver context with its own plans, tile shapes and
// Green's kernel.  A batch solve is captured once into a hipGraph: an origin
// stream forks to every problem's stream, each stream records its five pass
// launches, and all branches join back; replaying the executable graph is one
// hipGraphLaunch for the whole batch.  Problems of one grid (cells, bc, dx,
// reg) form a group whose sources and solutions are consecutive slices of one
// device block: when their options and pushed components agree and the
// operator is plain, the group is one solve whose components are the members'
// fields (five launches for the group instead of five per member), which fills
// the GPU where a small problem alone would not.  The graph is re-captured when the set
// of pushed components or the operator options change (the captured launch
// arguments depend on both); pushing new data of the same components reuses
// the graph, which reads the contexts' fixed device buffers.
#include <hip/hip_runtime.h>

#include <chrono>
#include <new>
#include <string>
#include <vector>

#include "../../include/ps_capi.h"

namespace psamd {
int ctx_mark_solved(ps_ctx* c);      // ps_capi.cpp
int ctx_rhs_mask(const ps_ctx* c);
int ctx_adopt_real(ps_ctx* c, void* rhs, void* lhs);
int ctx_solve_group(ps_ctx* const* m, int K);
void ctx_reset_held(ps_ctx* c);
unsigned long ctx_signature(const ps_ctx* c);
int ctx_check_solvable(ps_ctx* c);
}  // namespace psamd

struct ps_batch {
    std::vector<ps_ctx*> ctx;
    std::vector<int> ncell, bc, reg;    // per problem: 3, 3, 1 entries
    std::vector<double> dx;             // per problem: 3 entries
    hipStream_t origin = nullptr;
    std::vector<hipEvent_t> fork, join;
    std::vector<std::vector<int>> groups;   // problems of one grid, two or more
    std::vector<void*> blocks;              // per group: rhs block, lhs block
    int elem_bytes = 8;
    int ngrouped = 0;                       // problems solved inside group solves, last enqueue
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;          // graph must be (re)captured
    bool warmed = false;        // one eager solve done (buffers allocated)
    std::vector<unsigned long> sig;   // members' signatures when the graph was captured
    std::string err;
};

namespace {

int bfail(ps_batch* b, int st, const std::string& m)
{
    if (b) b->err = m;
    return st;
}

#define BK(call)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return bfail(b, PS_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void drop_graph(ps_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->dirty = true;
}

int check(ps_batch* b, int prob)
{
    if (!b) return PS_EINVAL;
    if (prob < 0 || prob >= (int)b->ctx.size()) return bfail(b, PS_EINVAL, "problem index out of range");
    return PS_OK;
}

int sub(ps_batch* b, ps_ctx* c, int st)
{
    if (st != PS_OK) b->err = ps_last_error(c);
    return st;
}

// a push that adds components changes the captured launches
void note_push(ps_batch* b, int prob, int mask_before)
{
    if (psamd::ctx_rhs_mask(b->ctx[prob]) != mask_before) {
        b->dirty = true;
        b->warmed = false;
    }
}

// Problems of one grid share a device block for their sources and solutions:
// member j's rhs and lhs are slices j of 3 n3 reals.  Called after every setup
// of the members (a setup returns a context to buffers of its own).
int form_groups(ps_batch* b)
{
    for (void* p : b->blocks)
        if (p) (void)hipFree(p);
    b->blocks.clear();
    b->groups.clear();
    const int n = (int)b->ctx.size();
    std::vector<bool> taken(n, false);
    for (int i = 0; i < n; ++i) {
        if (taken[i]) continue;
        std::vector<int> g{i};
        for (int j = i + 1; j < n; ++j) {
            bool same = b->reg[j] == b->reg[i];
            for (int d = 0; d < 3; ++d)
                same = same && b->ncell[3 * j + d] == b->ncell[3 * i + d] && b->bc[3 * j + d] == b->bc[3 * i + d] &&
                       b->dx[3 * j + d] == b->dx[3 * i + d];
            if (same) g.push_back(j);
        }
        if (g.size() < 2) continue;
        for (int j : g) taken[j] = true;
        const size_t slice = 3 * (size_t)b->ncell[3 * i] * b->ncell[3 * i + 1] * b->ncell[3 * i + 2] * b->elem_bytes;
        void* rb = nullptr;
        void* lb = nullptr;
        BK(hipMalloc(&rb, slice * g.size()));
        b->blocks.push_back(rb);
        BK(hipMalloc(&lb, slice * g.size()));
        b->blocks.push_back(lb);
        for (size_t k = 0; k < g.size(); ++k) {
            int st = sub(b, b->ctx[g[k]], psamd::ctx_adopt_real(b->ctx[g[k]], static_cast<char*>(rb) + k * slice,
                                                               static_cast<char*>(lb) + k * slice));
            if (st) return st;
        }
        b->groups.push_back(g);
    }
    return PS_OK;
}

bool grouped(const ps_batch* b, int prob)
{
    for (const auto& g : b->groups)
        for (int j : g)
            if (j == prob) return true;
    return false;
}

// Enqueue the batch on the problems' streams (forked from and joined to the
// origin stream): each group as one solve where it qualifies, every other
// problem (and the members of a group that does not) on its own.
int enqueue(ps_batch* b)
{
    int st = PS_OK;
    auto fork_to = [&](int i) -> int {
        hipStream_t s = (hipStream_t)ps_stream(b->ctx[i]);
        if (hipEventRecord(b->fork[i], b->origin) != hipSuccess || hipStreamWaitEvent(s, b->fork[i], 0) != hipSuccess)
            return bfail(b, PS_EHIP, "batch: fork failed");
        return PS_OK;
    };
    auto join_from = [&](int i) -> int {
        hipStream_t s = (hipStream_t)ps_stream(b->ctx[i]);
        if (hipEventRecord(b->join[i], s) != hipSuccess || hipStreamWaitEvent(b->origin, b->join[i], 0) != hipSuccess)
            return bfail(b, PS_EHIP, "batch: join failed");
        return PS_OK;
    };
    auto one = [&](int i) -> int {
        int r = fork_to(i);
        if (r == PS_OK) r = sub(b, b->ctx[i], ps_solve3d_async(b->ctx[i]));
        if (r == PS_OK) r = join_from(i);
        return r;
    };
    b->ngrouped = 0;
    for (const auto& g : b->groups) {
        std::vector<ps_ctx*> m;
        for (int j : g) m.push_back(b->ctx[j]);
        const int lead = g[0];
        if ((st = fork_to(lead))) return st;
        st = psamd::ctx_solve_group(m.data(), (int)m.size());
        if (st == PS_OK) {
            if ((st = join_from(lead))) return st;
            b->ngrouped += (int)g.size();
            continue;
        }
        if (st != PS_EUNSUPPORTED) return sub(b, m[0], st);
        if ((st = join_from(lead))) return st;
        for (int j : g)
            if ((st = one(j))) return st;
    }
    for (size_t i = 0; i < b->ctx.size(); ++i)
        if (!grouped(b, (int)i) && (st = one((int)i))) return st;
    return PS_OK;
}

int capture(ps_batch* b)
{
    drop_graph(b);
    BK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
    int st = enqueue(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->origin, &g);
    if (st != PS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return bfail(b, PS_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    BK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->dirty = false;
    return PS_OK;
}

}  // namespace

extern "C" {

int ps_batch_create(ps_batch** out, int nprob, const int* ncell, const int* bc, const double* dx,
                    const int* reg_order, int precision)
{
    if (!out || nprob < 1 || !ncell || !bc || !dx || !reg_order) return PS_EINVAL;
    *out = nullptr;
    ps_batch* b = new (std::nothrow) ps_batch();
    if (!b) return PS_ENOMEM;
    b->ncell.assign(ncell, ncell + 3 * nprob);
    b->bc.assign(bc, bc + 3 * nprob);
    b->dx.assign(dx, dx + 3 * nprob);
    b->reg.assign(reg_order, reg_order + nprob);
    int st = PS_OK;
    for (int i = 0; i < nprob && st == PS_OK; ++i) {
        ps_ctx* c = nullptr;
        st = ps_create(&c);
        if (st != PS_OK) break;
        b->ctx.push_back(c);
        st = ps_set_precision(c, precision);
        if (st == PS_OK) st = ps_setup3d(c, ncell + 3 * i, bc + 3 * i, dx + 3 * i, reg_order[i]);
        if (st != PS_OK) b->err = "problem " + std::to_string(i) + ": " + ps_last_error(c);
    }
    if (st == PS_OK && hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess) st = PS_EHIP;
    b->elem_bytes = precision == PS_FP64 ? 8 : 4;
    if (st == PS_OK) st = form_groups(b);
    if (st == PS_OK) {
        b->fork.resize(nprob);
        b->join.resize(nprob);
        for (int i = 0; i < nprob && st == PS_OK; ++i)
            if (hipEventCreateWithFlags(&b->fork[i], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess)
                st = PS_EHIP;
    }
    if (st != PS_OK) {
        ps_batch_destroy(b);
        return st;
    }
    *out = b;
    return PS_OK;
}

int ps_batch_autotune(ps_batch* b, int iters)
{
    if (!b || iters < 1) return PS_EINVAL;
    const int n = (int)b->ctx.size();
    std::vector<bool> done(n, false);
    for (int i = 0; i < n; ++i) {
        if (done[i]) continue;
        ps_tuning t;
        int st = sub(b, b->ctx[i], ps_autotune(b->ctx[i], iters, &t, nullptr));
        if (st) return st;
        done[i] = true;
        for (int j = i + 1; j < n; ++j) {
            bool same = true;
            for (int d = 0; d < 3; ++d)
                same = same && b->ncell[3 * j + d] == b->ncell[3 * i + d] && b->bc[3 * j + d] == b->bc[3 * i + d];
            if (!same) continue;
            ps_ctx* c = b->ctx[j];
            if ((st = sub(b, c, ps_set_tuning(c, &t)))) return st;
            if ((st = sub(b, c, ps_setup3d(c, &b->ncell[3 * j], &b->bc[3 * j], &b->dx[3 * j], b->reg[j])))) return st;
            done[j] = true;
        }
    }
    int st = form_groups(b);
    if (st) return st;
    drop_graph(b);
    b->warmed = false;
    return PS_OK;
}

int ps_batch_destroy(ps_batch* b)
{
    if (!b) return PS_EINVAL;
    if (b->origin) (void)hipStreamSynchronize(b->origin);
    drop_graph(b);
    for (hipEvent_t e : b->fork)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (ps_ctx* c : b->ctx) ps_destroy(c);
    for (void* p : b->blocks)
        if (p) (void)hipFree(p);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
    return PS_OK;
}

int ps_batch_set_ops(ps_batch* b, int grad, int div, int curl, int reproject)
{
    if (!b) return PS_EINVAL;
    for (ps_ctx* c : b->ctx) {
        ps_set_return_grad(c, grad);
        ps_set_return_div(c, div);
        ps_set_return_curl(c, curl);
        ps_set_reprojection(c, reproject);
    }
    b->dirty = true;
    b->warmed = false;
    return PS_OK;
}

int ps_batch_context(ps_batch* b, int prob, ps_ctx** ctx)
{
    int st = check(b, prob);
    if (st) return st;
    if (!ctx) return PS_EINVAL;
    *ctx = b->ctx[prob];
    return PS_OK;
}

int ps_batch_push(ps_batch* b, int prob, const double* bx, const double* by, const double* bz)
{
    int st = check(b, prob);
    if (st) return st;
    const int m = psamd::ctx_rhs_mask(b->ctx[prob]);
    st = sub(b, b->ctx[prob], ps_push(b->ctx[prob], bx, by, bz));
    note_push(b, prob, m);
    return st;
}

int ps_batch_push_device(ps_batch* b, int prob, const void* bx, const void* by, const void* bz)
{
    int st = check(b, prob);
    if (st) return st;
    const int m = psamd::ctx_rhs_mask(b->ctx[prob]);
    st = sub(b, b->ctx[prob], ps_push_device(b->ctx[prob], bx, by, bz));
    note_push(b, prob, m);
    return st;
}

int ps_batch_solve(ps_batch* b)
{
    if (!b) return PS_EINVAL;
    int st;
    // every member must be solvable before anything is launched
    std::vector<unsigned long> sig(b->ctx.size());
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        if ((st = sub(b, b->ctx[i], psamd::ctx_check_solvable(b->ctx[i])))) return st;
        psamd::ctx_reset_held(b->ctx[i]);
        sig[i] = psamd::ctx_signature(b->ctx[i]);
    }
    // a member set up again (new buffers), pushed with other components or
    // given other options since the capture: solve eagerly, then re-capture
    if (b->warmed && sig != b->sig) {
        b->warmed = false;
        drop_graph(b);
    }
    if (!b->warmed) {
        // eager pass: allocates every work buffer outside the capture
        if ((st = enqueue(b))) return st;
        BK(hipStreamSynchronize(b->origin));
        for (ps_ctx* c : b->ctx)
            if ((st = sub(b, c, psamd::ctx_mark_solved(c)))) return st;
        b->warmed = true;
        b->dirty = true;
        b->sig = sig;
        return PS_OK;
    }
    if (b->dirty && (st = capture(b))) return st;
    BK(hipGraphLaunch(b->exec, b->origin));
    BK(hipStreamSynchronize(b->origin));
    for (ps_ctx* c : b->ctx)
        if ((st = sub(b, c, psamd::ctx_mark_solved(c)))) return st;
    return PS_OK;
}

int ps_batch_pull(ps_batch* b, int prob, double* ax, double* ay, double* az, double* bx, double* by, double* bz)
{
    int st = check(b, prob);
    if (st) return st;
    return sub(b, b->ctx[prob], ps_pull(b->ctx[prob], ax, ay, az, bx, by, bz));
}

int ps_batch_profile(ps_batch* b, int iters, double* ms)
{
    if (!b || iters < 1 || !ms) return PS_EINVAL;
    int st;
    if (!b->warmed && (st = ps_batch_solve(b))) return st;
    if ((st = ps_batch_solve(b))) return st;      // capture + one replay
    hipEvent_t e0, e1;
    BK(hipEventCreate(&e0));
    BK(hipEventCreate(&e1));
    BK(hipEventRecord(e0, b->origin));
    for (int i = 0; i < iters; ++i) BK(hipGraphLaunch(b->exec, b->origin));
    BK(hipEventRecord(e1, b->origin));
    BK(hipEventSynchronize(e1));
    float t = 0.f;
    BK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / iters;
    return PS_OK;
}

int ps_batch_grouped(const ps_batch* b) { return b ? b->ngrouped : -1; }

const char* ps_batch_last_error(const ps_batch* b) { return b ? b->err.c_str() : "NULL batch"; }

}  // extern "C"
