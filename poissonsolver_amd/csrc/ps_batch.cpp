// Batches of independent solves of uneven size (C ABI in include/ps_capi.h).
//
// Each problem is a full solver context with its own plans, tile shapes and
// Green's kernel.  A batch solve is captured once into a hipGraph: an origin
// stream forks to every problem's stream, each stream records its five pass
// launches, and all branches join back; replaying the executable graph is one
// hipGraphLaunch for the whole batch.  The graph is re-captured when the set
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
}  // namespace psamd

struct ps_batch {
    std::vector<ps_ctx*> ctx;
    std::vector<int> ncell, bc, reg;    // per problem: 3, 3, 1 entries
    std::vector<double> dx;             // per problem: 3 entries
    hipStream_t origin = nullptr;
    std::vector<hipEvent_t> fork, join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;          // graph must be (re)captured
    bool warmed = false;        // one eager solve done (buffers allocated)
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

int capture(ps_batch* b)
{
    drop_graph(b);
    BK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
    int st = PS_OK;
    for (size_t i = 0; i < b->ctx.size() && st == PS_OK; ++i) {
        hipStream_t s = (hipStream_t)ps_stream(b->ctx[i]);
        if (hipEventRecord(b->fork[i], b->origin) != hipSuccess || hipStreamWaitEvent(s, b->fork[i], 0) != hipSuccess) {
            st = bfail(b, PS_EHIP, "graph capture: fork failed");
            break;
        }
        st = sub(b, b->ctx[i], ps_solve3d_async(b->ctx[i]));
        if (st == PS_OK &&
            (hipEventRecord(b->join[i], s) != hipSuccess || hipStreamWaitEvent(b->origin, b->join[i], 0) != hipSuccess))
            st = bfail(b, PS_EHIP, "graph capture: join failed");
    }
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
    if (!b->warmed) {
        // eager pass: allocates every work buffer outside the capture
        for (ps_ctx* c : b->ctx)
            if ((st = sub(b, c, ps_solve3d(c)))) return st;
        b->warmed = true;
        b->dirty = true;
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

const char* ps_batch_last_error(const ps_batch* b) { return b ? b->err.c_str() : "NULL batch"; }

}  // extern "C"
