// Solver context and C ABI (include/ps_capi.h).
//
// One context = one reference `poisson_solver` object on one GPU.  The
// per-solve pipeline is five kernel launches on the context's stream:
//   x pass   real rows -> half spectra, R2C two rows per pencil   (solve3d.cpp:118-173)
//   y pass   zero-padded My pencils                               (solve3d.cpp:183-237)
//   z pass   pad -> FFT -> Ghat*zeta*operator -> IFFT -> truncate (solve3d.cpp:243-441)
//   y^-1     inverse, keep ny                                     (solve3d.cpp:447-536)
//   x^-1     C2R, keep nx, real part                              (solve3d.cpp:541-629, pull.cpp:101-131)
// The two MPI transposes of the reference (map.cpp) are the strides of the
// strided passes: nothing is moved between passes except by the passes.
//
// A 2D context (setup2d) holds its grid (nx, ny) as (nx, 1, ny): the x passes
// are the same, there is no y pass, and the fused pass along the third axis
// is the y transform of solve2d.cpp:166-367 with the 2D operators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ps_capi.h"
#include "ps_greens.hpp"
#include "ps_io.hpp"
#include "ps_launch.hpp"
#include "ps_mfma_launch.hpp"
#include "ps_plan.hpp"
#include "ps_static.hpp"

using namespace psamd;

namespace {

struct DevPlan {
    Plan1D p;
    void* tw = nullptr;      // complex T table of length L
    int tile = 1;            // B (strided) or P (x pass)
    bool stat = false;       // compile-time specialised kernels for this length
};

// Tile shapes, layouts and kernel variants (ps_tuning in ps_capi.h).
using Tuning = ps_tuning;

int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    if (!v || !*v) return dflt;
    return atoi(v);
}

Tuning tuning_default()
{
    Tuning t;
    t.zblk = env_int("PS_TUNE_ZBLK", -2);
    t.tile_x = env_int("PS_TUNE_PX", 0);
    t.tile_y = env_int("PS_TUNE_BY", 0);
    t.tile_z = env_int("PS_TUNE_BZ", 0);
    t.radix_e = env_int("PS_TUNE_E", 16);
    t.lds_kb = env_int("PS_TUNE_LDS_KB", 80);
    t.xcd = env_int("PS_TUNE_XCD", 1);
    t.specialised = env_int("PS_TUNE_STATIC", 1);
    t.specialised3 = env_int("PS_TUNE_STATIC3", 1);
    t.twreg_z = env_int("PS_TUNE_TWREG_Z", 1);
    t.twreg_y = env_int("PS_TUNE_TWREG_Y", -1);
    t.occ4_z = env_int("PS_TUNE_OCC4_Z", 0);
    t.zwave = env_int("PS_TUNE_ZWAVE", 1);
    t.halfb_z = env_int("PS_TUNE_HALFB_Z", -1);
    t.halfb_y = env_int("PS_TUNE_HALFB_Y", 0);
    t.mfma = env_int("PS_TUNE_MFMA", 1);
    t.t2d = env_int("PS_TUNE_T2D", 1);
    t.zfan = env_int("PS_TUNE_ZFAN", 1);
    return t;
}

}  // namespace

struct ps_ctx {
    int prec = PS_FP64;
    bool grad = false, div = false, curl = false, reproj = false;
    bool ready = false;
    bool dim2 = false;      // set up by setup2d: grid (nx, ny) held as (nx, 1, ny)
    int ncell[3] = {0, 0, 0}, bc[3] = {0, 0, 0};
    double dx[3] = {0, 0, 0};
    int reg = 0;
    int M[3] = {0, 0, 0}, O[3] = {0, 0, 0};
    int H = 0, Hp = 0;
    long n3 = 0;
    int device = 0;
    hipStream_t stream = nullptr;

    void* rhs = nullptr;    // 3 x n3 real, pushed sources
    void* lhs = nullptr;    // 3 x n3 real, solutions
    bool ext_rl = false;    // rhs / lhs are slices of a batch's block (not freed here)
    void* rhsf = nullptr;   // 3 x n3 real, the held rhs as pull returns it (x < nx) when it is `spec`
    void* spec = nullptr;   // 3 x A1c complex: x spectra of the held rhs over the x-extended rows
    void* pot = nullptr;    // pot_cap x n3 real: outputs of the last potential solve (compact, by bit)
    int pot_cap = 0;
    int pot_mask = 0;       // quantities (PS_POT_*) that pot holds
    void* stage = nullptr;  // kStageReals doubles: host-pointer push / pull staging (FP32 forms)
    void* A1 = nullptr;     // ncomp_a1 x nz x ny x Hp complex
    void* A2 = nullptr;     // ncomp_a2 x nz x My x Hp complex (not used by 2D contexts)
    long A1c = 0, A2c = 0;  // complex elements per component
    int ncomp_a1 = 0, ncomp_a2 = 0;
    void* G = nullptr;      // octant of Ghat*zeta/N, real
    void* mfrag = nullptr;  // matrix-core DFT fragments (PS_MIXED)

    DevPlan px, py, pz1, pz3;
    Tuning tune = {};        // used by the last setup
    Tuning tune_req = {};    // set by ps_set_tuning
    bool tune_set = false;   // tune_req replaces the defaults
    int zb = 0;             // A2 layout: log2 z block (-1 = [ky][z][kx])
    int zwave = 0;          // z pass uses the pencil-per-wavefront kernel
    bool zres = false;      // FP32 z pencils of 2048 (nz = 1024 unbounded) as two residues of 1024
    void* tw_half = nullptr;   // 1024-point twiddle table of the two-residue z kernel
    bool zmfma = false;     // PS_MIXED: one-component z pass on the matrix cores
    int g_layout = 0;       // Gk octant: 0 [ky][kz][kx], 1 [ky][kx][kz]
    bool t2d = false;       // 2D: fused pass reads the x pass output transposed, A2 = [c][kx][nyp]
    long nyp = 0;           // 2D transposed rows: the y extent rounded up to 16
    Addr a2_ypass;          // A2 addressing seen by the y passes (slab = z, pos = ky)
    Addr a2_zpass;          // A2 addressing seen by the z pass (slab = ky, pos = z)

    bool rhs_present[3] = {false, false, false};
    bool lhs_present[3] = {false, false, false};
    // The held rhs (the input of the next solve and what pull returns as rhs),
    // as the reference keeps it in rhsX/Y/Z: after a push, the pushed field;
    // after a solve with a filter (regularisation, or reprojection of three
    // components), the filtered field on the x-extended rows
    // (solve3d.cpp:403-439, :447-629), which the next solve maps and
    // transforms again (map.cpp:50-60).  Here the filter is applied when
    // something reads the held rhs (filt_pending), and the filtered field is
    // kept as its x spectra (`spec`, src_spec), from which the next solve
    // starts at the y pass.
    bool src_spec = false;      // held rhs = spec (else the pushed `rhs`)
    bool filt_pending = false;  // held rhs = filter(held rhs), not yet applied
    bool filt_reproj = false;   // ... with reprojection
    bool rhsf_valid = false;    // rhsf holds the real rows of spec
    unsigned gen = 0;           // setup generation (a batch re-captures when it changes)

    size_t mem_used = 0;        // device bytes held by this context
    size_t mem_cap = ~(size_t)0;   // ps_set_device_memory_limit
    std::vector<std::pair<void*, size_t>> allocs;

    hipEvent_t ev[6] = {};
    bool ev_made = false;

    std::string err;
};

namespace {

int real_bytes(const ps_ctx* c) { return c->prec == PS_FP64 ? 8 : 4; }

int fail(ps_ctx* c, int status, const std::string& msg)
{
    if (c) c->err = msg;
    return status;
}

#define CK(call)                                                                                 \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(c, e_ == hipErrorOutOfMemory ? PS_ENOMEM : PS_EHIP,                      \
                        std::string(#call) + ": " + hipGetErrorString(e_));                      \
    } while (0)

// Device memory of a context: counted against its limit
// (ps_set_device_memory_limit); a request beyond it fails as the runtime's
// own out-of-memory does.
hipError_t dev_alloc(ps_ctx* c, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes > c->mem_cap || c->mem_used > c->mem_cap - bytes) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    c->allocs.emplace_back(*p, bytes);
    c->mem_used += bytes;
    return hipSuccess;
}

void free_ptr(ps_ctx* c, void*& p)
{
    if (!p) return;
    for (size_t i = 0; i < c->allocs.size(); ++i)
        if (c->allocs[i].first == p) {
            c->mem_used -= c->allocs[i].second;
            c->allocs[i] = c->allocs.back();
            c->allocs.pop_back();
            break;
        }
    (void)hipFree(p);
    p = nullptr;
}

void release(ps_ctx* c)
{
    if (c->ext_rl) {
        c->rhs = c->lhs = nullptr;
        c->ext_rl = false;
    } else {
        free_ptr(c, c->rhs);
        free_ptr(c, c->lhs);
    }
    free_ptr(c, c->rhsf);
    free_ptr(c, c->spec);
    free_ptr(c, c->pot);
    c->pot_mask = c->pot_cap = 0;
    free_ptr(c, c->stage);
    free_ptr(c, c->A1);
    free_ptr(c, c->A2);
    free_ptr(c, c->G);
    free_ptr(c, c->mfrag);
    free_ptr(c, c->tw_half);
    c->zres = false;
    for (DevPlan* d : {&c->px, &c->py, &c->pz1, &c->pz3}) free_ptr(c, d->tw);
    c->ncomp_a1 = c->ncomp_a2 = 0;
    c->ready = false;
    c->src_spec = c->filt_pending = c->filt_reproj = c->rhsf_valid = false;
}

template <typename T> int upload_twiddles(ps_ctx* c, DevPlan* d)
{
    std::vector<double> re, im;
    twiddles(d->p.L, re, im);
    std::vector<T> h(2 * (size_t)d->p.L);
    for (int j = 0; j < d->p.L; ++j) {
        h[2 * j] = (T)re[j];
        h[2 * j + 1] = (T)im[j];
    }
    CK(dev_alloc(c, &d->tw, h.size() * sizeof(T)));
    CK(hipMemcpy(d->tw, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return PS_OK;
}

int plan_dim(ps_ctx* c, DevPlan* d, int L, int ncomp, const char* what)
{
    std::string why;
    const int pref = c->tune.radix_e;
    if (make_plan(L, max_E_for(real_bytes(c), ncomp), &d->p, &why, pref)) return PS_OK;
    // a length whose radices need a larger E than the multi-component budget:
    // use the one-component bound (correct, more register pressure)
    if (ncomp > 1 && make_plan(L, max_E_for(real_bytes(c), 1), &d->p, &why, pref)) return PS_OK;
    return fail(c, PS_EUNSUPPORTED, std::string(what) + ": " + why);
}

// A1 of n1 and A2 of n2 components (kept when already as large)
int ensure_comp_buffers(ps_ctx* c, int n1, int n2)
{
    const size_t cb = 2 * (size_t)real_bytes(c);
    if (n1 > c->ncomp_a1) {
        free_ptr(c, c->A1);
        c->ncomp_a1 = 0;
        CK(dev_alloc(c, &c->A1, (size_t)n1 * c->A1c * cb));
        c->ncomp_a1 = n1;
    }
    if (n2 > c->ncomp_a2 && c->A2c > 0) {
        free_ptr(c, c->A2);
        c->ncomp_a2 = 0;
        CK(dev_alloc(c, &c->A2, (size_t)n2 * c->A2c * cb));
        c->ncomp_a2 = n2;
    }
    return PS_OK;
}

int ensure_real(ps_ctx* c, void** p)
{
    if (*p) return PS_OK;
    CK(dev_alloc(c, p, 3 * (size_t)c->n3 * real_bytes(c)));
    return PS_OK;
}

// x spectra of the held rhs: three components shaped like A1
int ensure_spec(ps_ctx* c)
{
    if (c->spec) return PS_OK;
    CK(dev_alloc(c, &c->spec, 3 * (size_t)c->A1c * 2 * real_bytes(c)));
    return PS_OK;
}

// Operator selection of solve3d.cpp:43-107 and solve2d.cpp:44-102
// (grad > div > curl > plain).
struct OpSel {
    int op = OP_PLAIN;
    int nc = 1;                 // 1: component-wise kernel, 3: mixes components
    int in_first = 0, in_step = 1, nin = 0;
    int out_first = 0, out_step = 1, nout = 0;
    int kmap = KMAP_3D;         // axis of each vector component (KMap)
};

bool arith(const bool present[3], int* first, int* step, int* n)
{
    int idx[3], k = 0;
    for (int d = 0; d < 3; ++d)
        if (present[d]) idx[k++] = d;
    *n = k;
    if (k == 0) return false;
    *first = idx[0];
    *step = (k >= 2) ? idx[1] - idx[0] : 1;
    return true;
}

// 2D operators (solve2d.cpp:292-326).  grad and curl differentiate one
// scalar in two directions.  In single precision with the runtime-plan
// kernel (long pencils) the scalar is transformed once and one launch writes
// both components (nc 2).  On the transposed path (long pencils on the
// specialised kernel) the scalar is also transformed once: the fused pass is
// out of place and each output component reads the same transposed copy.
// Otherwise the x pass transforms the scalar twice
// (input step 0) and each output is an in-place one-component z pass: the
// specialised, pencil-per-wavefront and matrix-core kernels are
// component-wise, and the FP64 two-output kernel would exceed 256 registers
// (its input spectrum and an output in flight together) and spill.
int select_op2d(ps_ctx* c, OpSel* s)
{
    const bool one_read = c->prec != PS_FP64 && !c->pz1.stat && !c->zmfma && !c->zwave;
    // transposed path: one x pass of the scalar, both outputs read it (nc 1, nin 1)
    const int nin1 = (one_read || c->t2d) ? 1 : 2;
    const bool rX = c->rhs_present[0], rY = c->rhs_present[1], rZ = c->rhs_present[2];
    if (!arith(c->rhs_present, &s->in_first, &s->in_step, &s->nin))
        return fail(c, PS_ESTATE, "[poisson_solver.solve2d]: no rhs field has been pushed");
    if (c->grad) {
        if (!rX)
            return fail(c, PS_ESTATE,
                        "[poisson_solver.solve2d]: Error rhs fields have not been pushed correctly for grad option");
        s->op = OP_GRAD;
        s->nc = one_read ? 2 : 1;
        s->kmap = KMAP_2D_GRAD;
        s->in_first = 0; s->in_step = 0; s->nin = nin1;
        s->out_first = 0; s->out_step = 1; s->nout = 2;
    } else if (c->div) {
        if (!(rX && rY))
            return fail(c, PS_ESTATE,
                        "[poisson_solver.solve2d]: Error rhs fields have not been pushed correctly for div option");
        s->op = OP_DIV;
        s->nc = 3;
        s->kmap = KMAP_2D_DIV;
        s->in_first = 0; s->in_step = 1; s->nin = 2;
        s->out_first = 0; s->out_step = 1; s->nout = 1;
    } else if (c->curl) {
        if (!rZ)
            return fail(c, PS_ESTATE,
                        "[poisson_solver.solve2d]: Error rhs fields have not been pushed correctly for curl option");
        s->op = OP_GRAD;
        s->nc = one_read ? 2 : 1;
        s->kmap = KMAP_2D_CURL;
        s->in_first = 2; s->in_step = 0; s->nin = nin1;
        s->out_first = 0; s->out_step = 1; s->nout = 2;
    } else {
        s->op = OP_PLAIN;
        s->nc = 1;
        s->out_first = s->in_first; s->out_step = s->in_step; s->nout = s->nin;
    }
    return PS_OK;
}

int select_op(ps_ctx* c, OpSel* s)
{
    if (c->dim2) return select_op2d(c, s);
    const bool rX = c->rhs_present[0], rY = c->rhs_present[1], rZ = c->rhs_present[2];
    if (!arith(c->rhs_present, &s->in_first, &s->in_step, &s->nin))
        return fail(c, PS_ESTATE, "[poisson_solver.solve3d]: no rhs field has been pushed");
    if (c->grad) {
        if (!rX)
            return fail(c, PS_ESTATE,
                        "[poisson_solver.solve3d]: Error rhs fields have not been pushed correctly for grad option");
        if (!(rY && rZ))
            return fail(c, PS_ESTATE,
                        "[poisson_solver.solve3d]: grad is component-wise and needs rhs X, Y and Z");
        s->op = OP_GRAD;
        s->nc = 1;
        s->out_first = 0; s->out_step = 1; s->nout = 3;
    } else if (c->div) {
        if (!(rX && rY && rZ))
            return fail(c, PS_ESTATE,
                        "[poisson_solver.solve3d]: Error rhs fields have not been pushed correctly for div option");
        s->op = OP_DIV;
        s->nc = 3;
        s->out_first = 0; s->out_step = 1; s->nout = 1;
    } else if (c->curl) {
        if (!(rX && rY && rZ))
            return fail(c, PS_ESTATE,
                        "[poisson_solver.solve3d]: Error rhs fields have not been pushed correctly for curl option");
        s->op = OP_CURL;
        s->nc = 3;
        s->out_first = 0; s->out_step = 1; s->nout = 3;
    } else {
        s->op = OP_PLAIN;
        s->nc = 1;
        s->out_first = s->in_first; s->out_step = s->in_step; s->nout = s->nin;
    }
    return PS_OK;
}

// The y^-1 outputs of the potential form that come from A's z output: A
// (ik mode 0), d/dx A (mode 1) and d/dy A (mode 2).  With xdx the x^-1
// pass derives d/dx A from A's y^-1 output (times i kx; FP32 storage,
// specialised x kernels, tuning zfan), so d/dx A has no y^-1 output of its own when A is
// wanted, and takes A's (mode 0) when it is not.  fan: these y^-1 outputs
// run as one launch from one load of A's z output (all three of them, FP32
// storage, the specialised y kernel, tuning zfan; for two, two launches are
// faster: 2 x 7.1 against 15.5 ms at 1024^3).
struct PotYPlan {
    int n, qn[3], mode[3];
    bool xdx, fan;
};

PotYPlan pot_yplan(const ps_ctx* c, int outputs)
{
    PotYPlan p{};
    p.xdx = (outputs & PS_POT_DX) && c->prec != PS_FP64 && c->px.stat && c->tune.zfan && !c->dim2;
    for (int qn = 0; qn < 3; ++qn) {
        if (!((outputs >> qn) & 1)) continue;
        int mode = qn;
        if (qn == 1 && p.xdx) {
            if (outputs & PS_POT_A) continue;
            mode = 0;
        }
        p.qn[p.n] = qn;
        p.mode[p.n] = mode;
        ++p.n;
    }
    p.fan = p.n >= 3 && c->prec != PS_FP64 && c->py.stat && c->tune.zfan && !c->dim2;
    return p;
}

template <typename T>
void fill_zargs(ps_ctx* c, PencilArgs<T>& a, const DevPlan& d, void* A1, void* A2)
{
    a.plan = device_plan<T>(d.p, reinterpret_cast<const cplx<T>*>(d.tw));
    a.in = a.out = c->dim2 ? A1 : A2;                 // 2D: the x pass output, in place
    a.ia = a.oa = c->a2_zpass;                        // slab = ky, pos = kz, col = kx
    a.xcd_remap = c->tune.xcd;
    a.slab_pair = 1;
    a.twreg = c->tune.twreg_z;
    a.occ4 = c->tune.occ4_z;
    a.n_in = a.n_out = c->ncell[2];
    a.ncol = c->H;
    a.B = d.tile;
    a.ncoltile = (c->H + a.B - 1) / a.B;
    a.nslab = c->M[1];
    a.load_mode = LOAD_COMPLEX;
    a.store_mode = STORE_COMPLEX;
    a.do_fwd = a.do_inv = 1;
    a.G = reinterpret_cast<const T*>(c->G);
    a.g_ky = (long)c->O[0] * c->O[2];
    a.g_kx = c->g_layout ? c->O[2] : 1;
    a.g_kz = c->g_layout ? 1 : c->O[0];
    a.zwave = c->zwave;
    a.Mx = c->M[0];
    a.My = c->M[1];
    a.Mz = c->M[2];
    for (int k = 0; k < 3; ++k) a.dk[k] = 1.0 / ((double)c->M[k] * c->dx[k]);
    const double h = std::max(std::max(c->dx[0], c->dx[1]), c->dx[2]);
    a.zarg = 0.5 * (2.0 * h) * (2.0 * h);
    int regm = c->reg;
    if (regm > 0 && regm % 2) regm += 1;
    a.reg = regm > 0 ? regm : 0;
    a.invN = 1.0 / ((double)c->M[0] * c->M[1] * c->M[2]);
    a.mfrag = c->mfrag;
    a.mfrag16 = c->mfrag ? static_cast<const unsigned short*>(c->mfrag) + kZmfmaFragHalves : nullptr;
    a.fone = 1.0f;
    a.fzero = 0.0f;
}

// Work buffers of a pipeline supplied by the caller instead of the
// context's own (A1, A2 of the shapes the context's setup chose).
struct WorkBufs {
    void* A1;
    void* A2;
};

// Where a pipeline starts and ends (run_pipeline flags).
enum PipeFlags : int {
    PF_SRC_SPEC = 1,   // no x pass: the y pass reads the held x spectra (spec), components as for src
    PF_DST_SPEC = 2,   // no x^-1 pass: the y^-1 pass writes x spectra into spec, components as for dst
    PF_X_ONLY = 4,     // only the x pass, writing the x spectra of src into spec (components of dst)
    PF_XINV_ONLY = 8,  // only the x^-1 pass, from spec (components of src) into dst
    PF_FANOUT = 16,    // one input (nin 1) through x and y once; the z pass reads it for each
                       // of the nout outputs (component stride 0) and writes them beside it
};

// The five passes (three for a 2D context).  `src` holds nin real components
// starting at in_first with step in_step (step 0: one component transformed
// nin times); `dst` receives nout components at out_first/out_step.
// zarg / reg >= 0 replace the regularisation filter's argument scale and
// order (the Gaussian of smooth3d is order 2 with zarg = h^2 / 2).
// `flags` (PipeFlags) start or end the pipeline at the held x spectra.
template <typename T>
int run_pipeline(ps_ctx* c, const void* src, int in_first, int in_step, int nin, int op, int nc, void* dst,
                 int out_first, int out_step, int nout, hipEvent_t* ev, int kmap = KMAP_3D, double zarg = -1.0,
                 int reg = -1, int flags = 0, const WorkBufs* wb = nullptr)
{
    hipStream_t s = c->stream;
    const int nx = c->ncell[0], ny = c->ncell[1], nz = c->ncell[2];
    const size_t cbytes = 2 * sizeof(T);
    if (flags & (PF_SRC_SPEC | PF_DST_SPEC | PF_X_ONLY | PF_XINV_ONLY)) {
        int r = ensure_spec(c);
        if (r) return r;
    }
    char* spec = static_cast<char*>(c->spec);
    if (flags & (PF_X_ONLY | PF_XINV_ONLY)) {
        // one strided pass between real rows and the held x spectra
        XArgs<T> xa{};
        xa.plan = device_plan<T>(c->px.p, reinterpret_cast<const cplx<T>*>(c->px.tw));
        xa.r_row = nx;
        xa.c_row = c->Hp;
        xa.nrows = ny * nz;
        xa.nvalid = nx;
        xa.H = c->H;
        xa.P = c->px.tile;
        xa.npairs = (ny * nz + 1) / 2;
        if (flags & PF_X_ONLY) {
            xa.in = reinterpret_cast<const T*>(src) + (long)in_first * c->n3;
            xa.r_comp = (long)in_step * c->n3;
            xa.out = spec + (size_t)out_first * c->A1c * cbytes;
            xa.c_comp = (long)out_step * c->A1c;
            xa.ncomp = nin;
            if (c->px.stat) CK(static_launch_x<T>(c->px.p.L, xa, 0, c->bc[0] == 0, s));
            else CK(launch_x_r2c<T>(c->px.p.E, xa, s));
        } else {
            xa.in = spec + (size_t)in_first * c->A1c * cbytes;
            xa.c_comp = (long)in_step * c->A1c;
            xa.out = reinterpret_cast<T*>(dst) + (long)out_first * c->n3;
            xa.r_comp = (long)out_step * c->n3;
            xa.ncomp = nout;
            if (c->px.stat) CK(static_launch_x<T>(c->px.p.L, xa, 1, c->bc[0] == 0, s));
            else CK(launch_x_c2r<T>(c->px.p.E, xa, s));
            // spec holds the unnormalised x spectra: the inverse divides by Mx
            for (int k = 0; k < nout; ++k) {
                T* p = reinterpret_cast<T*>(dst) + (long)(out_first + k * out_step) * c->n3;
                CK(launch_scale(p, p, sizeof(T), c->n3, 1.0 / c->M[0], s));
            }
        }
        return PS_OK;
    }
    void* A1 = wb ? wb->A1 : nullptr;
    void* A2 = wb ? wb->A2 : nullptr;
    const bool fan = (flags & PF_FANOUT) != 0;
    // potential fan-out: output c (the c-th set bit of kmap) goes to A1 /
    // dst component c.  kx and ky are constant along a z pencil, so d/dx A
    // and d/dy A are i kx and i ky times the z-pass spectrum of A: the z pass
    // transforms A (OP_PLAIN), d/dz A (OP_GRAD on k_z) and the Laplacian
    // (OP_LAPL) only, into A2 components 1 + zslot, and the y^-1 pass of
    // d/dx and d/dy reads A's and multiplies by i k on load (PencilArgs::ik_in).
    const bool potfan = fan && op == OP_POTGRAD;
    if ((op == OP_POTGRAD) != fan || (fan && (nin != 1 || nc != 1 || c->dim2)))
        return fail(c, PS_EINVAL, "the potential fan-out needs one input, one z component and a 3D context");
    int slot[5] = {-1, -1, -1, -1, -1}, zslot[5] = {-1, -1, -1, -1, -1}, nzq = 0;
    if (potfan) {
        for (int qn = 0, k = 0; qn < 5; ++qn) slot[qn] = ((kmap >> qn) & 1) ? k++ : -1;
        for (int qn : {0, 3, 4})
            if (qn == 0 ? (kmap & 7) != 0 : ((kmap >> qn) & 1) != 0) zslot[qn] = nzq++;
        zslot[1] = zslot[2] = zslot[0];
    }
    const PotYPlan yp = potfan ? pot_yplan(c, kmap) : PotYPlan{};
    if (!wb) {
        const int n = std::max(std::max(nin, nc), nout);
        int r = ensure_comp_buffers(c, potfan ? nout : n, potfan ? 1 + nzq : n);
        if (r) return r;
        A1 = c->A1;
        A2 = c->A2;
    }
    if (ev) CK(hipEventRecord(ev[0], s));
    if ((flags & PF_SRC_SPEC) && c->dim2) {
        // 2D: the fused pass reads the x pass output A1 (in place); start it
        // from copies of the held spectra
        for (int k = 0; k < nin; ++k)
            CK(hipMemcpyAsync(static_cast<char*>(A1) + (size_t)k * c->A1c * cbytes,
                              spec + (size_t)(in_first + k * in_step) * c->A1c * cbytes, c->A1c * cbytes,
                              hipMemcpyDeviceToDevice, s));
    }

    XArgs<T> xa{};
    xa.plan = device_plan<T>(c->px.p, reinterpret_cast<const cplx<T>*>(c->px.tw));
    xa.in = reinterpret_cast<const T*>(src) + (long)in_first * c->n3;
    xa.out = A1;
    xa.r_comp = (long)in_step * c->n3;
    xa.r_row = nx;
    xa.c_comp = c->A1c;
    xa.c_row = c->Hp;
    xa.nrows = ny * nz;
    xa.nvalid = nx;
    xa.H = c->H;
    xa.P = c->px.tile;
    xa.npairs = (ny * nz + 1) / 2;
    xa.ncomp = nin;
    if (!(flags & PF_SRC_SPEC)) {
        if (c->px.stat) CK(static_launch_x<T>(c->px.p.L, xa, 0, c->bc[0] == 0, s));
        else CK(launch_x_r2c<T>(c->px.p.E, xa, s));
    }
    if (ev) CK(hipEventRecord(ev[1], s));

    PencilArgs<T> ya{};
    if (!c->dim2) {
    ya.plan = device_plan<T>(c->py.p, reinterpret_cast<const cplx<T>*>(c->py.tw));
    ya.in = A1;
    ya.out = A2;
    ya.ia = addr_plain(c->A1c, (long)ny * c->Hp, c->Hp, 1);
    if (flags & PF_SRC_SPEC) {
        ya.in = spec + (size_t)in_first * c->A1c * cbytes;
        ya.ia.comp = (long)in_step * c->A1c;
    }
    ya.oa = c->a2_ypass;
    ya.xcd_remap = c->tune.xcd;
    ya.twreg = c->tune.twreg_y;
    ya.n_in = ny;
    ya.n_out = c->M[1];
    ya.ncol = c->H;
    ya.B = c->py.tile;
    ya.ncoltile = (c->H + ya.B - 1) / ya.B;
    ya.nslab = nz;
    ya.ncomp = nin;
    ya.load_mode = LOAD_COMPLEX;
    ya.store_mode = STORE_COMPLEX;
    ya.do_fwd = 1;
    ya.do_inv = 0;
    ya.op = OP_NONE;
    if (c->py.stat) CK(static_launch_pencil<T>(c->py.p.L, ya, 1, SM_FWD, c->bc[1] == 0, s));
    else CK(launch_pencil<T>(c->py.p.E, 1, ya, s));
    }
    if (ev) CK(hipEventRecord(ev[2], s));

    const DevPlan& zp = (nc == 3) ? c->pz3 : c->pz1;
    PencilArgs<T> za{};
    fill_zargs<T>(c, za, zp, A1, A2);
    za.op = op;
    za.kmap = kmap;
    if (zarg >= 0.0) za.zarg = zarg;
    if (reg >= 0) za.reg = reg;
    za.ncomp = (nc == 3) ? 1 : nin;     // nc 2: one input, outputs 0 and 1
    if (fan) {
        // the one input spectrum (A2 component 0) is read by each z launch below
        za.ia.comp = 0;
        za.ncomp = 1;
    }
    if (c->t2d) {
        // 2D transposed path: A1 [c][y][Hp] -> A2 [c][kx][nyp], then the fused
        // pass reads contiguous y pencils of A2 and writes A1; one input read by
        // several one-component outputs (grad, curl) has component stride 0
        CK(launch_transpose_c(A1, A2, 2 * (int)sizeof(T), nz, c->H, c->Hp, c->nyp, c->A1c, c->A2c, nin, s));
        if (ev) CK(hipEventRecord(ev[2], s));   // profiled as the (otherwise empty) y interval
        const bool shared_in = nc == 1 && nin == 1 && nout > 1;
        za.in = A2;
        za.ia = addr_plain(shared_in ? 0 : c->A2c, 0, 1, c->nyp);
        za.out = A1;
        if (nc == 1) za.ncomp = nout;
    }
    auto zlaunch = [&](const PencilArgs<T>& zz) -> int {
        bool mfma_done = false;
        if constexpr (sizeof(T) == 4) {
            if (c->zmfma && nc == 1 && zmfma_supported(zp.p.L, c->bc[2] == 0, zz.op)) {
                CK(launch_zmfma(zz, c->tune.mfma, s));
                mfma_done = true;
            } else if (c->zmfma && nc == 3 && zz.kmap == KMAP_3D && zmfma3_supported(zp.p.L, c->bc[2] == 0, zz.op)) {
                CK(launch_zmfma3(zz, c->tune.mfma, s));
                mfma_done = true;
            }
        }
        if (nc == 2 && zp.stat) return fail(c, PS_EINVAL, "one-input two-output z pass needs the runtime-plan kernel");
        if constexpr (sizeof(T) == 4) {
            if (!mfma_done && c->zres && nc == 1) {
                CK((s_launch_zres<float, 1024>(zz, c->tw_half, s)));
                mfma_done = true;
            }
        }
        if (mfma_done) {
        } else if (zp.stat) CK(static_launch_pencil<T>(zp.p.L, zz, nc, SM_FUSED, c->bc[2] == 0, s));
        else CK(launch_pencil<T>(zp.p.E, nc, zz, s));
        return PS_OK;
    };
    char* A2o = static_cast<char*>(A2) + (size_t)c->A2c * cbytes;
    bool zfan_done = false;
    if constexpr (sizeof(T) == 4) {
        // the z outputs from one forward transform where the kernel has it
        // (zslot 0 .. nzq - 1 in mask order)
        const int zmask = (zslot[0] >= 0 ? 1 : 0) | (zslot[3] >= 0 ? 2 : 0) | (zslot[4] >= 0 ? 4 : 0);
        if (potfan && c->zmfma && nzq > 1 && c->tune.zfan) {
            PencilArgs<T> zq = za;
            zq.ncomp = 1;
            zq.out = A2o;
            if (zmfma_fan_ok(zq, zmask, c->tune.mfma)) {
                CK(launch_zmfma_fan(zq, zmask, s));
                zfan_done = true;
            }
        }
    }
    if (zfan_done) {
    } else if (potfan) {
        for (int qn : {0, 3, 4}) {
            if (zslot[qn] < 0) continue;
            PencilArgs<T> zq = za;
            zq.op = qn == 0 ? OP_PLAIN : qn == 3 ? OP_GRAD : OP_LAPL;
            zq.kmap = qn == 3 ? KMAP_Z : KMAP_3D;
            zq.ncomp = 1;
            zq.out = A2o + (size_t)zslot[qn] * c->A2c * cbytes;
            if (const int zr = zlaunch(zq)) return zr;
        }
    } else if (const int zr = zlaunch(za)) {
        return zr;
    }
    if (ev) CK(hipEventRecord(ev[3], s));

    if (!c->dim2) {
    PencilArgs<T> yi = ya;
    yi.in = fan ? static_cast<void*>(A2o) : A2;
    yi.out = A1;
    yi.ia = c->a2_ypass;
    yi.oa = addr_plain(c->A1c, (long)ny * c->Hp, c->Hp, 1);
    if (flags & PF_DST_SPEC) {
        yi.out = spec + (size_t)out_first * c->A1c * cbytes;
        yi.oa.comp = (long)out_step * c->A1c;
    }
    // (the pipeline's 1/Mx is folded into the kernel multipliers: spec is
    // rescaled below to the unnormalised x spectra the x pass would give)
    yi.n_in = c->M[1];
    yi.n_out = ny;
    yi.ncomp = nout;
    yi.do_fwd = 0;
    yi.do_inv = 1;
    auto ylaunch = [&](const PencilArgs<T>& y) -> int {
        if (c->py.stat) CK(static_launch_pencil<T>(c->py.p.L, y, 1, SM_INV, c->bc[1] == 0, s));
        else CK(launch_pencil<T>(c->py.p.E, 1, y, s));
        return PS_OK;
    };
    if (potfan) {
        // one launch per output: d/dx and d/dy read A's slot
        yi.Mx = c->M[0];
        yi.My = c->M[1];
        for (int d = 0; d < 2; ++d) yi.dk[d] = 1.0 / ((double)c->M[d] * c->dx[d]);
        // the outputs from A's z output: one launch from one load of it
        // where the FP32 specialised kernel has that form, else one each
        if (yp.fan) {
            PencilArgs<T> y1 = yi;
            y1.ncomp = 1;
            y1.ik_in = 0;
            y1.ik_fan = 0;
            for (int i = 0; i < yp.n; ++i) {
                y1.ik_fan |= 1 << yp.mode[i];
                y1.fan_slot[yp.mode[i]] = slot[yp.qn[i]];
            }
            y1.in = A2o + (size_t)zslot[0] * c->A2c * cbytes;
            CK(static_launch_pencil<T>(c->py.p.L, y1, 1, SM_INVFAN, c->bc[1] == 0, s));
        } else {
            for (int i = 0; i < yp.n; ++i) {
                PencilArgs<T> y1 = yi;
                y1.ncomp = 1;
                y1.ik_in = yp.mode[i];
                y1.in = A2o + (size_t)zslot[0] * c->A2c * cbytes;
                y1.out = static_cast<char*>(yi.out) + (size_t)slot[yp.qn[i]] * yi.oa.comp * cbytes;
                if (const int yr = ylaunch(y1)) return yr;
            }
        }
        for (int qn = 3; qn < 5; ++qn) {
            if (slot[qn] < 0) continue;
            PencilArgs<T> y1 = yi;
            y1.ncomp = 1;
            y1.ik_in = 0;
            y1.in = A2o + (size_t)zslot[qn] * c->A2c * cbytes;
            y1.out = static_cast<char*>(yi.out) + (size_t)slot[qn] * yi.oa.comp * cbytes;
            if (const int yr = ylaunch(y1)) return yr;
        }
    } else if (const int yr = ylaunch(yi)) {
        return yr;
    }
    }
    if (flags & PF_DST_SPEC) {
        // y^-1 output (3D) or the fused pass's A1 (2D) times Mx: the
        // unnormalised x spectra of the held rhs
        for (int k = 0; k < nout; ++k) {
            char* o = spec + (size_t)(out_first + k * out_step) * c->A1c * cbytes;
            const char* i = c->dim2 ? static_cast<char*>(A1) + (size_t)k * c->A1c * cbytes : o;
            CK(launch_scale(i, o, sizeof(T), 2 * c->A1c, (double)c->M[0], s));
        }
    }
    if (ev) CK(hipEventRecord(ev[4], s));
    if (flags & PF_DST_SPEC) {
        if (ev) CK(hipEventRecord(ev[5], s));
        return PS_OK;
    }

    XArgs<T> xi = xa;
    xi.in = A1;
    xi.out = reinterpret_cast<T*>(dst) + (long)out_first * c->n3;
    xi.r_comp = (long)out_step * c->n3;
    xi.ncomp = nout;
    if (yp.xdx) {
        // d/dx A from A's y^-1 output: a second pass of A's workgroups, or,
        // when A is not wanted, its own grid component (its slot holds A's)
        xi.remap = 1;
        xi.dx_also = -1;
        int gc = 0, a_gc = -1;
        for (int qn = 0; qn < 5; ++qn) {
            if (slot[qn] < 0) continue;
            if (qn == 1 && a_gc >= 0) {
                xi.dx_also = a_gc;
                xi.dx_slot = slot[1];
                continue;
            }
            xi.src[gc] = slot[qn];
            xi.oslot[gc] = slot[qn];
            if (qn == 0) a_gc = gc;
            if (qn == 1) xi.ikx |= 1 << gc;
            ++gc;
        }
        xi.ncomp = gc;
        xi.Mx = c->M[0];
        xi.dk0 = 1.0 / ((double)c->M[0] * c->dx[0]);
    }
    if (c->px.stat) CK(static_launch_x<T>(c->px.p.L, xi, 1, c->bc[0] == 0, s));
    else CK(launch_x_c2r<T>(c->px.p.E, xi, s));
    if (ev) CK(hipEventRecord(ev[5], s));
    return PS_OK;
}

template <typename T> int setup_impl(ps_ctx* c)
{
    const int nx = c->ncell[0], ny = c->ncell[1], nz = c->ncell[2];
    for (int d = 0; d < 3; ++d) {
        c->M[d] = (c->bc[d] == 0) ? 2 * c->ncell[d] : c->ncell[d];     // partition_setup.cpp:71-96
        c->O[d] = c->M[d] / 2 + 1;
    }
    c->H = c->M[0] / 2 + 1;
    const int align = (real_bytes(c) == 4) ? 16 : 8;                    // 128-byte rows
    c->Hp = (c->H + align - 1) / align * align;
    c->n3 = (long)nx * ny * nz;
    c->A1c = (long)nz * ny * c->Hp;

    Tuning& tn = c->tune;
    tn = c->tune_set ? c->tune_req : tuning_default();
    // z tiles of half width where the z transform is 512 or shorter (measured
    // 3-10 % faster z passes for Mz = 256 ... 512, 2 % slower at 1024)
    if (tn.halfb_z < 0) tn.halfb_z = (c->M[2] <= 512) ? 1 : 0;
    // y-pass stage twiddles in registers for single-precision y transforms of
    // 1024 and longer (y^-1 of 512^3 0.82 -> 0.77 ms, y + y^-1 of 1024^3
    // 16.6 -> 15.1 ms); FP64 keeps them in the table (1024^3: 48.5 -> 86 ms)
    if (tn.twreg_y < 0) tn.twreg_y = (c->prec != PS_FP64 && c->M[1] >= 1024) ? 1 : 0;
    const int lds_bytes = tn.lds_kb * 1024;

    // A2 layout: z blocked by Zb = 2^zb planes, [z/Zb][ky][z%Zb][kx]; zb = 0 is
    // [z][ky][kx] (y pass writes contiguous planes, z pencils strided by a
    // plane), zb = -1 is [ky][z][kx] (z pencils contiguous, y pass strided).
    int zb = tn.zblk;
    if (zb == -2) {
        // 16-plane blocks (measured best at 512^3 on the FP32 kernels), fewer
        // when the specialised z kernels' threads per pencil (Mz / E) are fewer
        // than 16: a block must not split a thread's positions.  From 1024
        // planes 32-plane blocks (1024^3: z pass 25.8 -> 24.9 ms on the FP32
        // two-residue kernel, 26.0 -> 22.6 ms on the matrix cores; y + y^-1
        // 15.3 -> 14.0 ms against the plain layout).  The matrix-core z pass
        // of 512^3 takes the plain [z][ky][kx] layout (solve 4.34 -> 4.27 ms).
        const bool mf = c->prec == PS_MIXED && tn.mfma && zmfma_supported(c->M[2], c->bc[2] == 0, OP_PLAIN);
        zb = (nz >= 1024) ? 5 : (mf && !c->dim2) ? 0 : (nz >= 32) ? 4 : 0;
        int E1 = 0, B1 = 0, P1 = 0;
        if (zb > 0 && tn.specialised && static_policy(c->M[2], real_bytes(c), 1, &E1, &B1, &P1))
            while (zb > 0 && (c->M[2] / E1) % (1 << zb) != 0) --zb;
    }
    const long My = c->M[1];
    if (zb > 0 && (1 << zb) >= nz) zb = -1;
    if (c->dim2) zb = 0;
    c->zb = zb;
    if (c->dim2) {
        // no y pass: the fused pass reads the x pass output, rows of Hp
        c->A2c = 0;
        c->a2_ypass = addr_plain(0, 0, 0, 1);
        c->a2_zpass = addr_plain(c->A1c, c->Hp, c->Hp, 1);
    } else if (zb == 0) {
        c->A2c = (long)nz * My * c->Hp;
        c->a2_ypass = addr_plain(c->A2c, My * c->Hp, c->Hp, 1);
        c->a2_zpass = addr_plain(c->A2c, c->Hp, My * c->Hp, 1);
    } else if (zb < 0) {
        c->A2c = (long)nz * My * c->Hp;
        c->a2_ypass = addr_plain(c->A2c, c->Hp, (long)nz * c->Hp, 1);
        c->a2_zpass = addr_plain(c->A2c, (long)nz * c->Hp, c->Hp, 1);
    } else {
        const long Zb = 1L << zb, nzb = (nz + Zb - 1) / Zb;
        c->A2c = nzb * My * Zb * c->Hp;
        // y pass: slab = z (blocked), pos = ky
        c->a2_ypass = Addr{c->A2c, 1, c->Hp, My * Zb * c->Hp, Zb * c->Hp, 0, zb, 30};
        // z pass: slab = ky, pos = z (blocked)
        c->a2_zpass = Addr{c->A2c, 1, Zb * c->Hp, 0, c->Hp, My * Zb * c->Hp, 30, zb};
    }

    int r;
    if ((r = plan_dim(c, &c->px, c->M[0], 1, "x transform"))) return r;
    if ((r = plan_dim(c, &c->py, c->M[1], 1, "y transform"))) return r;
    if ((r = plan_dim(c, &c->pz1, c->M[2], 1, "z transform"))) return r;
    if ((r = plan_dim(c, &c->pz3, c->M[2], 3, "z transform (3 components)"))) return r;
    const int eb = real_bytes(c);
    c->px.tile = tn.tile_x > 0 ? tn.tile_x : choose_P(c->px.p, eb, (ny * nz + 1) / 2, lds_bytes);
    c->py.tile = tn.tile_y > 0 ? tn.tile_y : choose_B(c->py.p, eb, c->H, lds_bytes);
    c->pz1.tile = tn.tile_z > 0 ? tn.tile_z : choose_B(c->pz1.p, eb, c->H, lds_bytes);
    c->pz3.tile = choose_B(c->pz3.p, eb, c->H, lds_bytes);
    // specialised kernels: fixed E and tile shape per length
    for (DevPlan* d : {&c->px, &c->py, &c->pz1, &c->pz3}) {
        int E = 0, B = 0, P = 0;
        const bool allowed = tn.specialised && (d != &c->pz3 || tn.specialised3);
        const bool have = allowed && static_policy(d->p.L, eb, d == &c->pz3 ? 3 : 1, &E, &B, &P);
        // FP64 with 20 elements per thread (lengths 320, 640: 80 VGPRs of data
        // per thread) is slower on the specialised fused z and x kernels than on
        // the runtime-plan ones (z 0.39 / 4.51 ms against 0.28 / 2.41 ms at
        // 160^3 / 320^3); the y passes keep the specialised kernels (320^3:
        // 1.03 against 1.40 ms for y and y^-1)
        if (have && !(eb == 8 && E == 20 && d != &c->py)) {
            d->stat = true;
            d->p.E = E;
            d->p.TP = d->p.L / E;
            d->tile = (d == &c->px) ? P : B;
            // tile-width divisor: z (fused) pass 1, 2 or 4; y passes 1 or 2
            const bool zp = (d == &c->pz1 || d == &c->pz3);
            int div = zp ? tn.halfb_z : (d == &c->py ? tn.halfb_y : 0);
            div = zp ? (div >= 2 ? 4 : (div == 1 ? 2 : 1)) : (div ? 2 : 1);
            if (d != &c->px && B / div >= 1 && B % div == 0) d->tile = B / div;
        }
    }
    if (c->zb > 0) {
        for (DevPlan* d : {&c->pz1, &c->pz3})
            if (d->stat && (d->p.TP % (1 << c->zb)) != 0) d->stat = false;
    }
    for (DevPlan* d : {&c->px, &c->py, &c->pz1, &c->pz3}) {
        if (d->tile < 1 || d->p.TP * d->tile > kMaxThreads)
            return fail(c, PS_EINVAL, "tile shape exceeds the workgroup size bound");
        if (!d->stat && (size_t)d->tile * lds_len(d->p.L) * 2 * eb > 160 * 1024)
            return d->tile == 1 ? fail(c, PS_EUNSUPPORTED, "transform length " + std::to_string(d->p.L) +
                                                               " exceeds the pencil held in LDS at this precision")
                                : fail(c, PS_EINVAL, "tile shape exceeds the LDS of a compute unit");
    }
    for (DevPlan* d : {&c->px, &c->py, &c->pz1, &c->pz3})
        if ((r = upload_twiddles<T>(c, d))) return r;

    // pencil-per-wavefront z kernel: specialised one-component plan, pencil
    // inside one wavefront, two or more stages; it reads Gk along kz
    // (its elements per thread: 16, or 32 for FP32 pencils of 2048 (zwave_E))
    const int zwE = (c->pz1.p.L > 1024 && c->prec != PS_FP64) ? 32 : 16;
    // planner (zwave 1): FP64 pencils of 1024 run the fused specialised kernel
    // (z pass of 512^3 4.16 against 4.53 ms; the wavefront kernel is ahead up
    // to 256 and level at 512); zwave 2 still selects the wavefront kernel
    const bool zw_fp64_long = tn.zwave == 1 && c->prec == PS_FP64 && c->pz1.p.L >= 1024;
    c->zwave = tn.zwave && !zw_fp64_long && c->pz1.stat && c->pz1.p.L % zwE == 0 &&
               (64 % (c->pz1.p.L / zwE) == 0) && c->pz1.p.nstage >= 2 &&
               (c->zb <= 0 || (c->pz1.p.L / zwE) % (1 << c->zb) == 0);
    if (c->zwave && zwE == 32) c->pz1.tile = tn.halfb_z ? kZwaveB32 / 2 : kZwaveB32;
    if (c->zwave && zwE == 16) {
        // the wavefront kernel is built for the policy's tile width and half
        // of it; a quarter-width tile (halfb_z = 2) runs the fused kernel
        int E1 = 0, B1 = 0, P1 = 0;
        if (!static_policy(c->pz1.p.L, eb, 1, &E1, &B1, &P1) || (c->pz1.tile != B1 && c->pz1.tile != B1 / 2))
            c->zwave = 0;
    }
    // matrix-core z pass (PS_MIXED): its lanes read Gk along kx; positions
    // z0 + 32 u must be linear in u (z blocks of at most 32 planes, or plain).
    // The 1024-point kernel uses 32-bit byte offsets: one component of A2 and
    // the Gk octant below 4 GB; the 2048-point (two-residue) kernel 64-bit ones
    const long zcomp = c->dim2 ? c->A1c : c->A2c;
    const bool off32 = zcomp * 8 < (1L << 32) && (long)c->O[0] * c->O[1] * c->O[2] * 4 < (1L << 32);
    c->zmfma = c->prec == PS_MIXED && tn.mfma && zmfma_supported(c->M[2], c->bc[2] == 0, OP_PLAIN) &&
               c->zb <= 5 && (c->M[2] != 1024 || off32);     // the 1024 kernel's offsets are 32-bit
    // two-residue z kernel: FP32 pencils of 2048 holding nz = 1024 unbounded
    // cells run as two 1024-point wavefront transforms (zwave = 1; zwave = 2
    // keeps one pencil of 2048 per wavefront with 32 elements per thread)
    c->zres = c->prec != PS_FP64 && tn.zwave == 1 && !c->zmfma && c->pz1.stat && c->pz1.p.L == 2048 &&
              c->bc[2] == 0 && (c->zb <= 0 || (1024 / 16) % (1 << c->zb) == 0);
    if (c->zres) {
        int E1 = 0, B1 = 0, P1 = 0;
        static_policy(1024, eb, 1, &E1, &B1, &P1);
        c->zwave = 0;
        c->pz1.tile = tn.halfb_z ? B1 / 2 : B1;
        std::vector<double> re, im;
        twiddles(1024, re, im);
        std::vector<float> h(2048);
        for (int j = 0; j < 1024; ++j) {
            h[2 * j] = (float)re[j];
            h[2 * j + 1] = (float)im[j];
        }
        CK(dev_alloc(c, &c->tw_half, h.size() * sizeof(float)));
        CK(hipMemcpy(c->tw_half, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    // 2D, long y pencils (one or two columns per workgroup on the specialised
    // fused kernel): a column read in the x pass layout touches one element
    // per row; transposed, each pencil is contiguous.  The fused pass then
    // reads the transposed copy A2 [c][kx][nyp] and writes the x pass layout
    // of A1, so grad and curl need their scalar transformed only once.
    c->t2d = c->dim2 && tn.t2d && c->pz1.stat && !c->zwave && !c->zres && !c->zmfma && c->pz1.tile <= 2;
    if (c->t2d) {
        c->nyp = (nz + 15) / 16 * 16;
        c->A2c = (long)c->H * c->nyp;
    }
    // octant layout: [ky][kx][kz] for the kernels that read a pencil's row
    // along kz (wavefront, two-residue, transposed 2D and the register-pencil
    // matrix-core kernels), [ky][kz][kx] for those reading adjacent kx
    c->g_layout = (c->zmfma ? zmfma_reg_layout(c->M[2], tn.mfma) : (c->zwave || c->zres || c->t2d)) ? 1 : 0;
    if (c->zmfma) {
        std::vector<unsigned short> fr(kZmfmaFragHalves + 3 * kZmfma16FragHalves);
        zmfma_fragments(fr.data());
        for (int k = 0; k < 3; ++k) zmfma_fragments16k(fr.data() + kZmfmaFragHalves + k * kZmfma16FragHalves, k);
        CK(dev_alloc(c, &c->mfrag, fr.size() * sizeof(unsigned short)));
        CK(hipMemcpy(c->mfrag, fr.data(), fr.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    }
    CK(dev_alloc(c, &c->G, (size_t)c->O[0] * c->O[1] * c->O[2] * sizeof(T)));
    r = build_greens<T>(c->ncell, c->bc, c->dx, c->reg, reinterpret_cast<T*>(c->G), c->g_layout, c->stream,
                        &c->err, c->dim2 ? 1 : 0);
    if (r) return r;
    return PS_OK;
}

// Host bookkeeping of a solve: which lhs components it wrote, and the filter
// it leaves pending on the held rhs: regularisation (solve3d.cpp:329-343) and
// reprojection when all three components are held (:388-396; none in 2D).
void mark_lhs(ps_ctx* c, const OpSel& s)
{
    for (int k = 0; k < s.nout; ++k) c->lhs_present[s.out_first + k * s.out_step] = true;
}

void after_solve(ps_ctx* c, const OpSel& s)
{
    mark_lhs(c, s);
    const bool all3 = c->rhs_present[0] && c->rhs_present[1] && c->rhs_present[2];
    const bool reproj = !c->dim2 && c->reproj && all3;
    if (c->reg > 0 || reproj) {
        c->filt_pending = true;
        c->filt_reproj = reproj;
        c->rhsf_valid = false;
    }
}

// Apply the pending filter to the held rhs: one pipeline from the held rhs
// (pushed rows, or x spectra) to x spectra, i.e. the reference's rhs after
// its solve (the filtered field, z and y truncated, x extended).
template <typename T> int materialize_impl(ps_ctx* c)
{
    if (!c->filt_pending) return PS_OK;
    int first, step, n;
    if (!arith(c->rhs_present, &first, &step, &n)) {
        c->filt_pending = false;
        return PS_OK;
    }
    const int op = c->filt_reproj ? OP_RHSF_PROJ : OP_RHSF;
    const int nc = c->filt_reproj ? 3 : 1;
    int r = run_pipeline<T>(c, c->rhs, first, step, n, op, nc, nullptr, first, step, n, nullptr, KMAP_3D, -1.0, -1,
                            PF_DST_SPEC | (c->src_spec ? PF_SRC_SPEC : 0));
    if (r) return r;
    c->src_spec = true;
    c->filt_pending = false;
    c->rhsf_valid = false;
    return PS_OK;
}

int materialize(ps_ctx* c)
{
    if (!c->filt_pending) return PS_OK;
    return (c->prec == PS_FP64) ? materialize_impl<double>(c) : materialize_impl<float>(c);
}

// A solve of the held rhs.  mutate: the solve's filter becomes pending on the
// held rhs (a reference solve3d / solve2d); otherwise the held rhs is left as
// it was (timed repeats of one solve).
template <typename T> int solve_impl(ps_ctx* c, hipEvent_t* ev, bool mutate = true)
{
    int r = materialize(c);
    if (r) return r;
    OpSel s;
    if ((r = select_op(c, &s))) return r;
    if ((r = ensure_real(c, &c->lhs))) return r;
    r = run_pipeline<T>(c, c->rhs, s.in_first, s.in_step, s.nin, s.op, s.nc, c->lhs, s.out_first, s.out_step,
                        s.nout, ev, s.kmap, -1.0, -1, c->src_spec ? PF_SRC_SPEC : 0);
    if (r) return r;
    if (mutate) after_solve(c, s);
    else mark_lhs(c, s);
    return PS_OK;
}

// Real rows (x < nx) of the held x spectra, into rhsf, for pull.
template <typename T> int spec_rows_impl(ps_ctx* c)
{
    int first, step, n;
    if (!arith(c->rhs_present, &first, &step, &n)) return PS_OK;
    int r;
    if ((r = ensure_real(c, &c->rhsf))) return r;
    r = run_pipeline<T>(c, nullptr, first, step, n, OP_NONE, 1, c->rhsf, first, step, n, nullptr, KMAP_3D, -1.0, -1,
                        PF_XINV_ONLY);
    if (r) return r;
    c->rhsf_valid = true;
    return PS_OK;
}

// Staged host transfers: FP64 host data crosses PCIe as is and is converted
// on the device (chunks of kStageReals through one device buffer).
constexpr long kStageReals = 1L << 23;   // 64 MB of doubles

int ensure_stage(ps_ctx* c)
{
    if (c->stage) return PS_OK;
    CK(dev_alloc(c, &c->stage, (size_t)std::min(kStageReals, c->n3) * sizeof(double)));
    return PS_OK;
}

// device component d of `dst_base` (solve precision) <- host FP64 array
int push_host(ps_ctx* c, void* dst_base, int d, const double* src)
{
    char* dst = static_cast<char*>(dst_base) + (size_t)d * c->n3 * real_bytes(c);
    if (real_bytes(c) == 8) {
        CK(hipMemcpyAsync(dst, src, c->n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        return PS_OK;
    }
    int r = ensure_stage(c);
    if (r) return r;
    for (long i0 = 0; i0 < c->n3; i0 += kStageReals) {
        const long m = std::min(kStageReals, c->n3 - i0);
        CK(hipMemcpyAsync(c->stage, src + i0, m * sizeof(double), hipMemcpyHostToDevice, c->stream));
        CK(launch_cast(c->stage, 8, dst + (size_t)i0 * 4, 4, m, c->stream));
    }
    return PS_OK;
}

// host FP64 array <- device component d of `src_base`; waits for the copy
int pull_host(ps_ctx* c, double* dst, const void* src_base, int d)
{
    const char* src = static_cast<const char*>(src_base) + (size_t)d * c->n3 * real_bytes(c);
    if (real_bytes(c) == 8) {
        CK(hipMemcpyAsync(dst, src, c->n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CK(hipStreamSynchronize(c->stream));
        return PS_OK;
    }
    int r = ensure_stage(c);
    if (r) return r;
    for (long i0 = 0; i0 < c->n3; i0 += kStageReals) {
        const long m = std::min(kStageReals, c->n3 - i0);
        CK(launch_cast(src + (size_t)i0 * 4, 4, c->stage, 8, m, c->stream));
        CK(hipMemcpyAsync(dst + i0, c->stage, m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

// A push of the components in `mask`.  When a held component that is not
// pushed again carries a filtered state (a filtered solve preceded), it keeps
// it: the pending filter is applied before the new rows overwrite the pushed
// buffer (reprojection mixes components), and after the copy the pushed
// components join the held x spectra (their x pass runs then).  Otherwise the
// held rhs is simply the pushed rows.
bool push_keeps_spec(const ps_ctx* c, int mask)
{
    bool other = false;
    for (int d = 0; d < 3; ++d)
        if (c->rhs_present[d] && !((mask >> d) & 1)) other = true;
    return other && (c->src_spec || c->filt_pending);
}

int before_push(ps_ctx* c, int mask)
{
    return push_keeps_spec(c, mask) ? materialize(c) : PS_OK;
}

int after_push(ps_ctx* c, int mask)
{
    int r;
    if (push_keeps_spec(c, mask)) {
        bool m3[3] = {(mask & 1) != 0, (mask & 2) != 0, (mask & 4) != 0};
        int first, step, n;
        if (arith(m3, &first, &step, &n)) {
            r = (c->prec == PS_FP64)
                    ? run_pipeline<double>(c, c->rhs, first, step, n, OP_NONE, 1, nullptr, first, step, n, nullptr,
                                           KMAP_3D, -1.0, -1, PF_X_ONLY)
                    : run_pipeline<float>(c, c->rhs, first, step, n, OP_NONE, 1, nullptr, first, step, n, nullptr,
                                          KMAP_3D, -1.0, -1, PF_X_ONLY);
            if (r) return r;
        }
    } else {
        c->src_spec = false;
        c->filt_pending = false;
    }
    for (int d = 0; d < 3; ++d)
        if ((mask >> d) & 1) c->rhs_present[d] = true;
    c->rhsf_valid = false;
    return PS_OK;
}

// A, grad A and the Laplacian of A from the held scalar X: one x and one y
// pass, then the z pass writes each selected quantity from the one spectrum
// (OP_POTGRAD), and the inverse passes return them into pot (compact, in bit
// order).  The held rhs is left as it was.
template <typename T> int solve_potential_impl(ps_ctx* c, int mask, hipEvent_t* ev)
{
    int r = materialize(c);
    if (r) return r;
    if (!c->rhs || !c->rhs_present[0])
        return fail(c, PS_ESTATE, "solve3d_potential: the scalar rhs X has not been pushed");
    const int nout = __builtin_popcount((unsigned)mask);
    if (nout > c->pot_cap) {
        c->pot_mask = 0;
        free_ptr(c, c->pot);
        c->pot_cap = 0;
        CK(dev_alloc(c, &c->pot, (size_t)nout * c->n3 * sizeof(T)));
        c->pot_cap = nout;
    }
    r = run_pipeline<T>(c, c->rhs, 0, 1, 1, OP_POTGRAD, 1, c->pot, 0, 1, nout, ev, mask, -1.0, -1,
                        PF_FANOUT | (c->src_spec ? PF_SRC_SPEC : 0));
    if (r) return r;
    c->pot_mask = mask;
    return PS_OK;
}

int check_ready(ps_ctx* c)
{
    if (!c) return PS_EINVAL;
    if (!c->ready) return fail(c, PS_ESTATE, "setup3d or setup2d has not been called");
    return PS_OK;
}

// Set-up shared by setup3d and setup2d; ncell/bc/dx are the internal 3D form.
int setup_common(ps_ctx* c, const int ncell[3], const int bc[3], const double dx[3], int reg_order, bool dim2)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(c, PS_EHIP, "no HIP device available");
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    CK(hipGetDevice(&c->device));
    if (!c->stream) CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int d = 0; d < 3; ++d) {
        c->ncell[d] = ncell[d];
        c->bc[d] = bc[d];
        c->dx[d] = dx[d];
        c->rhs_present[d] = c->lhs_present[d] = false;
    }
    c->dim2 = dim2;
    c->reg = reg_order;
    ++c->gen;
    int r = (c->prec == PS_FP64) ? setup_impl<double>(c) : setup_impl<float>(c);
    if (r) {
        release(c);
        return r;
    }
    c->ready = true;
    return PS_OK;
}

// Where pull reads the held rhs (components as pushed, x < nx): the pushed
// rows, or the real rows of the held x spectra after a filtered solve
// (pull.cpp:101-131 takes the real part of the x pencils).
int rhs_source(ps_ctx* c, const void** src)
{
    int r = materialize(c);
    if (r) return r;
    *src = c->rhs;
    if (!c->src_spec) return PS_OK;
    if (!c->rhsf_valid) {
        r = (c->prec == PS_FP64) ? spec_rows_impl<double>(c) : spec_rows_impl<float>(c);
        if (r) return r;
    }
    *src = c->rhsf;
    return PS_OK;
}

// A strided block's cell (i, j, k) of the user grid in the context's 3D form:
// a 2D context holds (nx, ny) as (nx, 1, ny), so user j is internal k.
bool strided_args(const ps_ctx* c, const long stride[4], const long first[3], int ncomp, Strided* st, long* off,
                  long* span)
{
    if (ncomp < 1 || ncomp > 3) return false;
    long s[4] = {stride[0], stride[1], stride[2], stride[3]};
    long f[3] = {first[0], first[1], first[2]};
    const int n[3] = {c->ncell[0], c->dim2 ? c->ncell[2] : c->ncell[1], c->dim2 ? 1 : c->ncell[2]};
    for (int d = 0; d < 4; ++d)
        if (s[d] < 1) return false;
    for (int d = 0; d < 3; ++d)
        if (f[d] < 0) return false;
    *off = f[0] * s[1] + f[1] * s[2] + f[2] * s[3];
    *span = (long)(ncomp - 1) * s[0] + (f[0] + n[0] - 1) * s[1] + (f[1] + n[1] - 1) * s[2] +
            (f[2] + n[2] - 1) * s[3] + 1;
    if (c->dim2) *st = Strided{s[0], s[1], 0, s[2]};   // internal j = 0, internal k = user j
    else *st = Strided{s[0], s[1], s[2], s[3]};
    return true;
}

// a temporary device buffer of a context, freed at scope exit
struct CtxScratch {
    ps_ctx* c;
    void* p = nullptr;
    ~CtxScratch() { free_ptr(c, p); }
};

}  // namespace

extern "C" {

int ps_create(ps_ctx** out)
{
    if (!out) return PS_EINVAL;
    *out = new (std::nothrow) ps_ctx();
    return *out ? PS_OK : PS_ENOMEM;
}

int ps_destroy(ps_ctx* c)
{
    if (!c) return PS_EINVAL;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    if (c->ev_made)
        for (hipEvent_t& e : c->ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PS_OK;
}

int ps_set_precision(ps_ctx* c, int precision)
{
    if (!c) return PS_EINVAL;
    if (precision != PS_FP64 && precision != PS_FP32 && precision != PS_MIXED)
        return fail(c, PS_EINVAL, "unknown precision");
    if (c->ready) return fail(c, PS_ESTATE, "set the precision before setup3d");
    c->prec = precision;
    return PS_OK;
}

int ps_setup3d(ps_ctx* c, const int ncell[3], const int bc[3], const double dx[3], int reg_order)
{
    if (!c) return PS_EINVAL;
    if (!ncell || !bc || !dx) return fail(c, PS_EINVAL, "setup3d: NULL argument");
    for (int d = 0; d < 3; ++d) {
        if (ncell[d] < 1) return fail(c, PS_EINVAL, "setup3d: ncell must be positive");
        if (!(dx[d] > 0.0)) return fail(c, PS_EINVAL, "setup3d: dx must be positive");
        if (bc[d] != 0 && bc[d] != 1) return fail(c, PS_EBC, "setup3d: bc must be 0 (unbounded) or 1 (periodic)");
    }
    const bool uuu = !bc[0] && !bc[1] && !bc[2], ppu = bc[0] && bc[1] && !bc[2];
    const bool uup = !bc[0] && !bc[1] && bc[2], ppp = bc[0] && bc[1] && bc[2];
    if (!(uuu || ppu || uup || ppp)) {
        const int sum = bc[0] + bc[1] + bc[2];
        if (sum == 1)
            return fail(c, PS_EBC, "[poisson_solver.greens3d]: For unbounded-unbounded-periodic domains the "
                                   "z-direction must be the periodic direction");
        return fail(c, PS_EBC, "[poisson_solver.greens3d]: For unbounded-periodic-periodic domains the "
                               "z-direction must be the unbounded direction");
    }
    if (reg_order < 0) return fail(c, PS_EINVAL, "setup3d: reg_order must be >= 0");
    return setup_common(c, ncell, bc, dx, reg_order, false);
}

int ps_setup2d(ps_ctx* c, const int ncell[2], const int bc[2], const double dx[2], int reg_order)
{
    if (!c) return PS_EINVAL;
    if (!ncell || !bc || !dx) return fail(c, PS_EINVAL, "setup2d: NULL argument");
    for (int d = 0; d < 2; ++d) {
        if (ncell[d] < 1) return fail(c, PS_EINVAL, "setup2d: ncell must be positive");
        if (!(dx[d] > 0.0)) return fail(c, PS_EINVAL, "setup2d: dx must be positive");
        if (bc[d] != 0 && bc[d] != 1)
            return fail(c, PS_EBC, "[poisson_solver.greens2d]: Boundary condition configuration unknown.");
    }
    if (reg_order < 0) return fail(c, PS_EINVAL, "setup2d: reg_order must be >= 0");
    // (nx, ny) held as (nx, 1, ny): the middle axis is one periodic cell whose
    // spacing dx does not change max(dx, dy)
    const int n3[3] = {ncell[0], 1, ncell[1]};
    const int b3[3] = {bc[0], 1, bc[1]};
    const double d3[3] = {dx[0], dx[0], dx[1]};
    return setup_common(c, n3, b3, d3, reg_order, true);
}

int ps_set_return_grad(ps_ctx* c, int t) { if (!c) return PS_EINVAL; c->grad = t != 0; return PS_OK; }
int ps_set_return_div(ps_ctx* c, int t) { if (!c) return PS_EINVAL; c->div = t != 0; return PS_OK; }
int ps_set_return_curl(ps_ctx* c, int t) { if (!c) return PS_EINVAL; c->curl = t != 0; return PS_OK; }
int ps_set_reprojection(ps_ctx* c, int t) { if (!c) return PS_EINVAL; c->reproj = t != 0; return PS_OK; }

int ps_push(ps_ctx* c, const double* bx, const double* by, const double* bz)
{
    int r = check_ready(c);
    if (r) return r;
    if ((r = ensure_real(c, &c->rhs))) return r;
    const double* b[3] = {bx, by, bz};
    int mask = 0;
    for (int d = 0; d < 3; ++d)
        if (b[d]) mask |= 1 << d;
    if ((r = before_push(c, mask))) return r;
    for (int d = 0; d < 3; ++d)
        if (b[d] && (r = push_host(c, c->rhs, d, b[d]))) return r;
    if ((r = after_push(c, mask))) return r;
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_push_device(ps_ctx* c, const void* bx, const void* by, const void* bz)
{
    int r = check_ready(c);
    if (r) return r;
    if ((r = ensure_real(c, &c->rhs))) return r;
    const void* b[3] = {bx, by, bz};
    const size_t bytes = c->n3 * real_bytes(c);
    int mask = 0;
    for (int d = 0; d < 3; ++d)
        if (b[d]) mask |= 1 << d;
    if ((r = before_push(c, mask))) return r;
    for (int d = 0; d < 3; ++d)
        if (b[d])
            CK(hipMemcpyAsync(static_cast<char*>(c->rhs) + d * bytes, b[d], bytes, hipMemcpyDeviceToDevice,
                              c->stream));
    if ((r = after_push(c, mask))) return r;
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_push_strided(ps_ctx* c, const double* base, int ncomp, const long stride[4], const long first[3])
{
    int r = check_ready(c);
    if (r) return r;
    if (!base || !stride || !first) return fail(c, PS_EINVAL, "push_strided: NULL argument");
    Strided st;
    long off, span;
    if (!strided_args(c, stride, first, ncomp, &st, &off, &span))
        return fail(c, PS_EINVAL, "push_strided: ncomp must be 1..3, strides positive and offsets non-negative");
    if ((r = ensure_real(c, &c->rhs))) return r;
    if ((r = before_push(c, (1 << ncomp) - 1))) return r;
    CtxScratch stage{c};
    CK(dev_alloc(c, &stage.p, (size_t)span * sizeof(double)));
    CK(hipMemcpyAsync(stage.p, base, (size_t)span * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CK(launch_gather(static_cast<const double*>(stage.p) + off, st, c->rhs, real_bytes(c), ncomp, c->ncell[0],
                     c->ncell[1], c->ncell[2], c->stream));
    if ((r = after_push(c, (1 << ncomp) - 1))) return r;
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_pull_strided(ps_ctx* c, double* lhs_base, int lhs_ncomp, const long lhs_stride[4], double* rhs_base,
                    int rhs_ncomp, const long rhs_stride[4], const long first[3])
{
    int r = check_ready(c);
    if (r) return r;
    if (!first || (lhs_base && !lhs_stride) || (rhs_base && !rhs_stride))
        return fail(c, PS_EINVAL, "pull_strided: NULL argument");
    for (int which = 0; which < 2; ++which) {
        double* base = which ? rhs_base : lhs_base;
        if (!base) continue;
        const int ncomp = which ? rhs_ncomp : lhs_ncomp;
        Strided st;
        long off, span;
        if (!strided_args(c, which ? rhs_stride : lhs_stride, first, ncomp, &st, &off, &span))
            return fail(c, PS_EINVAL, "pull_strided: ncomp must be 1..3, strides positive and offsets non-negative");
        const bool* held = which ? c->rhs_present : c->lhs_present;
        int mask = 0;
        for (int d = 0; d < ncomp; ++d)
            if (held[d]) mask |= 1 << d;
        if (!mask) continue;
        const void* src = c->lhs;
        if (which && (r = rhs_source(c, &src))) return r;
        // the block's ghost cells and unheld components keep their values
        CtxScratch stage{c};
        CK(dev_alloc(c, &stage.p, (size_t)span * sizeof(double)));
        CK(hipMemcpyAsync(stage.p, base, (size_t)span * sizeof(double), hipMemcpyHostToDevice, c->stream));
        CK(launch_scatter(src, real_bytes(c), static_cast<double*>(stage.p) + off, st, mask, ncomp, c->ncell[0],
                          c->ncell[1], c->ncell[2], c->stream));
        CK(hipMemcpyAsync(base, stage.p, (size_t)span * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CK(hipStreamSynchronize(c->stream));
    }
    return PS_OK;
}

int ps_smooth3d(ps_ctx* c, double h)
{
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2) return fail(c, PS_ESTATE, "smooth3d: the solver was set up by setup2d");
    if (!(h >= 0.0) || !std::isfinite(h)) return fail(c, PS_EINVAL, "smooth3d: h must be finite and >= 0");
    int first, step, n;
    if (!c->rhs || !arith(c->rhs_present, &first, &step, &n)) return PS_OK;   // nothing held: no-op
    if ((r = materialize(c))) return r;
    // exp(-h^2 |k|^2 / 2) / N is the order-2 filter with argument scale h^2 / 2;
    // the smoothed field stays x-extended, as the Fortran routine leaves it
    // in its x pencils (poisson_solver_smooth3d.f:333-386)
    const double zarg = 0.5 * h * h;
    const int flags = PF_DST_SPEC | (c->src_spec ? PF_SRC_SPEC : 0);
    r = (c->prec == PS_FP64)
            ? run_pipeline<double>(c, c->rhs, first, step, n, OP_RHSF, 1, nullptr, first, step, n, nullptr,
                                   KMAP_3D, zarg, 2, flags)
            : run_pipeline<float>(c, c->rhs, first, step, n, OP_RHSF, 1, nullptr, first, step, n, nullptr, KMAP_3D,
                                  zarg, 2, flags);
    if (r) return r;
    CK(hipStreamSynchronize(c->stream));
    c->src_spec = true;
    c->rhsf_valid = false;
    return PS_OK;
}

int ps_solve3d(ps_ctx* c)
{
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2) return fail(c, PS_ESTATE, "solve3d: the solver was set up by setup2d");
    if (!c->rhs) return fail(c, PS_ESTATE, "[poisson_solver.solve3d]: no rhs field has been pushed");
    r = (c->prec == PS_FP64) ? solve_impl<double>(c, nullptr) : solve_impl<float>(c, nullptr);
    if (r) return r;
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_dft16_fragments(unsigned short* out, int n)
{
    if (!out || n < kZmfma16FragHalves) return PS_EINVAL;
    zmfma_fragments16(out);
    return PS_OK;
}

int ps_dft16_table(unsigned short* out, int n, int kind)
{
    if (!out || n < kZmfma16FragHalves || kind < 0 || kind > 2) return PS_EINVAL;
    zmfma_fragments16k(out, kind);
    return PS_OK;
}

int ps_dft32_fragments(unsigned short* out, int n)
{
    if (!out || n < kZmfmaFragHalves) return PS_EINVAL;
    zmfma_fragments(out);
    return PS_OK;
}

int ps_zpass_variant(const ps_ctx* c)
{
    if (!c || !c->ready) return -1;
    if (c->zres) return 4;
    if (c->zmfma) return 3;
    if (c->zwave) return 2;
    return c->pz1.stat ? 1 : 0;
}

int ps_potential_zlaunches(const ps_ctx* c, int outputs)
{
    if (!c || !c->ready) return -PS_ESTATE;
    if (c->dim2 || outputs < 1 || outputs > 31) return -PS_EINVAL;
    const int zmask = ((outputs & (PS_POT_A | PS_POT_DX | PS_POT_DY)) ? 1 : 0) | ((outputs & PS_POT_DZ) ? 2 : 0) |
                      ((outputs & PS_POT_LAP) ? 4 : 0);
    const int nzq = __builtin_popcount((unsigned)zmask);
    if (c->prec == PS_MIXED && c->zmfma && c->tune.zfan && nzq > 1) {
        // the fan-out condition of the pipeline, on the context's z arguments
        PencilArgs<float> za{};
        fill_zargs<float>(const_cast<ps_ctx*>(c), za, c->pz1, nullptr, nullptr);
        za.ia.comp = 0;
        za.ncomp = 1;
        if (zmfma_fan_ok(za, zmask, c->tune.mfma)) return 1;
    }
    return nzq;
}

int ps_potential_ylaunches(const ps_ctx* c, int outputs)
{
    if (!c || !c->ready) return -PS_ESTATE;
    if (c->dim2 || outputs < 1 || outputs > 31) return -PS_EINVAL;
    const PotYPlan yp = pot_yplan(c, outputs);
    return (yp.fan ? 1 : yp.n) + ((outputs & PS_POT_DZ) ? 1 : 0) + ((outputs & PS_POT_LAP) ? 1 : 0);
}

int ps_potential_youtputs(const ps_ctx* c, int outputs)
{
    if (!c || !c->ready) return -PS_ESTATE;
    if (c->dim2 || outputs < 1 || outputs > 31) return -PS_EINVAL;
    const PotYPlan yp = pot_yplan(c, outputs);
    return yp.n + ((outputs & PS_POT_DZ) ? 1 : 0) + ((outputs & PS_POT_LAP) ? 1 : 0);
}

int ps_solve3d_potential(ps_ctx* c, int outputs)
{
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2) return fail(c, PS_ESTATE, "solve3d_potential: the solver was set up by setup2d");
    if (outputs < 1 || outputs > 31) return fail(c, PS_EINVAL, "solve3d_potential: outputs must be a mask of 1..31");
    r = (c->prec == PS_FP64) ? solve_potential_impl<double>(c, outputs, nullptr)
                             : solve_potential_impl<float>(c, outputs, nullptr);
    if (r) return r;
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

// output slot of quantity bit qn in pot, or -1 if it was not computed
static int pot_slot(const ps_ctx* c, int qn)
{
    if (!((c->pot_mask >> qn) & 1)) return -1;
    return __builtin_popcount((unsigned)c->pot_mask & ((1u << qn) - 1));
}

int ps_pull_potential(ps_ctx* c, double* A, double* dx, double* dy, double* dz, double* lap)
{
    int r = check_ready(c);
    if (r) return r;
    double* o[5] = {A, dx, dy, dz, lap};
    for (int qn = 0; qn < 5; ++qn) {
        const int k = pot_slot(c, qn);
        if (!o[qn] || k < 0) continue;
        if ((r = pull_host(c, o[qn], c->pot, k))) return r;
    }
    return PS_OK;
}

int ps_pull_potential_device(ps_ctx* c, void* A, void* dx, void* dy, void* dz, void* lap)
{
    int r = check_ready(c);
    if (r) return r;
    void* o[5] = {A, dx, dy, dz, lap};
    const size_t bytes = c->n3 * real_bytes(c);
    for (int qn = 0; qn < 5; ++qn) {
        const int k = pot_slot(c, qn);
        if (!o[qn] || k < 0) continue;
        CK(hipMemcpyAsync(o[qn], static_cast<const char*>(c->pot) + k * bytes, bytes, hipMemcpyDeviceToDevice,
                          c->stream));
    }
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_profile_potential(ps_ctx* c, int outputs, int iters, double ms[6])
{
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2) return fail(c, PS_ESTATE, "profile_potential: the solver was set up by setup2d");
    if (iters < 1 || !ms || outputs < 1 || outputs > 31) return fail(c, PS_EINVAL, "profile_potential: bad argument");
    if (!c->ev_made) {
        for (hipEvent_t& e : c->ev) CK(hipEventCreate(&e));
        c->ev_made = true;
    }
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int it = 0; it < iters; ++it) {
        r = (c->prec == PS_FP64) ? solve_potential_impl<double>(c, outputs, c->ev)
                                 : solve_potential_impl<float>(c, outputs, c->ev);
        if (r) return r;
        CK(hipEventSynchronize(c->ev[5]));
        for (int k = 0; k < 5; ++k) {
            float t = 0.f;
            CK(hipEventElapsedTime(&t, c->ev[k], c->ev[k + 1]));
            acc[k] += t;
        }
        float tt = 0.f;
        CK(hipEventElapsedTime(&tt, c->ev[0], c->ev[5]));
        acc[5] += tt;
    }
    for (int k = 0; k < 6; ++k) ms[k] = acc[k] / iters;
    return PS_OK;
}

int ps_set_device_memory_limit(ps_ctx* c, unsigned long long bytes)
{
    if (!c) return PS_EINVAL;
    c->mem_cap = bytes ? (size_t)bytes : ~(size_t)0;
    return PS_OK;
}

int ps_device_memory_used(const ps_ctx* c, unsigned long long* bytes)
{
    if (!c || !bytes) return PS_EINVAL;
    *bytes = c->mem_used;
    return PS_OK;
}

}  // extern "C"

// ---- solve without allocation: work space in a static device array -------
namespace {

// Work space of ps_solve3d_noalloc: A1 and A2 of one solve.  Each device the
// library runs on holds its own copy of the array until the process ends.
constexpr size_t kNoallocScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` per device

// The array's device address on device `dev` (current), looked up once per
// device; host code uses this address, never the symbol itself.
int scratch_address(ps_ctx* c, int dev, void** out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(c, PS_EHIP, "solve3d_noalloc: device index beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(c, PS_EHIP, "solve3d_noalloc: no device address for the work space");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PS_OK;
}

template <typename T> int solve_noalloc_impl(ps_ctx* c, const void* src, int src_mask, void* dst)
{
    bool held[3];
    for (int d = 0; d < 3; ++d) {
        held[d] = c->rhs_present[d];
        c->rhs_present[d] = (src_mask >> d) & 1;
    }
    OpSel s;
    int r = select_op(c, &s);
    for (int d = 0; d < 3; ++d) c->rhs_present[d] = held[d];
    if (r) return r;
    const size_t cb = 2 * sizeof(T);
    const size_t ncomp = (size_t)std::max(std::max(s.nin, s.nc), s.nout);
    const size_t a1 = (ncomp * c->A1c * cb + 255) / 256 * 256;
    const size_t a2 = ncomp * c->A2c * cb;
    if (a1 + a2 > kNoallocScratchBytes)
        return fail(c, PS_EUNSUPPORTED,
                    "solve3d_noalloc: this grid and operator need " + std::to_string(a1 + a2) +
                        " bytes of work space; the static maximum is " + std::to_string(kNoallocScratchBytes));
    int dev = 0;
    CK(hipGetDevice(&dev));
    if (dev != c->device) return fail(c, PS_ESTATE, "solve3d_noalloc: the context's device is not current");
    void* base = nullptr;
    if ((r = scratch_address(c, dev, &base))) return r;
    const WorkBufs wb{base, static_cast<char*>(base) + a1};
    return run_pipeline<T>(c, src, s.in_first, s.in_step, s.nin, s.op, s.nc, dst, s.out_first, s.out_step, s.nout,
                           nullptr, s.kmap, -1.0, -1, 0, &wb);
}

}  // namespace

extern "C" {

int ps_solve3d_noalloc(ps_ctx* c, const void* src, int src_mask, void* dst)
{
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2) return fail(c, PS_ESTATE, "solve3d_noalloc: the solver was set up by setup2d");
    if (!src || !dst || src_mask < 1 || src_mask > 7)
        return fail(c, PS_EINVAL, "solve3d_noalloc: NULL block or component mask outside 1..7");
    return (c->prec == PS_FP64) ? solve_noalloc_impl<double>(c, src, src_mask, dst)
                                : solve_noalloc_impl<float>(c, src, src_mask, dst);
}

int ps_solve3d_async(ps_ctx* c)
{
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2) return fail(c, PS_ESTATE, "solve3d: the solver was set up by setup2d");
    if (!c->rhs) return fail(c, PS_ESTATE, "[poisson_solver.solve3d]: no rhs field has been pushed");
    return (c->prec == PS_FP64) ? solve_impl<double>(c, nullptr) : solve_impl<float>(c, nullptr);
}

int ps_solve2d(ps_ctx* c)
{
    int r = check_ready(c);
    if (r) return r;
    if (!c->dim2) return fail(c, PS_ESTATE, "solve2d: the solver was set up by setup3d");
    if (!c->rhs) return fail(c, PS_ESTATE, "[poisson_solver.solve2d]: no rhs field has been pushed");
    r = (c->prec == PS_FP64) ? solve_impl<double>(c, nullptr) : solve_impl<float>(c, nullptr);
    if (r) return r;
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_solve2d_async(ps_ctx* c)
{
    int r = check_ready(c);
    if (r) return r;
    if (!c->dim2) return fail(c, PS_ESTATE, "solve2d: the solver was set up by setup3d");
    if (!c->rhs) return fail(c, PS_ESTATE, "[poisson_solver.solve2d]: no rhs field has been pushed");
    return (c->prec == PS_FP64) ? solve_impl<double>(c, nullptr) : solve_impl<float>(c, nullptr);
}

int ps_is_2d(const ps_ctx* c) { return (c && c->ready && c->dim2) ? (c->t2d ? 2 : 1) : 0; }

int ps_pull(ps_ctx* c, double* ax, double* ay, double* az, double* bx, double* by, double* bz)
{
    int r = check_ready(c);
    if (r) return r;
    double* a[3] = {ax, ay, az};
    double* b[3] = {bx, by, bz};
    // copies run on the context's stream, after any solve enqueued on it
    for (int d = 0; d < 3; ++d) {
        if (!a[d] || !c->lhs_present[d]) continue;
        if ((r = pull_host(c, a[d], c->lhs, d))) return r;
    }
    bool want = false;
    for (int d = 0; d < 3; ++d) want = want || (b[d] && c->rhs_present[d]);
    if (!want) return PS_OK;
    const void* src = nullptr;
    if ((r = rhs_source(c, &src))) return r;
    for (int d = 0; d < 3; ++d) {
        if (!b[d] || !c->rhs_present[d]) continue;
        if ((r = pull_host(c, b[d], src, d))) return r;
    }
    return PS_OK;
}

int ps_pull_device(ps_ctx* c, void* ax, void* ay, void* az, void* bx, void* by, void* bz)
{
    int r = check_ready(c);
    if (r) return r;
    const size_t bytes = c->n3 * real_bytes(c);
    void* a[3] = {ax, ay, az};
    void* b[3] = {bx, by, bz};
    for (int d = 0; d < 3; ++d)
        if (a[d] && c->lhs_present[d])
            CK(hipMemcpyAsync(a[d], static_cast<const char*>(c->lhs) + d * bytes, bytes, hipMemcpyDeviceToDevice,
                              c->stream));
    bool want = false;
    for (int d = 0; d < 3; ++d) want = want || (b[d] && c->rhs_present[d]);
    if (want) {
        const void* src = nullptr;
        if ((r = rhs_source(c, &src))) return r;
        for (int d = 0; d < 3; ++d)
            if (b[d] && c->rhs_present[d])
                CK(hipMemcpyAsync(b[d], static_cast<const char*>(src) + d * bytes, bytes, hipMemcpyDeviceToDevice,
                                  c->stream));
    }
    CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_partition_count(const ps_ctx* c) { return (c && c->ready) ? 1 : 0; }

int ps_partition(const ps_ctx* c, int rank, int ncell[3], int icell[3], double dx[3])
{
    if (!c || !ncell || !icell || !dx) return PS_EINVAL;
    if (!c->ready) return PS_ESTATE;
    if (rank != 0) return PS_EINVAL;
    for (int d = 0; d < 3; ++d) {
        ncell[d] = c->ncell[d];
        icell[d] = 0;
        dx[d] = c->dx[d];
    }
    if (c->dim2) {   // setup2d.cpp:31-46: domain (nx, ny, 1), dx (dx, dy, 0)
        ncell[1] = c->ncell[2];
        ncell[2] = 1;
        dx[1] = c->dx[2];
        dx[2] = 0.0;
    }
    return PS_OK;
}

const char* ps_strerror(int s)
{
    switch (s) {
    case PS_OK: return "ok";
    case PS_EINVAL: return "invalid argument";
    case PS_EBC: return "unsupported boundary conditions";
    case PS_ESTATE: return "invalid solver state";
    case PS_ENOMEM: return "out of device memory";
    case PS_EHIP: return "HIP runtime error";
    case PS_EUNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
    }
}

const char* ps_last_error(const ps_ctx* c) { return c ? c->err.c_str() : "NULL context"; }

int ps_real_bytes(const ps_ctx* c) { return c ? real_bytes(c) : 0; }

void* ps_stream(ps_ctx* c) { return c ? (void*)c->stream : nullptr; }

int ps_synchronize(ps_ctx* c)
{
    if (!c) return PS_EINVAL;
    if (c->stream) CK(hipStreamSynchronize(c->stream));
    return PS_OK;
}

int ps_profile_solve(ps_ctx* c, int iters, double ms[6])
{
    int r = check_ready(c);
    if (r) return r;
    if (iters < 1 || !ms) return fail(c, PS_EINVAL, "profile: iters must be >= 1");
    if (!c->ev_made) {
        for (hipEvent_t& e : c->ev) CK(hipEventCreate(&e));
        c->ev_made = true;
    }
    // every timed solve solves the held rhs as it stands (the filter a
    // reference solve leaves on it is not applied between them)
    if ((r = materialize(c))) return r;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int it = 0; it < iters; ++it) {
        r = (c->prec == PS_FP64) ? solve_impl<double>(c, c->ev, false) : solve_impl<float>(c, c->ev, false);
        if (r) return r;
        CK(hipEventSynchronize(c->ev[5]));
        for (int k = 0; k < 5; ++k) {
            float t = 0.f;
            CK(hipEventElapsedTime(&t, c->ev[k], c->ev[k + 1]));
            acc[k] += t;
        }
        float tt = 0.f;
        CK(hipEventElapsedTime(&tt, c->ev[0], c->ev[5]));
        acc[5] += tt;
    }
    for (int k = 0; k < 6; ++k) ms[k] = acc[k] / iters;
    return PS_OK;
}

int ps_plan_describe(int L, int ncomp, int precision_bytes, int* E, int* tp, int* nstage, int radix[14])
{
    Plan1D p;
    std::string why;
    if (!make_plan(L, max_E_for(precision_bytes, ncomp), &p, &why) &&
        !(ncomp > 1 && make_plan(L, max_E_for(precision_bytes, 1), &p, &why)))
        return PS_EUNSUPPORTED;
    if (E) *E = p.E;
    if (tp) *tp = p.TP;
    if (nstage) *nstage = p.nstage;
    if (radix)
        for (int i = 0; i < kMaxStage; ++i) radix[i] = p.radix[i];
    return PS_OK;
}

int ps_eval_special(const double* x, double* y, long n, int which)
{
    if (!x || !y || n < 0 || (which != 0 && which != 1)) return PS_EINVAL;
    if (n == 0) return PS_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PS_EHIP;
    std::string err;
    return eval_special_device(x, y, n, which, &err);
}


}  // extern "C"

namespace psamd {
// State update of a solve whose launches were replayed from a captured graph.
int ctx_mark_solved(ps_ctx* c)
{
    OpSel s;
    int r = check_ready(c);
    if (r || (r = select_op(c, &s))) return r;
    after_solve(c, s);
    return PS_OK;
}
// Bit mask of the pushed rhs components.
int ctx_rhs_mask(const ps_ctx* c)
{
    return c ? (c->rhs_present[0] ? 1 : 0) | (c->rhs_present[1] ? 2 : 0) | (c->rhs_present[2] ? 4 : 0) : 0;
}

// rhs and lhs of a set-up context become slices of a caller's device blocks
// (3 n3 reals each, in the solve precision), which the context does not free;
// a later setup returns it to buffers of its own.  Pushed components and
// solutions are discarded.
int ctx_adopt_real(ps_ctx* c, void* rhs, void* lhs)
{
    int r = check_ready(c);
    if (r) return r;
    if (!rhs || !lhs) return fail(c, PS_EINVAL, "adopt: NULL block");
    if (!c->ext_rl) {
        free_ptr(c, c->rhs);
        free_ptr(c, c->lhs);
    }
    c->rhs = rhs;
    c->lhs = lhs;
    c->ext_rl = true;
    for (int d = 0; d < 3; ++d) c->rhs_present[d] = c->lhs_present[d] = false;
    c->src_spec = c->filt_pending = c->rhsf_valid = false;
    return PS_OK;
}

// A batch solve solves every member's pushed rows: the filtered rhs a
// previous solve left held is dropped.
void ctx_reset_held(ps_ctx* c)
{
    if (!c) return;
    c->src_spec = c->filt_pending = c->rhsf_valid = false;
}

// What the launches a batch captured for this context depend on: its setup
// generation, pushed components and options; 0 if it cannot be solved.
unsigned long ctx_signature(const ps_ctx* c)
{
    if (!c || !c->ready || !c->rhs) return 0;
    unsigned long sig = (unsigned long)c->gen << 16;
    sig |= (unsigned long)ctx_rhs_mask(c) << 8;
    sig |= (c->grad ? 1u : 0u) | (c->div ? 2u : 0u) | (c->curl ? 4u : 0u) | (c->reproj ? 8u : 0u) |
           (c->ext_rl ? 16u : 0u) | 32u;
    return sig;
}

// Checks the operator's inputs as a solve would, without launching.
int ctx_check_solvable(ps_ctx* c)
{
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2) return fail(c, PS_ESTATE, "solve3d: the solver was set up by setup2d");
    if (!c->rhs) return fail(c, PS_ESTATE, "[poisson_solver.solve3d]: no rhs field has been pushed");
    OpSel s;
    return select_op(c, &s);
}

// One solve over K contexts of one grid whose rhs / lhs are consecutive
// slices of 3 n3 reals of one block (ctx_adopt_real): the members' pushed
// components become the components of one plain solve on the first member's
// stream and work buffers (component stride 3 n3 for one pushed component,
// n3 for three).  Returns PS_EUNSUPPORTED, enqueuing nothing, when the members
// differ in grid, options, tuning or pushed components, or the operator is not
// plain; the caller then solves them one by one.  Host bookkeeping (which lhs
// components were written) is left to the caller (ctx_mark_solved).
template <typename T> int solve_group_impl(ps_ctx* const* m, int K)
{
    ps_ctx* c = m[0];
    OpSel s;
    int r = select_op(c, &s);
    if (r) return r;
    if (s.op != OP_PLAIN || s.nc != 1 || !(s.nin == 1 || (s.nin == 3 && s.in_step == 1)))
        return PS_EUNSUPPORTED;
    if (s.nin == 1)
        return run_pipeline<T>(c, c->rhs, s.in_first, 3, K, OP_PLAIN, 1, c->lhs, s.in_first, 3, K, nullptr);
    return run_pipeline<T>(c, c->rhs, 0, 1, 3 * K, OP_PLAIN, 1, c->lhs, 0, 1, 3 * K, nullptr);
}

int ctx_solve_group(ps_ctx* const* m, int K)
{
    if (!m || K < 1) return PS_EINVAL;
    ps_ctx* c = m[0];
    int r = check_ready(c);
    if (r) return r;
    if (c->dim2 || !c->ext_rl) return PS_EUNSUPPORTED;
    const size_t slice = 3 * (size_t)c->n3 * real_bytes(c);
    for (int j = 1; j < K; ++j) {
        const ps_ctx* o = m[j];
        if (!o || !o->ready || !o->ext_rl || o->dim2 || o->prec != c->prec) return PS_EUNSUPPORTED;
        if (o->grad != c->grad || o->div != c->div || o->curl != c->curl || o->reproj != c->reproj ||
            o->reg != c->reg || ctx_rhs_mask(o) != ctx_rhs_mask(c))
            return PS_EUNSUPPORTED;
        for (int d = 0; d < 3; ++d)
            if (o->ncell[d] != c->ncell[d] || o->bc[d] != c->bc[d] || o->dx[d] != c->dx[d])
                return PS_EUNSUPPORTED;
        if (std::memcmp(&o->tune, &c->tune, sizeof(Tuning)) != 0) return PS_EUNSUPPORTED;
        if (static_cast<const char*>(o->rhs) != static_cast<const char*>(c->rhs) + j * slice ||
            static_cast<const char*>(o->lhs) != static_cast<const char*>(c->lhs) + j * slice)
            return PS_EUNSUPPORTED;
    }
    return real_bytes(c) == 8 ? solve_group_impl<double>(m, K) : solve_group_impl<float>(m, K);
}
}  // namespace psamd

extern "C" {

int ps_tuning_default(ps_tuning* t)
{
    if (!t) return PS_EINVAL;
    *t = tuning_default();
    return PS_OK;
}

int ps_set_tuning(ps_ctx* c, const ps_tuning* t)
{
    if (!c) return PS_EINVAL;
    if (!t) {
        c->tune_set = false;
        return PS_OK;
    }
    if (t->zblk < -2 || t->zblk > 10 || t->tile_x < 0 || t->tile_x > 64 || t->tile_y < 0 || t->tile_y > 64 ||
        t->tile_z < 0 || t->tile_z > 64 || t->radix_e < 0 || t->radix_e > 70 || t->lds_kb < 16 || t->lds_kb > 160 ||
        t->halfb_z < -1 || t->halfb_z > 2 || t->halfb_y < 0 || t->halfb_y > 1 || t->mfma < 0 || t->mfma > 3 ||
        t->zfan < 0 || t->zfan > 1)
        return fail(c, PS_EINVAL, "set_tuning: field out of range");
    c->tune_req = *t;
    c->tune_set = true;
    return PS_OK;
}

int ps_get_tuning(const ps_ctx* c, ps_tuning* t)
{
    if (!c || !t) return PS_EINVAL;
    *t = c->ready ? c->tune : (c->tune_set ? c->tune_req : tuning_default());
    return PS_OK;
}

// Coordinate descent over the tuning fields: each field in turn takes the
// candidate value with the lowest average solve time (a change must win by
// 1% to count, so timing noise does not move it), the others held.  Fields
// that a setup ignores (tiles and radix split of specialised passes) are
// skipped.
int ps_autotune(ps_ctx* c, int iters, ps_tuning* best_out, double* best_ms_out)
{
    int r = check_ready(c);
    if (r) return r;
    if (iters < 1) return fail(c, PS_EINVAL, "autotune: iters must be >= 1");
    int ncell[3], bc[3];
    double dx[3];
    for (int d = 0; d < 3; ++d) {
        ncell[d] = c->ncell[d];
        bc[d] = c->bc[d];
        dx[d] = c->dx[d];
    }
    const int reg = c->reg;
    const bool dim2 = c->dim2;
    bool present[3] = {c->rhs_present[0], c->rhs_present[1], c->rhs_present[2]};
    if (c->grad || c->div || c->curl) present[0] = present[1] = present[2] = true;
    if (!present[0] && !present[1] && !present[2]) present[0] = true;

    auto trial = [&](const Tuning& t, double* ms) -> int {
        c->tune_req = t;
        c->tune_set = true;
        int st = setup_common(c, ncell, bc, dx, reg, dim2);
        if (st) return st;
        if ((st = ensure_real(c, &c->rhs))) return st;
        CK(hipMemset(c->rhs, 0, 3 * (size_t)c->n3 * real_bytes(c)));
        for (int d = 0; d < 3; ++d) c->rhs_present[d] = present[d];
        double t6[6];
        if ((st = ps_profile_solve(c, 1, t6))) return st;    // warm-up
        if ((st = ps_profile_solve(c, iters, t6))) return st;
        *ms = t6[5];
        return PS_OK;
    };

    Tuning best = c->tune;
    double best_ms = 0.0;
    if ((r = trial(best, &best_ms))) return r;
    const bool all_stat = c->px.stat && c->py.stat && c->pz1.stat;

    struct Knob {
        int Tuning::*field;
        std::vector<int> values;
        bool generic_only;   // only read by the generic (non-specialised) kernels
    };
    const std::vector<Knob> knobs = {
        {&Tuning::zblk, {-2, -1, 0, 2, 3, 4, 5}, false},
        {&Tuning::specialised, {1, 0}, false},
        {&Tuning::zwave, {1, 2, 0}, false},
        {&Tuning::halfb_z, {0, 1, 2}, false},
        {&Tuning::halfb_y, {0, 1}, false},
        {&Tuning::twreg_z, {1, 0}, false},
        {&Tuning::twreg_y, {0, 1}, false},
        {&Tuning::radix_e, {16, 8, 32}, true},
        {&Tuning::tile_y, {0, 4, 8, 16, 32}, true},
        {&Tuning::tile_z, {0, 4, 8, 16, 32}, true},
        {&Tuning::tile_x, {0, 2, 4, 8, 16}, true},
        {&Tuning::xcd, {1, 0}, false},
        {&Tuning::mfma, {1, 2, 3, 0}, false},
        {&Tuning::t2d, {1, 0}, false},
    };
    bool best_all_stat = all_stat;
    for (const Knob& k : knobs) {
        if (k.generic_only && best_all_stat && best.specialised) continue;
        if (k.field == &Tuning::t2d && !dim2) continue;              // read by 2D contexts only
        if (k.field == &Tuning::mfma && c->prec != PS_MIXED) continue;  // read by PS_MIXED only
        for (int v : k.values) {
            if (best.*k.field == v) continue;
            Tuning t = best;
            t.*k.field = v;
            double ms = 0.0;
            if (trial(t, &ms) != PS_OK) continue;     // shape not admissible here
            if (ms < 0.99 * best_ms) {
                best = t;
                best_ms = ms;
                best_all_stat = c->px.stat && c->py.stat && c->pz1.stat;
            }
        }
    }
    c->tune_req = best;
    c->tune_set = true;
    if ((r = setup_common(c, ncell, bc, dx, reg, dim2))) return r;
    if (best_out) *best_out = c->tune;
    if (best_ms_out) *best_ms_out = best_ms;
    return PS_OK;
}

}  // extern "C"
