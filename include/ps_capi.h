/*
 * ps_capi.h -- C ABI of the MI355X solve path of the Poisson solver.
 *
 * Plain C: opaque handle, plain pointers and sizes, integer status codes.
 * No exception crosses this boundary.  One context owns one HIP stream and is
 * not thread-safe (like the reference object); distinct contexts may be used
 * from distinct threads.  Each entry point names the reference interface it
 * replaces (paths relative to the reference checkout, c++/source/).
 *
 * Single-GPU semantics: nproc = 1, so the partition list has one entry that
 * covers the whole grid.  Host arrays are real double, x fastest,
 * ijk = (k*ny + j)*nx + i, as in push.cpp:82-90.
 */
#ifndef PS_CAPI_H
#define PS_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ps_ctx ps_ctx;

/* status codes */
enum {
    PS_OK = 0,
    PS_EINVAL = 1,        /* bad argument (NULL handle, non-positive size, ...) */
    PS_EBC = 2,           /* unsupported boundary-condition tuple (greens3d.cpp:522-541) */
    PS_ESTATE = 3,        /* call out of order, or rhs missing for the selected operator
                             (solve3d.cpp:57-60, :71-74, :87-90) */
    PS_ENOMEM = 4,        /* device allocation failed */
    PS_EHIP = 5,          /* HIP runtime error, or no device */
    PS_EUNSUPPORTED = 6   /* grid length with a prime factor above 7, or a transform whose pencil
                             does not fit the LDS of a compute unit */
};

/* precision of the solve (extension; the reference is FP64 only) */
enum {
    PS_FP64 = 0,          /* FP64 storage and arithmetic (the reference's precision) */
    PS_FP32 = 1,          /* FP32 storage and arithmetic, FP64-built kernel */
    PS_MIXED = 2          /* FP32 storage, MFMA DFT stages with split operands */
};

/* replaces poisson_solver::poisson_solver() (poisson_solver.hpp:107) */
int ps_create(ps_ctx** out);
/* replaces poisson_solver::~poisson_solver() (poisson_solver.hpp:108, destructor.cpp:8-90) */
int ps_destroy(ps_ctx* ctx);

/* extension: solve precision, set before ps_setup3d (default PS_FP64) */
int ps_set_precision(ps_ctx* ctx, int precision);

/* replaces poisson_solver::setup3d(int[3], int[3], double[3], int)
   (poisson_solver.hpp:111, setup3d.cpp:8-71): ncell, bc (0 unbounded,
   1 periodic), cell size, regularisation order */
int ps_setup3d(ps_ctx* ctx, const int ncell[3], const int bc[3], const double dx[3], int reg_order);
/* replaces poisson_solver::setup2d(int[2], int[2], double[2], int)
   (poisson_solver.hpp:110, setup2d.cpp:8-72, greens2d.cpp:8-441): ncell,
   bc (0 unbounded, 1 periodic; all four pairs), cell size, regularisation
   order.  The UP kernel uses the intended sample index (greens2d.cpp:400). */
int ps_setup2d(ps_ctx* ctx, const int ncell[2], const int bc[2], const double dx[2], int reg_order);

/* replace set_return_grad/div/curl and set_reprojection (settings.cpp:9-30) */
int ps_set_return_grad(ps_ctx* ctx, int toggle);
int ps_set_return_div(ps_ctx* ctx, int toggle);
int ps_set_return_curl(ps_ctx* ctx, int toggle);
int ps_set_reprojection(ps_ctx* ctx, int toggle);

/* replaces poisson_solver::push(double*, double*, double*) (push.cpp:9-397);
   NULL = component absent */
int ps_push(ps_ctx* ctx, const double* bx, const double* by, const double* bz);
/* replaces poisson_solver::solve3d() (solve3d.cpp:8-635).  Solves the held
   rhs: after a push, the pushed field; after a solve with regularisation (or
   reprojection of three components), the filtered field that solve left in
   the reference's rhs arrays -- z and y truncated, x extended over the doubled
   rows (solve3d.cpp:403-629), mapped and transformed again by the next solve
   (map.cpp:50-60).  The reference's x pencils also carry an imaginary part
   where its reprojection uses the negative Nyquist wavenumber; the held field
   here is their real part (what pull returns). */
int ps_solve3d(ps_ctx* ctx);
/* replaces poisson_solver::solve2d() (poisson_solver.hpp:118, solve2d.cpp:8-475):
   plain (every pushed component), grad (of X, two components), div (X, Y ->
   one component), curl (of the scalar Z, two components); no reprojection */
int ps_solve2d(ps_ctx* ctx);
/* replaces poisson_solver::pull(double* x6) (pull.cpp:9-469): lhs X,Y,Z then
   rhs X,Y,Z; a component is written only when the pointer is non-NULL and the
   solver holds that component */
int ps_pull(ps_ctx* ctx, double* ax, double* ay, double* az, double* bx, double* by, double* bz);

/* replaces the public member std::vector<class_partition> partition
   (poisson_solver.hpp:102, partition_setup.cpp:9-199) */
int ps_partition_count(const ps_ctx* ctx);
int ps_partition(const ps_ctx* ctx, int rank, int ncell[3], int icell[3], double dx[3]);

/* status text; the last error message recorded on a context */
const char* ps_strerror(int status);
const char* ps_last_error(const ps_ctx* ctx);

/* --- the reference's Fortran interface (fortran/source/poisson_solver/) ---
   A Fortran binding keeps one ps_ctx per solver index of
   poisson_solver_initialise(nsolvers) (poisson_solver_initialise.f:8) and
   binds the routines below; the 3-array push/pull, setup2d/3d, the setters
   and solve2d/3d bind the functions above.
   Strided block: element (c, i, j, k) of component c < ncomp and physical
   cell (i, j, k) is base[c*stride[0] + (first[0]+i)*stride[1] +
   (first[1]+j)*stride[2] + (first[2]+k)*stride[3]] -- for a Fortran array
   real_rhs(ncomp, lx:ux, ly:uy, lz:uz) passed as c_loc(real_rhs(1,lx,ly,lz)):
   stride = (1, ncomp, ncomp*nx', ncomp*nx'*ny'), first = offset - (lx, ly, lz)
   with nx' = ux-lx+1 etc.  ncomp 1, 2, 3 = X; X, Y; X, Y, Z.  2D contexts
   take k = 0 (first[2] = 0). */
/* replaces poisson_solver_push_1pointer3d (poisson_solver_push.f:8-419) */
int ps_push_strided(ps_ctx* ctx, const double* base, int ncomp, const long stride[4], const long first[3]);
/* replaces poisson_solver_pull_1pointer3d (poisson_solver_pull.f:8-493): lhs
   components X..ncomp-1 the solver holds, then (rhs_base non-NULL) the rhs
   ones; ghost cells and components the solver does not hold are left as they
   are */
int ps_pull_strided(ps_ctx* ctx, double* lhs_base, int lhs_ncomp, const long lhs_stride[4], double* rhs_base,
                    int rhs_ncomp, const long rhs_stride[4], const long first[3]);
/* output: the VTK writer of the reference's test programs
   (c++/test/unbounded3d/sources/main.cpp:329-455): <dir>/<stem>_P00.vti, an
   ImageData piece whose CellData holds nfield 3-component Float64 arrays
   (host arrays comps[3*a + d], x fastest, NULL = zeros), and <dir>/<stem>.pvti
   listing the piece.  binary = 0: ascii as the reference writes it (%17.8e);
   1: raw appended data (full precision).  Host only; no context needed. */
int ps_write_vti(const char* dir, const char* stem, const int ncell[3], const double origin[3], const double dx[3],
                 int nfield, const char* const* names, const double* const* comps, int binary);
/* replaces poisson_solver_smooth3d(ps, h) (poisson_solver_smooth3d.f:8-401):
   every held rhs component becomes Re(IFFT(exp(-h^2 |k|^2 / 2) FFT(rhs))) on
   the domain-doubled grid; 3D contexts only */
int ps_smooth3d(ps_ctx* ctx, double h);

/* --- extensions: device-resident data, streams, timing (no reference counterpart) --- */
/* real element type of device arrays: double for PS_FP64, float otherwise */
int ps_real_bytes(const ps_ctx* ctx);
/* device pointers of that element type, same layout as the host arrays */
int ps_push_device(ps_ctx* ctx, const void* bx, const void* by, const void* bz);
int ps_pull_device(ps_ctx* ctx, void* ax, void* ay, void* az, void* bx, void* by, void* bz);
/* the context's HIP stream (hipStream_t) */
void* ps_stream(ps_ctx* ctx);
int ps_synchronize(ps_ctx* ctx);
/* run `iters` solves of the held rhs and time each pass with HIP events on
   the context's stream: ms[0..4] = x, y, z, y^-1, x^-1 pass averages, ms[5] =
   whole solve.  Every timed solve solves the held rhs as it stands: the
   filter a solve3d leaves on the held rhs is not applied between them. */
int ps_profile_solve(ps_ctx* ctx, int iters, double ms[6]);
/* describe the plan chosen for transform length L: E, threads per pencil,
   number of stages and radices (up to 14); returns PS_EUNSUPPORTED if none */
int ps_plan_describe(int L, int ncomp, int precision_bytes, int* E, int* tp, int* nstage, int radix[14]);
/* evaluate the device Si (which = 0) or Ji0 (which = 1) at n host arguments */
int ps_eval_special(const double* x, double* y, long n, int which);
/* --- extension: the potential and its gradient from one scalar -----------
   No reference counterpart: the reference's grad is component-wise
   (solve3d.cpp:348-353), so a caller wanting A and grad A pushes the scalar
   three times and solves twice (c++/test/mixed_ffp3d/sources/main.cpp:164-178).
   ps_solve3d_potential transforms the held scalar rhs X once in x and y and
   computes every selected quantity from that one spectrum:
     PS_POT_A    A = G * B                  (the plain solve of X)
     PS_POT_DX.. dA/dx, dA/dy, dA/dz         (grad of X, as solve3d's grad)
     PS_POT_LAP  Laplacian of A, -|k|^2 Ghat  (div of grad A)
   with the context's regularisation.  The z pass runs one inverse z
   transform each for A, dA/dz and the Laplacian (in the mixed form at
   Mz = 512 and 2048, all of them in one launch from one forward transform;
   tuning field zfan); dA/dx and dA/dy are i kx and i ky times A's z-pass
   output, applied as their y^-1 pass loads it.  Results stay in
   the context until the next potential solve.  The held rhs is left as it
   was.  3D contexts. */
enum { PS_POT_A = 1, PS_POT_DX = 2, PS_POT_DY = 4, PS_POT_DZ = 8, PS_POT_LAP = 16 };
int ps_solve3d_potential(ps_ctx* ctx, int outputs);
/* host arrays (x fastest) of the quantities computed; NULL skips one */
int ps_pull_potential(ps_ctx* ctx, double* A, double* dAdx, double* dAdy, double* dAdz, double* lapA);
/* device arrays of the solve precision's real type */
int ps_pull_potential_device(ps_ctx* ctx, void* A, void* dAdx, void* dAdy, void* dAdz, void* lapA);
/* number of z-pass launches a potential solve of `outputs` makes on this
   context: 1 when its z outputs run from one forward transform, else one per
   distinct z multiplier (A, dA/dz, Laplacian); -PS_ESTATE before setup3d,
   -PS_EINVAL for a 2D context or a mask outside 1..31 */
int ps_potential_zlaunches(const ps_ctx* ctx, int outputs);
/* number of y^-1 launches a potential solve of `outputs` makes: one per
   y^-1 output, except that three from A's z-pass output (A, dA/dx, dA/dy)
   run as one launch from one load of it in the FP32 and mixed forms where
   the y length has a specialised kernel (tuning field zfan); negative as
   above */
int ps_potential_ylaunches(const ps_ctx* ctx, int outputs);
/* number of y^-1 outputs written: one per output, except that dA/dx is taken
   by the x^-1 pass from A's y^-1 output (times i kx) in the FP32 and mixed
   forms where the x length has a specialised kernel (tuning field zfan);
   negative as above */
int ps_potential_youtputs(const ps_ctx* ctx, int outputs);
/* `iters` potential solves timed per pass as ps_profile_solve */
int ps_profile_potential(ps_ctx* ctx, int outputs, int iters, double ms[6]);

/* device memory a context may hold (setup, work buffers, staging): a request
   beyond it fails with PS_ENOMEM, as a device out of memory does; 0 = no
   limit (the default).  Buffers already held are kept. */
int ps_set_device_memory_limit(ps_ctx* ctx, unsigned long long bytes);
/* device bytes the context holds now */
int ps_device_memory_used(const ps_ctx* ctx, unsigned long long* bytes);
/* no counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  Enqueues on the context's stream the solve of the
   context's operator on `src` (3 x n3 reals of the solve precision, component
   c at src + c n3, those in src_mask bits 0..2 present) into `dst` (same
   shape; only the components the operator writes are written), and returns
   without waiting.  Calls no allocator: the work space is a static device
   array of 256 MiB (per device, held until the process ends); a grid and
   operator needing more return PS_EUNSUPPORTED before launching.  The
   context must be set up (setup allocates) and its device current.  The held
   rhs / lhs of the context are not touched.  Calls that overlap on one device
   (any contexts, any streams) share the array and race. */
int ps_solve3d_noalloc(ps_ctx* ctx, const void* src, int src_mask, void* dst);
/* enqueue ps_solve3d on the context's stream and return without waiting;
   push, pull and solve calls that follow run after it (same stream) */
int ps_solve3d_async(ps_ctx* ctx);
/* enqueue ps_solve2d on the context's stream and return without waiting */
int ps_solve2d_async(ps_ctx* ctx);
/* nonzero if the context was set up by ps_setup2d (its passes: x, fused y, x^-1;
   ps_profile_solve reports the fused y pass as the z entry): 2 when its fused
   pass reads a transposed copy of the x pass output (tuning field t2d; the
   transpose is reported as the y entry), else 1; 0 for a 3D or unset context */
int ps_is_2d(const ps_ctx* ctx);
/* the FP16 hi/lo fragments of the 32-point DFT matrix that the matrix-core
   kernels multiply by (binary16 bit patterns; n >= 4096 entries, layout
   [lane 64][K chunk 2][Yr hi, Yr lo, Yi hi, Yi lo][8]) */
int ps_dft32_fragments(unsigned short* out, int n);
/* the same for the 16-point DFT matrix (transforms of 512; n >= 2048
   entries, layout [lane 64][Yr hi, Yr lo, Yi hi, Yi lo][8]) */
int ps_dft16_fragments(unsigned short* out, int n);
/* 16 x 16 tables W32^f(i,k) (W32 = exp(-2 pi i / 32)) in the same layout, which the
   matrix-core z pass of 1024 uses for its even/odd 16-point products: kind 0 the
   16-point DFT (f = 2ik), 1 P (f = k(2i+1)), 2 Q (f = i(2k+1)); n >= 2048 */
int ps_dft16_table(unsigned short* out, int n, int kind);
/* kernel family of the one-component z pass chosen at setup: 0 runtime-plan,
   1 compile-time specialised, 2 pencil-per-wavefront, 3 matrix-core (PS_MIXED),
   4 two-residue (FP32 z pencils of 2048 as two 1024-point transforms);
   negative: not set up */
int ps_zpass_variant(const ps_ctx* ctx);

/* --- extension: kernel tuning and the autotuner (no reference counterpart) ---
   Shapes and layouts the setup picks for the pass kernels.  Every field has a
   planner default; the PS_TUNE_* environment variables of the same names
   override the defaults.  No setting changes results beyond round-off. */
typedef struct ps_tuning {
    int zblk;         /* z blocking of the y/z-transposed spectrum: log2 block > 0,
                         0 = [z][ky][kx], -1 = [ky][z][kx], -2 = planner: 16-plane blocks,
                         fewer for short z pencils, 32-plane blocks from 1024 planes, none
                         for the matrix-core z pass of 512 planes (PS_TUNE_ZBLK) */
    int tile_x;       /* row pairs per workgroup, x passes; 0 = planner (PS_TUNE_PX) */
    int tile_y;       /* columns per workgroup, y passes; 0 = planner (PS_TUNE_BY) */
    int tile_z;       /* columns per workgroup, z pass; 0 = planner (PS_TUNE_BZ) */
    int radix_e;      /* radix split: preferred elements per thread of the generic
                         kernels (PS_TUNE_E, default 16) */
    int lds_kb;       /* LDS budget per workgroup for the planner, KB (PS_TUNE_LDS_KB, 80) */
    int xcd;          /* XCD-contiguous work ids in the strided passes (PS_TUNE_XCD, 1) */
    int specialised;  /* compile-time specialised kernels where the length has one (PS_TUNE_STATIC, 1) */
    int specialised3; /* ... also for the 3-component z pass (PS_TUNE_STATIC3, 1) */
    int twreg_z;      /* z pass keeps stage twiddles in registers (PS_TUNE_TWREG_Z, 1) */
    int twreg_y;      /* y passes keep stage twiddles in registers; -1 = planner: yes for single
                         precision y transforms of 1024 and longer (PS_TUNE_TWREG_Y, -1) */
    int occ4_z;       /* z pass targets 4 waves per SIMD instead of 2 (PS_TUNE_OCC4_Z, 0) */
    int zwave;        /* pencil-per-wavefront z kernel where the length allows (PS_TUNE_ZWAVE, 1);
                         1 leaves FP64 pencils of 1024 on the fused kernel (2 does not);
                         FP32 z pencils of 2048 (nz = 1024 unbounded): 1 two 1024-point residues,
                         2 one 2048-point pencil per wavefront, 0 neither */
    int halfb_z;      /* z tile-width divisor 0/1/2 -> B, B/2, B/4; -1 = planner: B/2 for z
                         transforms of 512 or shorter, else B (PS_TUNE_HALFB_Z, -1) */
    int halfb_y;      /* y tile-width divisor 0/1 -> B, B/2 (PS_TUNE_HALFB_Y, 0) */
    int mfma;         /* PS_MIXED: matrix-core (FP16 hi/lo DFT) kernels where the length
                         has them: z transforms of 512 and 1024: 1 the persistent
                         register-pencil kernel (one pencil per wavefront, 16-column tiles),
                         2 one tile per workgroup; of 2048: 1 the register-pencil kernel (two
                         wavefronts per pencil, 8-column tiles), 2 eight pencils per
                         workgroup, 3 four; 0 the FP32 kernels (PS_TUNE_MFMA, 1) */
    int t2d;          /* 2D contexts whose fused y pass holds one or two columns per workgroup
                         (long pencils): the x pass output is transposed so that y pencils are
                         contiguous, and grad / curl transform their scalar once (PS_TUNE_T2D, 1) */
    int zfan;         /* potential form (ps_solve3d_potential): 1 runs its z outputs (A, d/dz A,
                         Laplacian) from one forward transform of the scalar where the
                         matrix-core kernel has that form (PS_MIXED; z transforms of 512 and
                         2048, mfma 1), and the y^-1 passes of A, d/dx A and d/dy A from one
                         load of A's z output (FP32 and PS_MIXED, specialised y lengths); 0 one
                         launch per output (PS_TUNE_ZFAN, 1) */
} ps_tuning;
/* the defaults (with PS_TUNE_* overrides applied) */
int ps_tuning_default(ps_tuning* t);
/* tuning used by the next ps_setup3d on this context; NULL returns to the defaults */
int ps_set_tuning(ps_ctx* ctx, const ps_tuning* t);
/* tuning the last ps_setup3d used */
int ps_get_tuning(const ps_ctx* ctx, ps_tuning* t);
/* time candidate tunings on this context's grid, options and pushed components
   (X only when none is pushed), by coordinate descent from the current tuning
   over layout, tile, radix-split and kernel-variant choices, `iters` solves per
   candidate; leaves the context set up with the fastest (pushed data and
   solutions are discarded: push again).  best / best_ms may be NULL. */
int ps_autotune(ps_ctx* ctx, int iters, ps_tuning* best, double* best_ms);

/* --- extension: batches of independent solves of uneven size ---------------
   One context per problem (own plans, tiles and Green's kernel); ps_batch_solve
   replays every problem's passes as one captured hipGraph launch, the problems
   running on parallel branches of the graph.  Options apply to all problems.
   Problems of one grid (ncell, bc, dx, reg_order) keep their sources and
   solutions in one device block; when the operator is plain and their pushed
   components agree they are solved as one multi-component solve (five launches
   for the group), with results identical to solving each alone.  Every batch
   solve solves each problem's pushed rows (a filtered rhs that an earlier
   solve left held is dropped).  Before a replay every member is checked; a
   member set up again, pushed with other components or given other options
   since the capture makes the batch solve eagerly once and capture again. */
typedef struct ps_batch ps_batch;
/* ncell, bc, dx: nprob x 3; reg_order: nprob entries */
int ps_batch_create(ps_batch** out, int nprob, const int* ncell, const int* bc, const double* dx,
                    const int* reg_order, int precision);
/* autotune every distinct (ncell, bc) once with the batch's options and apply the
   result to all problems of that shape; pushed data is discarded (push again) */
int ps_batch_autotune(ps_batch* b, int iters);
int ps_batch_destroy(ps_batch* b);
int ps_batch_set_ops(ps_batch* b, int grad, int div, int curl, int reproject);
int ps_batch_context(ps_batch* b, int prob, ps_ctx** ctx);
int ps_batch_push(ps_batch* b, int prob, const double* bx, const double* by, const double* bz);
int ps_batch_push_device(ps_batch* b, int prob, const void* bx, const void* by, const void* bz);
/* solve every problem: one graph launch (captured on the first call after a push
   pattern or option change), then wait */
int ps_batch_solve(ps_batch* b);
int ps_batch_pull(ps_batch* b, int prob, double* ax, double* ay, double* az, double* bx, double* by, double* bz);
/* run `iters` batch solves, return the average wall time of one graph launch in ms */
int ps_batch_profile(ps_batch* b, int iters, double* ms);
/* problems solved inside group solves by the last solve (0 if none), -1 for NULL */
int ps_batch_grouped(const ps_batch* b);
const char* ps_batch_last_error(const ps_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* PS_CAPI_H */
