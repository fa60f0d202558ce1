// cp_problem.h -- host-side view of one device-resident plasticity problem.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "cp_abi.h"
#include "cp_internal.h"
#include "cp_kernels.h"

#define CP_HIP(expr)                                                                                         \
    do {                                                                                                     \
        hipError_t _e = (expr);                                                                              \
        if (_e != hipSuccess) return cp::failf(CP_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

struct cp_ctx {
    int device;
    hipDeviceProp_t prop;
    // device address of this GPU's copy of the library's static staging
    // array (cp_internal_force_noalloc), from hipGetSymbolAddress
    void* scratch = nullptr;
};

namespace cp {

// Bump allocator over one hipMalloc (256-byte aligned pieces).
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
    template <typename T>
    T* take(size_t count)
    {
        size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        return p;
    }
};

// Device memory groups of a problem: common data, the FP64 solve and the
// mixed-precision solve (allocated on first use).
struct Layout {
    size_t common, fp64, mixed;
};

Layout problem_layout(int nn, int ne, long long nblk_stored, long long ninc, int nslices, int nbmax, int npe = 4);

struct MgState;   // multigrid hierarchy (cp_mg.hip)
struct cp_mg_batch;   // lockstep launches of several problems (cp_mg.hip)

}  // namespace cp

struct cp_problem {
    cp_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    int nn = 0, ne = 0, nslices = 0, max_w = 0;
    int npe = 4;           // nodes per element: 4 (bilinear) or 8 (8-node quadratic)
    int nbn = 0, nbe = 0;  // blocks over nodes / elements
    long long nblk = 0, nblk_stored = 0;
    cpk::MatParams M{};
    double load = 1.0;
    std::vector<double> rp0;   // host copy for ||b||^2
    // ||b||^2 at load factor bn_load (one O(nn) host pass per load level,
    // not one per linear solve)
    mutable double bn_load = -1., bn_val = 0.;
    std::vector<int> adj_ptr, adj_node;
    std::vector<int> conn_host;   // 8-node elements: the connectivity (multigrid's corner-lattice transfer)
    long long last_cg_iters = 0;
    double mg_lrate = 0.;   // multigrid CG: mean log reduction of (r, r) per iteration (0: none yet)
    double last_cg_ms_per_iter = 0.;
    bool have_fp64_values = false;
    int precond = 0;          // current solve: 0 plain CG, 1 block-Jacobi, 2 multigrid
    cp::MgState* mg = nullptr;  // set by cp_problem_set_grid
    bool solved = false;      // u / psi hold a converged level at `load`
    double lam_prev = 0.;     // load factor of u_prev / psi_prev

    cp::Arena common, a64, amx, aq8;
    // 8-node elements: element blocks and their gather lists (cp_q8.h)
    int* q8_inc = nullptr;      // node -> 8 e + i, element order
    int *q8_bptr = nullptr, *q8_bsrc = nullptr, *q8_bdst = nullptr;
    double4* q8_keb = nullptr;
    double2* q8_fe = nullptr;
    // common
    double2 *xy = nullptr, *ecnu = nullptr, *rp0d = nullptr, *u0d = nullptr, *b = nullptr, *u = nullptr,
            *u_prev = nullptr;
    double* psi_prev = nullptr;
    int4* conn = nullptr;
    uchar2* mask = nullptr;
    double *psi = nullptr, *F = nullptr, *G = nullptr, *rcur = nullptr, *rprev = nullptr, *partials = nullptr,
           *partials_alt = nullptr;
    int *inc_ptr = nullptr, *col = nullptr;
    int2 *inc = nullptr, *slice = nullptr;
    cpk::CgScal *sc64 = nullptr, *sc32 = nullptr;
    cpk::RmScal* rms = nullptr;
    cpk::RefScal* refs = nullptr;
    // FP64 solve
    double2 *va = nullptr, *vb = nullptr, *p64 = nullptr, *r64[2] = {}, *w64[2] = {}, *s64[2] = {};
    double4* minv64 = nullptr;
    // mixed solve
    float4* vf = nullptr;
    float4* minv32 = nullptr;
    float2 *x32 = nullptr, *p32 = nullptr, *r32[2] = {}, *w32[2] = {}, *s32[2] = {}, *b32 = nullptr;
    double2* res64 = nullptr;
    // pinned host mirrors of the device scalars
    cpk::CgScal* h_cg = nullptr;
    cpk::RmScal* h_rm = nullptr;
    cpk::RefScal* h_ref = nullptr;
    hipEvent_t ev[8] = {};

    cpk::CgArgs<double> cg64_args() const;
    cpk::CgArgs<float> cg32_args() const;
};

namespace cp {
int problem_alloc_solver(cp_problem* P, int precision);
int problem_create(cp_ctx* ctx, int nn, int ne, const double* x, const double* y, const int* conn,
                   const double* rp, const double* u0, const unsigned char* mask, const cp_material* mat,
                   hipStream_t stream, cp_problem** out, int npe = 4);
int solve_level(cp_problem* P, const cp_options* opt, int elastic, cp_stats* st);
// mg_full (multigrid only): 0 lets the coarse levels be reused (see mg_setup).
int launch_assemble(cp_problem* P, int precision, int mg_full = 1);
double host_bnorm2(const cp_problem* P);
int set_load(cp_problem* P, double factor, int reset);
int advance_load(cp_problem* P, double factor, int predictor);
int reset_state(cp_problem* P);
cpk::ResArgs residual_args(cp_problem* P);
// r = mask (b - f(u)) (b NULL: f(u)) in FP64 with the element kernels of the
// problem's element type; partials for ||r||^2 when ra.sc is set.
int launch_residual(cp_problem* P, const cpk::ResArgs& ra, hipStream_t st);
int enqueue_cg_init(cp_problem* P, int precision, double bnorm2, double tol, int maxit);
int enqueue_residual(cp_problem* P);

// One pass of the FP64 iterative refinement of the mixed-precision solve,
// decided from the FP64 residual of the current u (the decision of cg_mixed
// and of the serving loop's refinement phase, so that both take the same
// passes).
struct Refine {
    double ref_norm = 0.;   // ||r_0||^2 of the solve (||b||^2, or the first residual when b = 0)
    double prev = -1.;      // ||r||^2 of the previous pass
    int stalled = 0;        // passes that reduced ||r||^2 by less than 4x
};
enum { REF_DONE = 0, REF_SOLVE = 1, REF_STALLED = 2, REF_NAN = 3 };
// rn = ||r||^2, mag2 = ||(|K| |u| + |b|)||^2 over free DOFs (0 if unknown).
// REF_DONE: the residual meets tol, or lies at the rounding floor of its own
// FP64 evaluation (kRefFloor eps64 ||(|K| |u| + |b|)||), below which no pass
// can be verified; REF_SOLVE: run one FP32 solve asking for *itol;
// REF_STALLED: three passes without progress above that floor.
constexpr double kRefFloor = 16.;
int refine_decide(Refine& R, double rn, double mag2, double tol, double inner_tol, double* itol);
int enqueue_correction(cp_problem* P);
int enqueue_plastic(cp_problem* P);
int enqueue_rm_reset(cp_problem* P);
int launch_return_map(cp_problem* P, int k, int aitken, int fixed_outer, int max_outer, int elastic, double tol,
                      int solved_tight);
template <typename T>
void launch_cg_iters(int max_w, int pre, const cpk::CgArgs<T>& a, int count, dim3 g, hipStream_t st);
// Multigrid set-up after an assembly: the finest level's smoother always;
// the coarse levels (Galerkin operators, smoothers, coarsest inverse) when
// full is set, on the first set-up of a precision, or every
// MgState::rebuild-th set-up, else the previous ones are kept.
int mg_setup(cp_problem* P, int precision, int full);
// The FP32 assembly of a set-up that keeps the coarse levels, straight into
// level 0's symmetric lattice layout (k_assemble_sym) where that layout
// exists and level 0 smooths by segments: *done = 1 when launched (the
// set-up then skips its own conversion to that layout), else 0 and the
// caller assembles the full rows.
int mg_assemble_direct(cp_problem* P, int precision, int full, const cpk::AsmArgs& a, int* done);
void mg_destroy(cp_problem* P);
// zero_start: x_0 = 0, so (r_0, r_0) = bnorm2 (the refinement passes).
template <typename T>
int run_mg_cg(cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc, double tol2, double bnorm2,
              long long* iters, int* status, double* ms, bool zero_start = false);
// The same solve in pieces, for callers that drive it without waiting
// (the serving loop): enqueue the start (after enqueue_cg_init), enqueue n
// iterations, the predicted first chunk, and the bookkeeping once done.
int mg_cg_enqueue_start(cp_problem* P, int precision);
// The buffers of `precision` and the two CG graphs (start, iteration) ahead
// of the first solve, so that no later call captures a graph (the serving
// loop prepares problems on a worker thread, where a capture cannot overlap
// another thread's device calls).
int mg_prepare(cp_problem* P, int precision);
int mg_cg_enqueue_iterations(cp_problem* P, int precision, int n);
int mg_cg_first_chunk(const cp_problem* P, double tol2, bool zero_start);
// Upper bound of the device memory cp_problem_set_grid and the multigrid
// solves add to a problem of nn nodes (the coarse levels hold at most nn
// nodes together).
size_t mg_bytes_estimate(int nn, int precision);
void mg_cg_observe(cp_problem* P, const cpk::CgScal& h);
// Lockstep CG iterations of several prepared multigrid problems of one
// precision (the serving loop): `iters` iterations of every problem, each
// step of the iteration one launch over the tiles of all of them, enqueued
// on st.  *batch caches the recorded launches while the problems stay the
// same, as named by ids (names never reused for another problem: a freed
// problem's address may come back); created on first use, freed by
// mg_batch_destroy.
int mg_batch_run(cp_mg_batch** batch, cp_problem* const* probs, const long long* ids, int n, int precision, int iters,
                 hipStream_t st);
void mg_batch_destroy(cp_mg_batch* batch);
int mg_batch_launches(const cp_mg_batch* batch);
}  // namespace cp
