// cp_server.hip -- serving loop: continuous batching of independent
// plasticity problems on one GPU.
//
// Requests are admitted (FIFO) while a slot and the HBM budget allow; each
// admitted request keeps its mesh, matrix, state and load level on the
// device and advances through a small state machine
//   level start -> assemble -> CG (or FP64 residual / FP32 CG refinement)
//   -> return map -> next secant iteration / next level / done.
// Block-Jacobi requests share the server's stream and advance in rounds: one
// round enqueues every such request's next device work, then one chunk of
// batched CG launches (a single launch advances one CG iteration of every
// problem in its CG phase, tiles of all problems in one grid, each problem
// with its own scalars and partial sums, so a result does not depend on what
// else is in the batch), then one host synchronisation.  Multigrid requests
// work on streams of their own and advance independently: each one's device
// work ends with a readback of its scalars into a pinned host slot and an
// event, and every step advances the requests whose event has completed.
// Their CG iterations (the V-cycles) run in lockstep batches: the requests
// that entered their CG phase in a step advance together on one of the
// batch streams (taken in turn, so that batches overlap), every step of the iteration (each level's smoothing steps, the
// restrictions and prolongations, the one-workgroup tails, the vector
// updates and dots) one launch over the tiles of all of them, so N small
// problems cost the launches of one.  Requests that
// arrive while others run are admitted (FIFO) at the next step as slots and
// memory free up; a worker thread builds an admitted request's device problem
// (CSR pattern, upload, multigrid lattice) while the steps go on advancing
// the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cp_problem.h"

namespace {

enum Phase { PH_LEVEL, PH_CG, PH_WAIT_RES, PH_WAIT_RM, PH_DONE };
enum State { RQ_QUEUED, RQ_ACTIVE, RQ_DONE, RQ_FAILED };

struct Request {
    long long id = 0;
    int nn = 0, ne = 0;
    std::vector<double> x, y, rp, u0;
    std::vector<int> conn;
    std::vector<unsigned char> mask;
    int n1 = 0, n2 = 0;              // multigrid lattice (precond 2)
    int per1 = 0;                    // the lattice is periodic in direction 1
    std::vector<int> node_of;        // its node numbering (empty: identity)
    cp_material mat{};
    int nsteps = 1;
    double final_factor = 1.;
    cp_options opt{};
    size_t bytes = 0;
    State state = RQ_QUEUED;
    cp_problem* P = nullptr;
    hipEvent_t ev = nullptr;         // multigrid requests: their scalars' readback has landed (also
                                     // marks the end of their own stream's work for a batch to wait on)
    int slot = -1;                   // multigrid requests: mirror slot of the readback
    bool pending = false;            // multigrid requests: a readback is in flight
    int phase = PH_LEVEL;
    int level = 0, k = 0, tight = 1;
    double prev_rel = 1., ctol = 0., bnorm2 = 0., itol = 0., rn = 0.;
    cp::Refine refine;               // refinement passes of the current mixed-precision solve
    bool lin_ok = true;              // the current secant iteration's linear solve converged
    std::string lin_msg;             // why it did not
    long long cg_iters = 0;
    int outer_total = 0;
    int status = CP_OK;
    int prep_status = CP_OK;         // result of the worker's preparation
    std::string err;                 // message of the failure that ended the request
    double t_submit = 0., t_admit = -1., t_done = -1.;
    double cg_rho_prev = -1.;  // CG residual at the last observation
    int cg_iters_seen = 0;      // CG iteration count at the last observation
    int next_chunk = 16;        // predicted iterations left in the current CG
    std::vector<double> u_out, sig_out;
};

// Host mirrors of the device scalars of one request, filled each round.
struct Mirror {
    cpk::CgScal cg;
    cpk::RmScal rm;
    cpk::RefScal ref;
};

struct PackItem {
    const char* src;
    int dst_off;
    int nbytes;
};

__global__ void k_pack(const PackItem* items, int n, char* dst)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const PackItem it = items[i];
    for (int b = threadIdx.x; b < it.nbytes; b += blockDim.x) dst[it.dst_off + b] = it.src[b];
}

// The three scalar blocks of one request into its mirror slot (grid of 3).
struct Pack3 {
    PackItem it[3];
};

__global__ void k_pack3(Pack3 p, char* dst)
{
    const PackItem it = p.it[blockIdx.x];
    for (int b = threadIdx.x; b < it.nbytes; b += blockDim.x) dst[it.dst_off + b] = it.src[b];
}

constexpr size_t kMirStride = (sizeof(Mirror) + 255) / 256 * 256;

double clock_now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct cp_server {
    cp_ctx* ctx = nullptr;
    cp_server_config cfg{};
    hipStream_t stream = nullptr;
    double t0 = 0.;
    long long next_id = 1;
    size_t used = 0;
    std::map<long long, std::unique_ptr<Request>> reqs;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    // preparation of admitted requests on the worker thread: to_prep and
    // prepared are shared with it under mu; `preparing` counts both plus the
    // one in the worker's hands (main thread only)
    std::thread worker;
    // held by every step and by the worker's graph captures: a capture must
    // not overlap another thread's device calls (a synchronous copy, an
    // allocation, a free), which would invalidate it
    std::mutex step_mu;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> to_prep, prepared;
    bool stop = false;
    int preparing = 0;
    // finished problems whose memory is not yet returned (a free may wait for
    // the whole device, so frees are batched): problem and budgeted bytes
    std::vector<std::pair<cp_problem*, size_t>> grave;
    // readback slots of the multigrid requests, one per admission slot
    char* mir_dev = nullptr;
    char* mir_host = nullptr;
    std::vector<int> free_slots;
    // lockstep multigrid CG batches on streams of their own, taken in turn
    // so that the batches of successive steps overlap; per stream and
    // precision the recorded launches of its last batch
    static constexpr int kBatchStreams = 4;
    hipStream_t mg_stream[kBatchStreams] = {};
    cp::cp_mg_batch* mgb[kBatchStreams][2] = {};
    int mg_next = 0;
    std::vector<Request*> mg_ready;   // multigrid requests whose CG iterations go into this step's batches
    long long mg_rounds = 0, mg_members = 0, mg_iters = 0;   // batches run, their members, lockstep iterations
    // device tables for the batched CG and the scalar pack
    char* dev = nullptr;
    size_t dev_size = 0;
    char* host = nullptr;  // pinned staging, same layout as dev
    size_t host_size = 0;
    double now() const { return clock_now() - t0; }
};

namespace {

int ensure_buffers(cp_server* S, size_t bytes)
{
    if (bytes <= S->dev_size) return CP_OK;
    size_t n = std::max(bytes, 2 * S->dev_size);
    if (S->dev) (void)hipFree(S->dev);
    if (S->host) (void)hipHostFree(S->host);
    S->dev = nullptr;
    S->host = nullptr;
    CP_HIP(hipMalloc(&S->dev, n));
    CP_HIP(hipHostMalloc((void**)&S->host, n, hipHostMallocDefault));
    S->dev_size = S->host_size = n;
    return CP_OK;
}

// Return the memory of every finished problem to the device and the budget.
void flush_grave(cp_server* S)
{
    for (auto& g : S->grave) {
        cp_problem_destroy(g.first);
        S->used -= g.second;
    }
    S->grave.clear();
}

void release(cp_server* S, Request* r)
{
    if (r->P) {
        S->grave.emplace_back(r->P, r->bytes);
        r->P = nullptr;
    }
    if (r->ev) {
        (void)hipEventDestroy(r->ev);
        r->ev = nullptr;
    }
    if (r->slot >= 0) {
        S->free_slots.push_back(r->slot);
        r->slot = -1;
    }
    r->pending = false;
    if ((int)S->grave.size() >= S->cfg.max_slots) flush_grave(S);
}

int fail_request(cp_server* S, Request* r, int code)
{
    if (r->err.empty()) r->err = cp_last_error();
    r->status = code;
    r->state = RQ_FAILED;
    r->t_done = S->now();
    release(S, r);
    return CP_OK;
}

double level_factor(const Request* r, int level) { return r->final_factor * double(level) / double(r->nsteps); }

// Enqueue the start of secant iteration k of the current level.
int start_iteration(Request* r)
{
    cp_problem* P = r->P;
    // multigrid: a fresh hierarchy at the first secant iteration of a level,
    // as cp_solve does
    if (int s = cp::launch_assemble(P, r->opt.precision, r->k == 0)) return s;
    double ctol = r->opt.cg_tol;
    if (r->opt.cg_forcing > 0. && r->opt.fixed_outer <= 0) ctol = std::max(r->opt.cg_tol, r->opt.cg_forcing * r->prev_rel);
    r->ctol = ctol;
    r->tight = ctol <= r->opt.cg_tol;
    r->bnorm2 = cp::host_bnorm2(P);
    r->lin_ok = true;
    r->lin_msg.clear();
    const int maxit = r->opt.cg_maxit > 0 ? r->opt.cg_maxit : (P->precond == 2 ? 2000 : 1000000000);
    if (r->opt.precision == CP_FP64) {
        if (int s = cp::enqueue_cg_init(P, CP_FP64, r->bnorm2, ctol, maxit)) return s;
        if (P->precond == 2)
            if (int s = cp::mg_cg_enqueue_start(P, CP_FP64)) return s;
        r->phase = PH_CG;
        r->cg_rho_prev = -1.;
        r->cg_iters_seen = 0;
        r->next_chunk = P->precond == 2 ? cp::mg_cg_first_chunk(P, ctol * ctol, false) : 16;
    } else {
        r->refine = cp::Refine();
        r->refine.ref_norm = r->bnorm2;
        if (int s = cp::enqueue_residual(P)) return s;
        r->phase = PH_WAIT_RES;
    }
    return CP_OK;
}

int start_level(Request* r)
{
    cp_problem* P = r->P;
    const double lam = level_factor(r, r->level);
    if (r->level == 1) {
        P->load = lam;
        if (int s = cp::reset_state(P)) return s;
    } else {
        if (int s = cp::advance_load(P, lam, r->opt.predictor)) return s;
    }
    P->precond = r->opt.precond;
    if (int s = cp::enqueue_rm_reset(P)) return s;
    r->k = 0;
    r->prev_rel = 1.;
    return start_iteration(r);
}

int max_outer_of(const Request* r)
{
    return r->opt.fixed_outer > 0 ? r->opt.fixed_outer : std::max(1, r->opt.max_outer);
}

// Use the scalars read back this round; enqueue the request's next work.
int advance(cp_server* S, Request* r, const Mirror& m)
{
    cp_problem* P = r->P;
    switch (r->phase) {
    case PH_CG: {
        const cpk::CgScal& c = m.cg;
        if (!c.done) return CP_OK;
        r->cg_iters += c.iters;
        if (c.status == 2) return fail_request(S, r, cp::fail(CP_ERR_NAN, "CG: NaN in the residual"));
        P->last_cg_iters = c.iters;
        if (P->precond == 2) cp::mg_cg_observe(P, c);
        if (r->opt.precision == CP_FP64) {
            // as cp_solve: a solve that stopped at cg_maxit is not a tight one
            r->lin_ok = c.status == 0;
            if (int s = cp::launch_return_map(P, r->k, r->opt.aitken, r->opt.fixed_outer, max_outer_of(r), 0,
                                              r->opt.tol, r->tight && r->lin_ok))
                return s;
            r->phase = PH_WAIT_RM;
        } else {
            if (int s = cp::enqueue_correction(P)) return s;
            // as cp_solve: an inexact secant iteration's FP32 solve that met
            // the whole reduction ends the refinement
            const double inner = r->opt.inner_tol > 0 ? r->opt.inner_tol : 1e-3;
            if (!r->tight && r->itol > inner && c.status == 0 && c.rho0 > 0. && c.rho <= r->itol * r->itol * c.rho0) {
                if (int s = cp::launch_return_map(P, r->k, r->opt.aitken, r->opt.fixed_outer, max_outer_of(r), 0,
                                                  r->opt.tol, r->tight))
                    return s;
                r->phase = PH_WAIT_RM;
                return CP_OK;
            }
            if (int s = cp::enqueue_residual(P)) return s;
            r->phase = PH_WAIT_RES;
        }
        return CP_OK;
    }
    case PH_WAIT_RES: {
        const double rn = m.ref.rnorm2;
        const double inner = r->opt.inner_tol > 0 ? r->opt.inner_tol : 1e-3;
        double itol = 0.;
        const int act = cp::refine_decide(r->refine, rn, m.ref.mag2, r->ctol, inner, &itol);
        if (act == cp::REF_NAN) return fail_request(S, r, cp::fail(CP_ERR_NAN, "refinement: NaN in the residual"));
        if (act == cp::REF_DONE || act == cp::REF_STALLED) {
            // as cp_solve: a stalled refinement is not a tight solve
            r->lin_ok = act == cp::REF_DONE;
            if (!r->lin_ok)
                r->lin_msg = "refinement stalled at relative residual " + std::to_string(std::sqrt(rn / r->refine.ref_norm));
            if (int s = cp::launch_return_map(P, r->k, r->opt.aitken, r->opt.fixed_outer, max_outer_of(r), 0,
                                              r->opt.tol, r->tight && r->lin_ok))
                return s;
            r->phase = PH_WAIT_RM;
            return CP_OK;
        }
        r->itol = itol;
        r->rn = rn;
        const int maxit = r->opt.cg_maxit > 0 ? r->opt.cg_maxit : (P->precond == 2 ? 2000 : 1000000000);
        if (int s = cp::enqueue_cg_init(P, CP_MIXED, rn, itol, maxit)) return s;
        if (P->precond == 2)
            if (int s = cp::mg_cg_enqueue_start(P, CP_MIXED)) return s;
        r->phase = PH_CG;
        r->cg_rho_prev = -1.;
        r->cg_iters_seen = 0;
        r->next_chunk = P->precond == 2 ? cp::mg_cg_first_chunk(P, itol * itol, true) : 16;
        return CP_OK;
    }
    case PH_WAIT_RM: {
        const cpk::RmScal& h = m.rm;
        ++r->outer_total;
        if (h.status == 2) return fail_request(S, r, cp::fail(CP_ERR_NAN, "return map: NaN in the secant update"));
        if (!h.last) {
            ++r->k;
            r->prev_rel = h.rel_err <= r->opt.tol ? 0. : h.rel_err;
            return start_iteration(r);
        }
        if (!h.converged) return fail_request(S, r, cp::fail(CP_ERR_NOT_CONVERGED, "secant iterations did not converge"));
        if (!r->lin_ok)
            return fail_request(S, r, cp::failf(CP_ERR_NOT_CONVERGED, "the level's last linear solve did not converge (%s)",
                                                r->lin_msg.empty() ? "CG: cg_maxit iterations" : r->lin_msg.c_str()));
        // end of a level, as in cp_solve_steps: plastic strains, then the next level
        if (int s = cp::enqueue_plastic(P)) return s;
        P->solved = true;
        if (r->level < r->nsteps) {
            ++r->level;
            return start_level(r);
        }
        r->phase = PH_DONE;
        return CP_OK;
    }
    default:
        return CP_OK;
    }
}

// Remaining-iteration estimate of a request in its CG phase from the
// observed decay of (r, r) since the last observation; `chunk` is the
// iteration count of the batched round that just ran (block-Jacobi).
void update_chunk(Request* r, const cpk::CgScal& g, int chunk)
{
    const bool mg = r->opt.precond == 2;
    int next = mg ? std::min(64, 2 * r->next_chunk) : std::min(1024, 2 * chunk);
    const double r0 = r->cg_rho_prev > 0. ? r->cg_rho_prev : (mg ? g.rho0 : -1.);
    if (r0 > 0. && g.rho > 0. && g.rho < r0 && g.iters > r->cg_iters_seen) {
        const double target = g.tol2 * g.bnorm2;
        const double rate = std::log(r0 / g.rho) / double(g.iters - r->cg_iters_seen);
        const double rem = std::log(g.rho / target) / rate;
        if (rem > 0 && rem < 1e7)
            next = mg ? (int)std::min(64., std::max(1., std::ceil(rem)))
                      : (int)std::min(1024., std::max(4., rem * 1.05 + 2.));
    }
    r->cg_rho_prev = g.rho;
    r->cg_iters_seen = g.iters;
    r->next_chunk = next;
}

// Readback of a multigrid request's scalars into its mirror slot on stream
// st, and the event that says it has landed.
int enqueue_readback(cp_server* S, Request* r, hipStream_t st)
{
    cp_problem* P = r->P;
    Pack3 pk;
    pk.it[0] = PackItem{(const char*)(r->opt.precision == CP_FP64 ? (const void*)P->sc64 : (const void*)P->sc32),
                        (int)offsetof(Mirror, cg), (int)sizeof(cpk::CgScal)};
    pk.it[1] = PackItem{(const char*)P->rms, (int)offsetof(Mirror, rm), (int)sizeof(cpk::RmScal)};
    pk.it[2] = PackItem{(const char*)P->refs, (int)offsetof(Mirror, ref), (int)sizeof(cpk::RefScal)};
    char* d = S->mir_dev + (size_t)r->slot * kMirStride;
    hipLaunchKernelGGL(k_pack3, dim3(3), dim3(64), 0, st, pk, d);
    CP_HIP(hipGetLastError());
    CP_HIP(hipMemcpyAsync(S->mir_host + (size_t)r->slot * kMirStride, d, sizeof(Mirror), hipMemcpyDeviceToHost, st));
    CP_HIP(hipEventRecord(r->ev, st));
    r->pending = true;
    return CP_OK;
}

// Multigrid request: its CG iterations if it is in the CG phase, then the
// readback of its scalars.  With lockstep batches a request in its CG phase
// waits for this step's batches instead.
int enqueue_mg_round(cp_server* S, Request* r)
{
    cp_problem* P = r->P;
    if (r->phase == PH_CG && S->cfg.batch_mg && (S->cfg.batch_mg_max_nodes <= 0 || r->nn <= S->cfg.batch_mg_max_nodes)) {
        S->mg_ready.push_back(r);
        return CP_OK;
    }
    if (r->phase == PH_CG)
        if (int s = cp::mg_cg_enqueue_iterations(P, r->opt.precision, r->next_chunk)) return s;
    return enqueue_readback(S, r, P->stream);
}

// This step's lockstep batches: per precision, every multigrid request in
// its CG phase runs the chunk of iterations the slowest of them is
// predicted to need (the others' steps return at once once they are done),
// after its own stream's work, then its readback.  A request alone replays
// its own graphs.
void run_mg_batches(cp_server* S)
{
    std::vector<Request*> ready;
    ready.swap(S->mg_ready);
    for (int pi = 0; pi < 2; ++pi) {
        const int prec = pi == 0 ? CP_FP64 : CP_MIXED;
        std::vector<Request*> g;
        for (Request* r : ready)
            if (r->state == RQ_ACTIVE && r->phase == PH_CG && r->opt.precision == prec) g.push_back(r);
        if (g.empty()) continue;
        if (g.size() == 1) {
            Request* r = g[0];
            int s = cp::mg_cg_enqueue_iterations(r->P, prec, r->next_chunk);
            if (s == CP_OK) s = enqueue_readback(S, r, r->P->stream);
            if (s != CP_OK) fail_request(S, r, s);
            continue;
        }
        const int bs = S->mg_next;
        S->mg_next = (S->mg_next + 1) % cp_server::kBatchStreams;
        hipStream_t bst = S->mg_stream[bs];
        int chunk = 1;
        std::vector<cp_problem*> probs;
        std::vector<long long> ids;
        int s = CP_OK;
        for (Request* r : g) {
            chunk = std::max(chunk, r->next_chunk);
            probs.push_back(r->P);
            ids.push_back(r->id);
            if (s == CP_OK && (hipEventRecord(r->ev, r->P->stream) != hipSuccess ||
                               hipStreamWaitEvent(bst, r->ev, 0) != hipSuccess))
                s = cp::fail(CP_ERR_DEVICE, "cp_server: batch dependency failed");
        }
        if (s == CP_OK) s = cp::mg_batch_run(&S->mgb[bs][pi], probs.data(), ids.data(), (int)g.size(), prec, chunk, bst);
        if (s == CP_OK) {
            ++S->mg_rounds;
            S->mg_members += (long long)g.size();
            S->mg_iters += chunk;
        }
        for (Request* r : g) {
            int q = s;
            if (q == CP_OK) q = enqueue_readback(S, r, bst);
            if (q != CP_OK) fail_request(S, r, q);
        }
    }
}

// A request whose work has finished: its results to the host, its memory to
// the deferred frees.
void finish(cp_server* S, Request* r)
{
    r->u_out.resize(2 * (size_t)r->nn);
    r->sig_out.resize(r->ne);
    if (cp_download(r->P, CP_F_U, r->u_out.data()) || cp_download(r->P, CP_F_SIGI, r->sig_out.data())) {
        fail_request(S, r, CP_ERR_DEVICE);
    } else {
        r->state = RQ_DONE;
        r->t_done = S->now();
        release(S, r);
    }
}

// Worker thread: the device problem of an admitted request (host CSR
// pattern and incidences, upload, solver vectors, multigrid lattice and
// hierarchy); the request's host copies of its mesh are dropped after.
void prepare(cp_server* S, Request* r)
{
    const bool mg = r->opt.precond == 2;
    cp_problem* P = nullptr;
    // a multigrid request works on a stream of its own, so the V-cycles of
    // several requests (small problems leave most of the GPU idle) run side
    // by side; block-Jacobi requests share the server's stream with the
    // batched CG kernels
    int s = cp::problem_create(S->ctx, r->nn, r->ne, r->x.data(), r->y.data(), r->conn.data(), r->rp.data(),
                               r->u0.data(), r->mask.data(), &r->mat, mg ? nullptr : S->stream, &P);
    r->P = P;   // from here on the caller destroys it if anything fails
    if (s == CP_OK && mg && hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess)
        s = cp::fail(CP_ERR_DEVICE, "cp_server: event creation failed");
    if (s == CP_OK) s = cp::problem_alloc_solver(P, r->opt.precision);
    if (s == CP_OK && mg)
        s = (r->per1 ? cp_problem_set_grid_periodic : cp_problem_set_grid)(P, r->n1, r->n2,
                                                                          r->node_of.empty() ? nullptr : r->node_of.data());
    if (s == CP_OK && mg) {
        std::lock_guard<std::mutex> lk(S->step_mu);
        s = cp::mg_prepare(P, r->opt.precision);
    }
    if (s != CP_OK && r->P) {
        cp_problem_destroy(r->P);
        r->P = nullptr;
    }
    r->prep_status = s;
    if (s != CP_OK) r->err = cp_last_error();
    std::vector<double>().swap(r->x);
    std::vector<double>().swap(r->y);
    std::vector<int>().swap(r->conn);
}

void worker_main(cp_server* S)
{
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [S] { return S->stop || !S->to_prep.empty(); });
            if (S->stop) return;
            r = S->to_prep.front();
            S->to_prep.pop_front();
        }
        try {
            prepare(S, r);
        } catch (...) {   // host memory for the pattern or the hierarchy
            if (r->P) {
                cp_problem_destroy(r->P);
                r->P = nullptr;
            }
            r->prep_status = cp::fail(CP_ERR_NOMEM, "cp_server: preparation ran out of host memory");
            r->err = cp_last_error();
        }
        std::lock_guard<std::mutex> lk(S->mu);
        S->prepared.push_back(r);
    }
}

int admit(cp_server* S)
{
    // prepared requests start their first level
    std::deque<Request*> ready;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        ready.swap(S->prepared);
    }
    for (Request* r : ready) {
        --S->preparing;
        if (r->prep_status != CP_OK) {
            S->used -= r->bytes;
            fail_request(S, r, r->prep_status);
            continue;
        }
        r->state = RQ_ACTIVE;
        r->level = 1;
        int q = start_level(r);
        if (q == CP_OK && r->ev) q = enqueue_mg_round(S, r);
        if (q != CP_OK) {
            fail_request(S, r, q);
            continue;
        }
        S->active.push_back(r);
    }
    // FIFO: hand the next requests to the worker while a slot and the
    // budget allow
    while (!S->queue.empty() && (int)S->active.size() + S->preparing < S->cfg.max_slots) {
        Request* r = S->queue.front();
        if (r->bytes > S->cfg.hbm_budget) {
            S->queue.pop_front();
            fail_request(S, r, cp::fail(CP_ERR_BUDGET, "cp_server: the request needs more than the HBM budget"));
            continue;
        }
        if (S->used + r->bytes > S->cfg.hbm_budget) {
            if (S->grave.empty()) break;
            flush_grave(S);
            continue;
        }
        S->queue.pop_front();
        S->used += r->bytes;
        if (r->opt.precond == 2) {
            r->slot = S->free_slots.back();
            S->free_slots.pop_back();
        }
        r->t_admit = S->now();
        ++S->preparing;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->to_prep.push_back(r);
        }
        S->cv.notify_one();
    }
    return CP_OK;
}

// Batched CG over the requests currently in their CG phase.
template <typename T>
struct Group {
    std::vector<Request*> reqs;
    int pre = 0, maxw = 0;
};

template <int W, bool PRE, typename T>
void launch_batch_w(const cpk::CgArgs<T>* d_args, const int2* d_tiles, int ntiles, int nprob, int chunk, hipStream_t st)
{
    for (int j = 0; j < chunk; ++j) {
        hipLaunchKernelGGL((cpk::k_cg_iter_batched<W, PRE, T>), dim3(ntiles), dim3(cpk::kBlock), 0, st, d_args, d_tiles,
                           ntiles);
        hipLaunchKernelGGL((cpk::k_cg_step_batched<T>), dim3(nprob), dim3(cpk::kBlock), 0, st, d_args, nprob);
    }
}

template <typename T>
void launch_batch(int maxw, int pre, const cpk::CgArgs<T>* d_args, const int2* d_tiles, int ntiles, int nprob,
                  int chunk, hipStream_t st)
{
    if (pre) {
        if (maxw <= 9) launch_batch_w<9, true, T>(d_args, d_tiles, ntiles, nprob, chunk, st);
        else launch_batch_w<0, true, T>(d_args, d_tiles, ntiles, nprob, chunk, st);
    } else {
        if (maxw <= 9) launch_batch_w<9, false, T>(d_args, d_tiles, ntiles, nprob, chunk, st);
        else launch_batch_w<0, false, T>(d_args, d_tiles, ntiles, nprob, chunk, st);
    }
}

}  // namespace

extern "C" {

int cp_server_create(cp_ctx* ctx, const cp_server_config* cfg, cp_server** out)
{
    if (!out) return cp::fail(CP_ERR_INVALID, "null out");
    *out = nullptr;
    if (!ctx || !cfg || cfg->max_slots < 1 || cfg->hbm_budget == 0)
        return cp::fail(CP_ERR_INVALID, "cp_server_create: bad configuration");
    CP_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<cp_server> S(new cp_server());
    S->ctx = ctx;
    S->cfg = *cfg;
    S->t0 = clock_now();
    CP_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    for (hipStream_t& s : S->mg_stream) CP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (int s = ensure_buffers(S.get(), 1 << 20)) return s;
    CP_HIP(hipMalloc(&S->mir_dev, kMirStride * cfg->max_slots));
    CP_HIP(hipHostMalloc((void**)&S->mir_host, kMirStride * cfg->max_slots, hipHostMallocDefault));
    for (int i = cfg->max_slots - 1; i >= 0; --i) S->free_slots.push_back(i);
    try {
        S->worker = std::thread(worker_main, S.get());
    } catch (...) {
        return cp::fail(CP_ERR_DEVICE, "cp_server_create: cannot start the preparation thread");
    }
    *out = S.release();
    return CP_OK;
}

int cp_server_destroy(cp_server* S)
{
    if (!S) return CP_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipStreamSynchronize(S->stream);
    for (int i = 0; i < cp_server::kBatchStreams; ++i) {
        (void)hipStreamSynchronize(S->mg_stream[i]);
        for (cp::cp_mg_batch*& b : S->mgb[i]) {
            cp::mg_batch_destroy(b);
            b = nullptr;
        }
    }
    for (auto& kv : S->reqs) {
        if (kv.second->P) cp_problem_destroy(kv.second->P);   // waits for its stream
        if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
    }
    flush_grave(S);
    if (S->dev) (void)hipFree(S->dev);
    if (S->host) (void)hipHostFree(S->host);
    if (S->mir_dev) (void)hipFree(S->mir_dev);
    if (S->mir_host) (void)hipHostFree(S->mir_host);
    (void)hipStreamDestroy(S->stream);
    for (hipStream_t s : S->mg_stream) (void)hipStreamDestroy(s);
    delete S;
    return CP_OK;
}

int cp_server_batch_stats(cp_server* S, long long* rounds, long long* members, long long* iters)
{
    if (!S) return cp::fail(CP_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> step_lock(S->step_mu);
    if (rounds) *rounds = S->mg_rounds;
    if (members) *members = S->mg_members;
    if (iters) *iters = S->mg_iters;
    return CP_OK;
}

int cp_server_clock(cp_server* S, double* now)
{
    if (!S || !now) return cp::fail(CP_ERR_INVALID, "null argument");
    *now = S->now();
    return CP_OK;
}

int cp_server_submit(cp_server* S, int nn, int ne, const double* x, const double* y, const int* conn, const double* rp,
                     const double* u0, const unsigned char* mask, const cp_material* mat, int nsteps,
                     double final_factor, const cp_options* opt, long long* id)
{
    if (opt && opt->precond == 2)
        return cp::fail(CP_ERR_INVALID, "cp_server_submit: multigrid (precond 2) needs the lattice: cp_server_submit_grid");
    return cp_server_submit_grid(S, nn, ne, x, y, conn, rp, u0, mask, mat, nsteps, final_factor, opt, 0, 0, nullptr, id);
}

namespace {

int submit_grid(cp_server* S, int nn, int ne, const double* x, const double* y, const int* conn, const double* rp,
                const double* u0, const unsigned char* mask, const cp_material* mat, int nsteps, double final_factor,
                const cp_options* opt, int n1, int n2, const int* node_of, int per1, long long* id)
{
    if (!S || !x || !y || !conn || !rp || !mask || !mat || !opt || nsteps < 1 || nn < 1 || ne < 1)
        return cp::fail(CP_ERR_INVALID, "cp_server_submit: bad arguments");
    if (opt->precision != CP_FP64 && opt->precision != CP_MIXED) return cp::fail(CP_ERR_INVALID, "bad precision");
    if (opt->precond < 0 || opt->precond > 2) return cp::fail(CP_ERR_INVALID, "cp_server_submit: precond must be 0, 1 or 2");
    if (opt->precond == 2 && (n1 < 2 || n2 < 2 || (long long)n1 * n2 != nn))
        return cp::failf(CP_ERR_INVALID, "cp_server_submit_grid: %d x %d lattice for %d nodes", n1, n2, nn);
    std::unique_ptr<Request> r(new Request());
    if (opt->precond == 2) {
        r->n1 = n1;
        r->n2 = n2;
        r->per1 = per1;
        if (node_of) r->node_of.assign(node_of, node_of + nn);
    }
    r->id = S->next_id++;
    r->nn = nn;
    r->ne = ne;
    r->x.assign(x, x + nn);
    r->y.assign(y, y + nn);
    r->conn.assign(conn, conn + 4 * (size_t)ne);
    r->rp.assign(rp, rp + 2 * (size_t)nn);
    r->u0.assign(2 * (size_t)nn, 0.);
    if (u0) r->u0.assign(u0, u0 + 2 * (size_t)nn);
    r->mask.assign(mask, mask + 2 * (size_t)nn);
    r->mat = *mat;
    r->nsteps = nsteps;
    r->final_factor = final_factor;
    r->opt = *opt;
    // matrix blocks of a quad mesh ~ 9 per node: budget with the layout estimate
    size_t bytes = 0;
    if (int s = cp_problem_bytes(nn, ne, 4LL * 9 * nn, opt->precision, &bytes)) return s;
    if (opt->precond == 2) bytes += cp::mg_bytes_estimate(nn, opt->precision);
    r->bytes = bytes;
    r->t_submit = S->now();
    Request* raw = r.get();
    S->reqs[raw->id] = std::move(r);
    S->queue.push_back(raw);
    if (id) *id = raw->id;
    return CP_OK;
}

}  // namespace

int cp_server_submit_grid(cp_server* S, int nn, int ne, const double* x, const double* y, const int* conn,
                          const double* rp, const double* u0, const unsigned char* mask, const cp_material* mat,
                          int nsteps, double final_factor, const cp_options* opt, int n1, int n2, const int* node_of,
                          long long* id)
{
    return submit_grid(S, nn, ne, x, y, conn, rp, u0, mask, mat, nsteps, final_factor, opt, n1, n2, node_of, 0, id);
}

int cp_server_submit_grid_periodic(cp_server* S, int nn, int ne, const double* x, const double* y, const int* conn,
                                   const double* rp, const double* u0, const unsigned char* mask,
                                   const cp_material* mat, int nsteps, double final_factor, const cp_options* opt,
                                   int n1, int n2, const int* node_of, long long* id)
{
    return submit_grid(S, nn, ne, x, y, conn, rp, u0, mask, mat, nsteps, final_factor, opt, n1, n2, node_of, 1, id);
}

int cp_server_step(cp_server* S, int* in_flight)
{
    if (!S) return cp::fail(CP_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> step_lock(S->step_mu);
    if (int s = admit(S)) return s;
    hipStream_t st = S->stream;
    // multigrid requests: advance every one whose last readback has landed
    for (Request* r : S->active) {
        if (!r->pending) continue;
        const hipError_t q = hipEventQuery(r->ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) {
            fail_request(S, r, cp::failf(CP_ERR_DEVICE, "cp_server: event query: %s", hipGetErrorString(q)));
            continue;
        }
        r->pending = false;
        const Mirror m = *reinterpret_cast<const Mirror*>(S->mir_host + (size_t)r->slot * kMirStride);
        if (r->phase == PH_CG) update_chunk(r, m.cg, 0);
        if (int s = advance(S, r, m)) fail_request(S, r, s);
        if (r->state != RQ_ACTIVE) continue;
        if (r->phase == PH_DONE) finish(S, r);
        else if (int s = enqueue_mg_round(S, r)) fail_request(S, r, s);
    }
    run_mg_batches(S);
    // block-Jacobi requests: one batched round on the server's stream
    std::vector<Request*> act;
    for (Request* r : S->active)
        if (r->state == RQ_ACTIVE && !r->ev) act.push_back(r);
    if (!act.empty()) {
        // requests in their CG phase, grouped by (precision, preconditioner)
        std::vector<Request*> g64[2], g32[2];
        for (Request* r : act) {
            if (r->phase != PH_CG) continue;
            (r->opt.precision == CP_FP64 ? g64 : g32)[r->opt.precond ? 1 : 0].push_back(r);
        }
        // device tables: [CgArgs groups][tiles][pack items] then the pack output
        size_t off = 0;
        auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
        struct Launch {
            int prec, pre, maxw, ntiles, nprob;
            size_t args_off, tiles_off;
        };
        std::vector<Launch> launches;
        size_t need = 0;
        for (int p = 0; p < 2; ++p)
            for (int pre = 0; pre < 2; ++pre) {
                auto& g = p == 0 ? g64[pre] : g32[pre];
                if (g.empty()) continue;
                int nt = 0;
                for (Request* r : g) nt += r->P->nbn;
                need += align(g.size() * sizeof(cpk::CgArgs<double>)) + align(nt * sizeof(int2));
            }
        need += align(act.size() * 3 * sizeof(PackItem)) + align(act.size() * sizeof(Mirror));
        if (int s = ensure_buffers(S, need)) return s;
        for (int p = 0; p < 2; ++p)
            for (int pre = 0; pre < 2; ++pre) {
                auto& g = p == 0 ? g64[pre] : g32[pre];
                if (g.empty()) continue;
                Launch L{p, pre, 0, 0, (int)g.size(), off, 0};
                for (size_t i = 0; i < g.size(); ++i) {
                    if (p == 0) reinterpret_cast<cpk::CgArgs<double>*>(S->host + off)[i] = g[i]->P->cg64_args();
                    else reinterpret_cast<cpk::CgArgs<float>*>(S->host + off)[i] = g[i]->P->cg32_args();
                    L.maxw = std::max(L.maxw, g[i]->P->max_w);
                }
                off += align(g.size() * (p == 0 ? sizeof(cpk::CgArgs<double>) : sizeof(cpk::CgArgs<float>)));
                L.tiles_off = off;
                int2* tiles = reinterpret_cast<int2*>(S->host + off);
                for (size_t i = 0; i < g.size(); ++i)
                    for (int b = 0; b < g[i]->P->nbn; ++b) tiles[L.ntiles++] = make_int2((int)i, b);
                off += align(L.ntiles * sizeof(int2));
                launches.push_back(L);
            }
        const size_t pack_off = off;
        PackItem* items = reinterpret_cast<PackItem*>(S->host + pack_off);
        off += align(act.size() * 3 * sizeof(PackItem));
        const size_t mir_off = off;
        int ni = 0;
        for (size_t i = 0; i < act.size(); ++i) {
            cp_problem* P = act[i]->P;
            const int base = (int)(i * sizeof(Mirror));
            const char* cg = (const char*)(act[i]->opt.precision == CP_FP64 ? P->sc64 : P->sc32);
            items[ni++] = PackItem{cg, base + (int)offsetof(Mirror, cg), (int)sizeof(cpk::CgScal)};
            items[ni++] = PackItem{(const char*)P->rms, base + (int)offsetof(Mirror, rm), (int)sizeof(cpk::RmScal)};
            items[ni++] = PackItem{(const char*)P->refs, base + (int)offsetof(Mirror, ref), (int)sizeof(cpk::RefScal)};
        }
        off += align(act.size() * sizeof(Mirror));
        CP_HIP(hipMemcpyAsync(S->dev, S->host, pack_off + align(ni * sizeof(PackItem)), hipMemcpyHostToDevice, st));

        // one chunk of CG iterations for every request in its CG phase: as
        // many as the member closest to convergence is predicted to need
        int chunk = 0;
        for (Request* r : act)
            if (r->phase == PH_CG) chunk = chunk == 0 ? r->next_chunk : std::min(chunk, r->next_chunk);
        chunk = std::max(4, chunk == 0 ? 16 : chunk);
        for (const Launch& L : launches) {
            if (S->cfg.batch_cg) {
                if (L.prec == 0)
                    launch_batch<double>(L.maxw, L.pre, reinterpret_cast<const cpk::CgArgs<double>*>(S->dev + L.args_off),
                                         reinterpret_cast<const int2*>(S->dev + L.tiles_off), L.ntiles, L.nprob, chunk, st);
                else
                    launch_batch<float>(L.maxw, L.pre, reinterpret_cast<const cpk::CgArgs<float>*>(S->dev + L.args_off),
                                        reinterpret_cast<const int2*>(S->dev + L.tiles_off), L.ntiles, L.nprob, chunk, st);
            } else {
                auto& g = L.prec == 0 ? g64[L.pre] : g32[L.pre];
                for (Request* r : g) {
                    cp_problem* P = r->P;
                    if (L.prec == 0) cp::launch_cg_iters<double>(P->max_w, L.pre, P->cg64_args(), chunk, dim3(P->nbn), st);
                    else cp::launch_cg_iters<float>(P->max_w, L.pre, P->cg32_args(), chunk, dim3(P->nbn), st);
                }
            }
        }
        CP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_pack, dim3(ni), dim3(64), 0, st, reinterpret_cast<const PackItem*>(S->dev + pack_off), ni,
                           S->dev + mir_off);
        CP_HIP(hipGetLastError());
        CP_HIP(hipMemcpyAsync(S->host + mir_off, S->dev + mir_off, act.size() * sizeof(Mirror), hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
        const Mirror* mir = reinterpret_cast<const Mirror*>(S->host + mir_off);

        // advance every request with this round's scalars
        std::vector<Mirror> mcopy(mir, mir + act.size());
        for (size_t i = 0; i < act.size(); ++i) {
            Request* r = act[i];
            if (r->phase == PH_CG) update_chunk(r, mcopy[i].cg, chunk);
            if (int s = advance(S, r, mcopy[i])) fail_request(S, r, s);
            if (r->state == RQ_ACTIVE && r->phase == PH_DONE) finish(S, r);
        }
    }
    std::vector<Request*> keep;
    for (Request* r : S->active)
        if (r->state == RQ_ACTIVE) keep.push_back(r);
    S->active.swap(keep);
    // idle: return the finished problems' memory now
    if (S->active.empty() && S->queue.empty() && S->preparing == 0 && !S->grave.empty()) flush_grave(S);
    if (in_flight) *in_flight = (int)(S->active.size() + S->queue.size()) + S->preparing;
    return CP_OK;
}

int cp_server_result(cp_server* S, long long id, cp_request_info* info, double* u, double* sig_i)
{
    if (!S) return cp::fail(CP_ERR_INVALID, "null server");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return cp::fail(CP_ERR_INVALID, "unknown request id");
    Request* r = it->second.get();
    if (info) {
        info->id = r->id;
        info->status = r->state == RQ_DONE ? CP_OK : (r->state == RQ_FAILED ? r->status : CP_ERR_STATE);
        info->steps_done = r->state == RQ_DONE ? r->nsteps : std::max(0, r->level - 1);
        info->cg_iters = r->cg_iters;
        info->outer_iters = r->outer_total;
        info->t_submit = r->t_submit;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
    }
    if (r->state == RQ_FAILED) return cp::failf(r->status, "request %lld failed: %s", id, r->err.c_str());
    if (r->state != RQ_DONE) return cp::fail(CP_ERR_STATE, "request not finished");
    if (u) std::memcpy(u, r->u_out.data(), sizeof(double) * r->u_out.size());
    if (sig_i) std::memcpy(sig_i, r->sig_out.data(), sizeof(double) * r->sig_out.size());
    return CP_OK;
}

}  // extern "C"
