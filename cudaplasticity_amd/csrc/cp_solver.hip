// cp_solver.hip -- device problems and the secant (variable elasticity
// parameters) loop of PlasticitySolver::solveCUDA (PlasticitySolver.cu:469-512)
// on gfx950: node-centric assembly into 2x2-block sliced ELL, fused
// single-launch CG (FP64, or FP32 with FP64 iterative refinement), element
// return map with Aitken acceleration.  No host round trip inside a CG
// iteration; the host polls a device flag between chunks of launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "cp_problem.h"
#include "cp_q8.h"

namespace {
// Device staging of cp_internal_force_noalloc: the internal force of a
// problem of up to kScratchBytes / 16 nodes.  Each GPU holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (cp_ctx::scratch).
constexpr size_t kScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int blocks_for(long long n) { return (int)((n + cpk::kBlock - 1) / cpk::kBlock); }
}  // namespace

namespace cp {

Layout problem_layout(int nn, int ne, long long nblk_stored, long long ninc, int nslices, int nbmax, int npe)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    Layout L{};
    L.common = al(sizeof(double2) * nn) * 6 +      // xy, rp0, u0, b, u, u_prev
               al(sizeof(uchar2) * nn) +
               al(sizeof(int4) * (size_t)ne * (npe / 4)) + al(sizeof(double2) * ne) +
               al(sizeof(double) * ne) * (5 + CP_NFIELDS) +   // psi, psi_prev, G, rcur, rprev, fields
               al(sizeof(int) * (nn + 1)) + al(sizeof(int2) * ninc) +
               al(sizeof(int) * nblk_stored) + al(sizeof(int2) * nslices) +
               al(sizeof(double) * 5 * (size_t)nbmax) + al(sizeof(double) * 3 * (size_t)nbmax) +
               4 * 256;                                        // scalars
    L.fp64 = al(sizeof(double2) * nblk_stored) * 2 + al(sizeof(double2) * nn) * 7 + al(sizeof(double4) * nn);
    L.mixed = al(sizeof(float4) * nblk_stored) + al(sizeof(float2) * nn) * 9 + al(sizeof(double2) * nn) +
              al(sizeof(float4) * nn);
    return L;
}

}  // namespace cp

cpk::CgArgs<double> cp_problem::cg64_args() const
{
    cpk::CgArgs<double> a{};
    a.col = col;
    a.slice = slice;
    a.va = va;
    a.vb = vb;
    a.x = u;
    a.p = p64;
    for (int i = 0; i < 2; ++i) {
        a.r[i] = r64[i];
        a.w[i] = w64[i];
        a.s[i] = s64[i];
    }
    a.b = b;
    a.mask = mask;
    a.minv = minv64;
    a.sc = sc64;
    a.partials = partials;
    a.partials_alt = partials_alt;
    a.nn = nn;
    a.nblocks = nbn;
    return a;
}

cpk::CgArgs<float> cp_problem::cg32_args() const
{
    cpk::CgArgs<float> a{};
    a.col = col;
    a.slice = slice;
    a.va = vf;
    a.vb = nullptr;
    a.x = x32;
    a.p = p32;
    for (int i = 0; i < 2; ++i) {
        a.r[i] = r32[i];
        a.w[i] = w32[i];
        a.s[i] = s32[i];
    }
    a.b = b32;
    a.mask = mask;
    a.minv = minv32;
    a.sc = sc32;
    a.partials = partials;
    a.partials_alt = partials_alt;
    a.nn = nn;
    a.nblocks = nbn;
    return a;
}

namespace cp {

int problem_alloc_solver(cp_problem* P, int precision)
{
    if (precision == CP_FP64) {
        if (P->a64.base) return CP_OK;
        Layout L = problem_layout(P->nn, P->ne, P->nblk_stored, 0, P->nslices, 0);
        void* mem = nullptr;
        if (hipMalloc(&mem, L.fp64) != hipSuccess) {
            (void)hipGetLastError();
            return failf(CP_ERR_NOMEM, "hipMalloc of %zu bytes (FP64 solver) failed", L.fp64);
        }
        P->a64.base = (char*)mem;
        P->a64.size = L.fp64;
        P->va = P->a64.take<double2>(P->nblk_stored);
        P->vb = P->a64.take<double2>(P->nblk_stored);
        if (P->npe == 8) {   // the gather writes stored blocks only: padding stays zero
            CP_HIP(hipMemsetAsync(P->va, 0, sizeof(double2) * P->nblk_stored, P->stream));
            CP_HIP(hipMemsetAsync(P->vb, 0, sizeof(double2) * P->nblk_stored, P->stream));
        }
        P->p64 = P->a64.take<double2>(P->nn);
        for (int i = 0; i < 2; ++i) {
            P->r64[i] = P->a64.take<double2>(P->nn);
            P->w64[i] = P->a64.take<double2>(P->nn);
            P->s64[i] = P->a64.take<double2>(P->nn);
        }
        P->minv64 = P->a64.take<double4>(P->nn);
        return CP_OK;
    }
    if (precision == CP_MIXED) {
        if (P->amx.base) return CP_OK;
        Layout L = problem_layout(P->nn, P->ne, P->nblk_stored, 0, P->nslices, 0);
        void* mem = nullptr;
        if (hipMalloc(&mem, L.mixed) != hipSuccess) {
            (void)hipGetLastError();
            return failf(CP_ERR_NOMEM, "hipMalloc of %zu bytes (mixed solver) failed", L.mixed);
        }
        P->amx.base = (char*)mem;
        P->amx.size = L.mixed;
        P->vf = P->amx.take<float4>(P->nblk_stored);
        if (P->npe == 8) CP_HIP(hipMemsetAsync(P->vf, 0, sizeof(float4) * P->nblk_stored, P->stream));
        P->x32 = P->amx.take<float2>(P->nn);
        P->p32 = P->amx.take<float2>(P->nn);
        for (int i = 0; i < 2; ++i) {
            P->r32[i] = P->amx.take<float2>(P->nn);
            P->w32[i] = P->amx.take<float2>(P->nn);
            P->s32[i] = P->amx.take<float2>(P->nn);
        }
        P->b32 = P->amx.take<float2>(P->nn);
        P->res64 = P->amx.take<double2>(P->nn);
        P->minv32 = P->amx.take<float4>(P->nn);
        return CP_OK;
    }
    return failf(CP_ERR_INVALID, "unknown precision %d", precision);
}

static int max_width_instantiated(int w)
{
    if (w <= 9) return 9;
    if (w <= 12) return 12;
    if (w <= 16) return 16;
    if (w <= 24) return 24;
    if (w <= 32) return 32;
    return 0;
}

int problem_create(cp_ctx* ctx, int nn, int ne, const double* x, const double* y, const int* conn,
                   const double* rp, const double* u0, const unsigned char* mask, const cp_material* mat,
                   hipStream_t stream, cp_problem** out, int npe)
{
    *out = nullptr;
    if (!ctx || nn < 1 || ne < 1 || !x || !y || !conn || !rp || !mask || !mat)
        return fail(CP_ERR_INVALID, "cp_problem_create: null argument or empty mesh");
    if (npe != 4 && npe != 8) return fail(CP_ERR_INVALID, "cp_problem_create: elements must have 4 or 8 nodes");
    if (!(mat->E > 0) || !(mat->sigmaT > 0) || mat->law < 0 || mat->law > 2 || mat->plane < 0 || mat->plane > 1)
        return fail(CP_ERR_INVALID, "cp_problem_create: invalid material (E and sigmaT must be positive)");
    // the elasticity matrix (formC, PlasticitySolver.cu:226-242) is positive
    // definite only for -1 < nu < 1/2
    if (!(mat->nu > -1.) || !(mat->nu < 0.5) || !(mat->h > 0))
        return fail(CP_ERR_INVALID, "cp_problem_create: invalid material (need -1 < nu < 0.5 and h > 0)");
    if (nn > (1 << 30) / 2 || ne > (npe == 8 ? (1 << 24) : (1 << 29)))
        return fail(CP_ERR_INVALID, "cp_problem_create: mesh too large");
    Adjacency adj;
    if (int s = build_adjacency(nn, ne, npe, conn, adj)) return s;

    std::unique_ptr<cp_problem> P(new cp_problem());
    P->ctx = ctx;
    P->nn = nn;
    P->ne = ne;
    P->npe = npe;
    P->M = cpk::MatParams{mat->E, mat->nu, mat->h, mat->sigmaT, mat->law, mat->plane, mat->K_T, mat->A, mat->m, mat->n};
    P->nslices = (nn + cpk::kWave - 1) / cpk::kWave;
    P->nbn = blocks_for(nn);
    P->nbe = blocks_for(ne);
    if (npe == 8) P->conn_host.assign(conn, conn + 8 * (size_t)ne);

    // sliced ELL layout
    std::vector<int2> slices(P->nslices);
    long long off = 0;
    int maxw = 0;
    for (int s = 0; s < P->nslices; ++s) {
        int w = 0;
        for (int n = s * cpk::kWave; n < std::min(nn, (s + 1) * cpk::kWave); ++n)
            w = std::max(w, adj.ptr[n + 1] - adj.ptr[n]);
        if (off > (long long)INT32_MAX - (long long)w * cpk::kWave)
            return fail(CP_ERR_INVALID, "cp_problem_create: matrix too large for 32-bit slice offsets");
        slices[s] = make_int2((int)off, w);
        off += (long long)w * cpk::kWave;
        maxw = std::max(maxw, w);
    }
    P->nblk_stored = off;
    P->nblk = adj.ptr[nn];
    P->max_w = maxw;
    if (!max_width_instantiated(maxw))
        return failf(CP_ERR_INVALID, "cp_problem_create: node with %d neighbours (max 31 supported)", maxw - 1);
    std::vector<int> hcol(off);
    for (int s = 0; s < P->nslices; ++s)
        for (int lane = 0; lane < cpk::kWave; ++lane) {
            const int n = s * cpk::kWave + lane;
            for (int k = 0; k < slices[s].y; ++k) {
                int c = n < nn ? n : 0;
                if (n < nn && k < adj.ptr[n + 1] - adj.ptr[n]) c = adj.node[adj.ptr[n] + k];
                hcol[(size_t)slices[s].x + (size_t)k * cpk::kWave + lane] = c;
            }
        }
    // node -> (element, slots of K_e(i, j) in the node's row)
    const long long ninc = adj.incPtr[nn];
    std::vector<int2> hinc(npe == 4 ? ninc : 0);
    // 8-node elements: per stored block (compact order adj.ptr[n] + k) the
    // element blocks 64 e + 8 i + j that land there, in element order
    std::vector<int> bptr, bsrc, bdst;
    if (npe == 4) {
        for (int n = 0; n < nn; ++n) {
            const int* row = adj.node.data() + adj.ptr[n];
            const int deg = adj.ptr[n + 1] - adj.ptr[n];
            for (int q = adj.incPtr[n]; q < adj.incPtr[n + 1]; ++q) {
                const int e = adj.inc[q] >> 2;
                unsigned packed = 0;
                for (int j = 0; j < 4; ++j) {
                    const int v = conn[4 * e + j];
                    int k = 0;
                    while (k < deg && row[k] != v) ++k;
                    packed |= (unsigned)k << (8 * j);
                }
                hinc[q] = make_int2(e, (int)packed);
            }
        }
    } else {
        const long long nblk = adj.ptr[nn];
        bptr.assign(nblk + 1, 0);
        bsrc.resize(64 * (size_t)ne);
        bdst.resize(nblk);
        std::vector<int> slot(8 * (size_t)ninc);
        for (int n = 0; n < nn; ++n) {
            const int* row = adj.node.data() + adj.ptr[n];
            const int deg = adj.ptr[n + 1] - adj.ptr[n];
            for (int k = 0; k < deg; ++k)
                bdst[adj.ptr[n] + k] = slices[n / cpk::kWave].x + k * cpk::kWave + (n % cpk::kWave);
            for (int q = adj.incPtr[n]; q < adj.incPtr[n + 1]; ++q) {
                const int e = adj.inc[q] / 8;
                for (int j = 0; j < 8; ++j) {
                    const int v = conn[8 * (size_t)e + j];
                    int k = 0;
                    while (k < deg && row[k] != v) ++k;
                    slot[8 * (size_t)q + j] = k;
                    ++bptr[adj.ptr[n] + k + 1];
                }
            }
        }
        for (long long b = 0; b < nblk; ++b) bptr[b + 1] += bptr[b];
        std::vector<int> fill(bptr.begin(), bptr.end() - 1);
        for (int n = 0; n < nn; ++n)
            for (int q = adj.incPtr[n]; q < adj.incPtr[n + 1]; ++q) {
                const int e = adj.inc[q] / 8, i = adj.inc[q] % 8;
                for (int j = 0; j < 8; ++j) bsrc[fill[adj.ptr[n] + slot[8 * (size_t)q + j]]++] = 64 * e + 8 * i + j;
            }
    }

    Layout L = problem_layout(nn, ne, P->nblk_stored, npe == 4 ? ninc : 0, P->nslices, std::max(P->nbn, P->nbe),
                              npe);
    void* mem = nullptr;
    if (hipMalloc(&mem, L.common) != hipSuccess) {
        (void)hipGetLastError();
        return failf(CP_ERR_NOMEM, "hipMalloc of %zu bytes failed", L.common);
    }
    P->common.base = (char*)mem;
    P->common.size = L.common;
    Arena& A = P->common;
    P->xy = A.take<double2>(nn);
    P->rp0d = A.take<double2>(nn);
    P->u0d = A.take<double2>(nn);
    P->b = A.take<double2>(nn);
    P->u = A.take<double2>(nn);
    P->u_prev = A.take<double2>(nn);
    P->mask = A.take<uchar2>(nn);
    P->conn = A.take<int4>((size_t)ne * (npe / 4));
    P->ecnu = A.take<double2>(ne);
    P->psi = A.take<double>(ne);
    P->psi_prev = A.take<double>(ne);
    P->G = A.take<double>(ne);
    P->rcur = A.take<double>(ne);
    P->rprev = A.take<double>(ne);
    P->F = A.take<double>((size_t)ne * CP_NFIELDS);
    P->inc_ptr = A.take<int>(nn + 1);
    P->inc = A.take<int2>(npe == 4 ? ninc : 0);
    P->col = A.take<int>(P->nblk_stored);
    P->slice = A.take<int2>(P->nslices);
    P->partials = A.take<double>(5 * (size_t)std::max(P->nbn, P->nbe));
    P->partials_alt = A.take<double>(3 * (size_t)std::max(P->nbn, P->nbe));
    P->sc64 = A.take<cpk::CgScal>(2);
    P->sc32 = A.take<cpk::CgScal>(2);
    P->rms = A.take<cpk::RmScal>(1);
    P->refs = A.take<cpk::RefScal>(1);

    if (stream) {
        P->stream = stream;
        P->own_stream = false;
    } else {
        CP_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    }
    for (auto& e : P->ev) CP_HIP(hipEventCreate(&e));
    void* pinned = nullptr;
    CP_HIP(hipHostMalloc(&pinned, 1024, hipHostMallocDefault));
    P->h_cg = (cpk::CgScal*)pinned;
    P->h_rm = (cpk::RmScal*)((char*)pinned + 256);
    P->h_ref = (cpk::RefScal*)((char*)pinned + 512);

    if (npe == 8) {
        auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t bytes = al(sizeof(int) * ninc) + al(sizeof(int) * bptr.size()) + al(sizeof(int) * bsrc.size()) +
                             al(sizeof(int) * bdst.size()) + al(sizeof(double4) * 64 * (size_t)ne) +
                             al(sizeof(double2) * 8 * (size_t)ne);
        void* m8 = nullptr;
        if (hipMalloc(&m8, bytes) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(P->common.base);
            return failf(CP_ERR_NOMEM, "hipMalloc of %zu bytes (8-node element blocks) failed", bytes);
        }
        P->aq8.base = (char*)m8;
        P->aq8.size = bytes;
        P->q8_inc = P->aq8.take<int>(ninc);
        P->q8_bptr = P->aq8.take<int>(bptr.size());
        P->q8_bsrc = P->aq8.take<int>(bsrc.size());
        P->q8_bdst = P->aq8.take<int>(bdst.size());
        P->q8_keb = P->aq8.take<double4>(64 * (size_t)ne);
        P->q8_fe = P->aq8.take<double2>(8 * (size_t)ne);
    }

    std::vector<double2> hxy(nn), hrp(nn), hu0(nn);
    std::vector<uchar2> hmask(nn);
    for (int n = 0; n < nn; ++n) {
        hxy[n] = make_double2(x[n], y[n]);
        hrp[n] = make_double2(rp[2 * n], rp[2 * n + 1]);
        hu0[n] = u0 ? make_double2(u0[2 * n], u0[2 * n + 1]) : make_double2(0., 0.);
        hmask[n] = make_uchar2(mask[2 * n] ? 1 : 0, mask[2 * n + 1] ? 1 : 0);
    }
    hipStream_t st = P->stream;
    CP_HIP(hipMemcpyAsync(P->xy, hxy.data(), sizeof(double2) * nn, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(P->rp0d, hrp.data(), sizeof(double2) * nn, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(P->u0d, hu0.data(), sizeof(double2) * nn, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(P->mask, hmask.data(), sizeof(uchar2) * nn, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(P->conn, conn, sizeof(int) * npe * (size_t)ne, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(P->inc_ptr, adj.incPtr.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice, st));
    if (npe == 4) {
        CP_HIP(hipMemcpyAsync(P->inc, hinc.data(), sizeof(int2) * ninc, hipMemcpyHostToDevice, st));
    } else {
        CP_HIP(hipMemcpyAsync(P->q8_inc, adj.inc.data(), sizeof(int) * ninc, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(P->q8_bptr, bptr.data(), sizeof(int) * bptr.size(), hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(P->q8_bsrc, bsrc.data(), sizeof(int) * bsrc.size(), hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(P->q8_bdst, bdst.data(), sizeof(int) * bdst.size(), hipMemcpyHostToDevice, st));
    }
    CP_HIP(hipMemcpyAsync(P->col, hcol.data(), sizeof(int) * off, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(P->slice, slices.data(), sizeof(int2) * P->nslices, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemsetAsync(P->F, 0, sizeof(double) * (size_t)ne * CP_NFIELDS, st));
    CP_HIP(hipMemsetAsync(P->sc64, 0, 2 * sizeof(cpk::CgScal), st));
    CP_HIP(hipMemsetAsync(P->sc32, 0, 2 * sizeof(cpk::CgScal), st));
    CP_HIP(hipMemsetAsync(P->rms, 0, sizeof(cpk::RmScal), st));
    CP_HIP(hipMemsetAsync(P->refs, 0, sizeof(cpk::RefScal), st));
    const int nbe = P->nbe, nbn = P->nbn;
    hipLaunchKernelGGL(cpk::k_init_state, dim3(nbe), dim3(cpk::kBlock), 0, st, P->psi, P->ecnu, P->rprev, P->M.E,
                       P->M.nu, ne);
    hipLaunchKernelGGL(cpk::k_init_state, dim3(nbe), dim3(cpk::kBlock), 0, st, P->psi_prev, P->ecnu, P->rprev,
                       P->M.E, P->M.nu, ne);
    CP_HIP(hipMemsetAsync(P->u_prev, 0, sizeof(double2) * nn, st));
    hipLaunchKernelGGL(cpk::k_set_load, dim3(nbn), dim3(cpk::kBlock), 0, st, P->b, P->u, P->rp0d, P->u0d, P->mask,
                       1.0, 1, nn);
    CP_HIP(hipGetLastError());
    CP_HIP(hipStreamSynchronize(st));
    P->rp0.assign(rp, rp + 2 * (size_t)nn);
    P->adj_ptr.swap(adj.ptr);
    P->adj_node.swap(adj.node);
    *out = P.release();
    return CP_OK;
}

// ---------------------------------------------------------------- kernels drivers

static int launch_assemble_q8(cp_problem* P, bool f64)
{
    cpk::Q8Args a{};
    a.conn = P->conn;
    a.xy = P->xy;
    a.ecnu = P->ecnu;
    a.keb = P->q8_keb;
    a.bptr = P->q8_bptr;
    a.bsrc = P->q8_bsrc;
    a.bdst = P->q8_bdst;
    a.va = P->va;
    a.vb = P->vb;
    a.vf = P->vf;
    a.ne = P->ne;
    a.nblk = (int)P->nblk;
    a.nbe8 = blocks_for(8LL * P->ne);
    a.nbb = blocks_for(P->nblk);
    a.plane = P->M.plane;
    a.h = P->M.h;
    hipLaunchKernelGGL(cpk::k_q8_ke, dim3(a.nbe8), dim3(cpk::kBlock), 0, P->stream, a);
    if (f64) hipLaunchKernelGGL(cpk::k_q8_gather<true>, dim3(a.nbb), dim3(cpk::kBlock), 0, P->stream, a);
    else hipLaunchKernelGGL(cpk::k_q8_gather<false>, dim3(a.nbb), dim3(cpk::kBlock), 0, P->stream, a);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int launch_residual(cp_problem* P, const cpk::ResArgs& ra, hipStream_t st)
{
    if (P->npe == 4) {
        hipLaunchKernelGGL(cpk::k_residual64, dim3(P->nbn), dim3(cpk::kBlock), 0, st, ra);
        CP_HIP(hipGetLastError());
        return CP_OK;
    }
    cpk::Q8ResArgs q{};
    q.conn = P->conn;
    q.xy = P->xy;
    q.ecnu = P->ecnu;
    q.u = ra.u;
    q.fe = P->q8_fe;
    q.incPtr = P->inc_ptr;
    q.inc = P->q8_inc;
    q.b = ra.b;
    q.mask = ra.mask;
    q.r64 = ra.r64;
    q.r32 = ra.r32;
    q.sc = ra.sc;
    q.partials = ra.partials;
    q.ne = P->ne;
    q.nn = P->nn;
    q.nbe = P->nbe;
    q.nbn = P->nbn;
    q.plane = P->M.plane;
    q.h = P->M.h;
    hipLaunchKernelGGL(cpk::k_q8_force, dim3(P->nbe), dim3(cpk::kBlock), 0, st, q);
    hipLaunchKernelGGL(cpk::k_q8_fgather, dim3(P->nbn), dim3(cpk::kBlock), 0, st, q);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int launch_assemble(cp_problem* P, int precision, int mg_full)
{
    if (int s = problem_alloc_solver(P, precision)) return s;
    if (P->npe == 8) {
        if (int s = launch_assemble_q8(P, precision == CP_FP64)) return s;
        const dim3 g(P->nbn), b(cpk::kBlock);
        if (P->precond == 1) {
            if (precision == CP_FP64)
                hipLaunchKernelGGL(cpk::k_block_jacobi<double>, g, b, 0, P->stream, P->cg64_args(), P->minv64);
            else hipLaunchKernelGGL(cpk::k_block_jacobi<float>, g, b, 0, P->stream, P->cg32_args(), P->minv32);
            CP_HIP(hipGetLastError());
        } else if (P->precond == 2) {
            if (int s = mg_setup(P, precision, mg_full)) return s;
        }
        if (precision == CP_FP64) P->have_fp64_values = true;
        return CP_OK;
    }
    cpk::AsmArgs a{};
    a.incPtr = P->inc_ptr;
    a.inc = P->inc;
    a.conn = P->conn;
    a.xy = P->xy;
    a.ecnu = P->ecnu;
    a.slice = P->slice;
    a.va = P->va;
    a.vb = P->vb;
    a.vf = P->vf;
    a.nn = P->nn;
    a.nblocks = P->nbn;
    a.plane = P->M.plane;
    a.h = P->M.h;
    const dim3 g(P->nbn), b(cpk::kBlock);
    const bool f64 = precision == CP_FP64;
    int direct = 0;
    if (P->precond == 2)
        if (int s = mg_assemble_direct(P, precision, mg_full, a, &direct)) return s;
    if (!direct) switch (max_width_instantiated(P->max_w)) {
#define CP_ASM_CASE(W)                                                                                  \
    case W:                                                                                             \
        if (f64) hipLaunchKernelGGL((cpk::k_assemble<W, true>), g, b, 0, P->stream, a);                 \
        else hipLaunchKernelGGL((cpk::k_assemble<W, false>), g, b, 0, P->stream, a);                   \
        break;
        CP_ASM_CASE(9)
        CP_ASM_CASE(12)
        CP_ASM_CASE(16)
        CP_ASM_CASE(24)
        CP_ASM_CASE(32)
#undef CP_ASM_CASE
    default:
        return fail(CP_ERR_INVALID, "row width too large");
    }
    CP_HIP(hipGetLastError());
    if (P->precond == 1) {
        if (f64) hipLaunchKernelGGL(cpk::k_block_jacobi<double>, g, b, 0, P->stream, P->cg64_args(), P->minv64);
        else hipLaunchKernelGGL(cpk::k_block_jacobi<float>, g, b, 0, P->stream, P->cg32_args(), P->minv32);
        CP_HIP(hipGetLastError());
    } else if (P->precond == 2) {
        if (int s = mg_setup(P, precision, mg_full)) return s;
    }
    if (f64) P->have_fp64_values = true;
    return CP_OK;
}

double host_bnorm2(const cp_problem* P)
{
    if (P->bn_load == P->load) return P->bn_val;
    double s = 0.;
    for (double v : P->rp0) s += (P->load * v) * (P->load * v);
    P->bn_load = P->load;
    P->bn_val = s;
    return s;
}

// Launches `count` CG iterations with the row width unrolled to the mesh's
// maximum (9 for the structured quad meshes: a node and its 8 neighbours).
template <int W, bool PRE, typename T>
static void launch_iters_w(const cpk::CgArgs<T>& a, int count, dim3 g, hipStream_t st)
{
    for (int j = 0; j < count; ++j) {
        hipLaunchKernelGGL((cpk::k_cg_iter<W, PRE, T>), g, dim3(cpk::kBlock), 0, st, a);
        hipLaunchKernelGGL((cpk::k_cg_step<T>), dim3(1), dim3(cpk::kBlock), 0, st, a);
    }
}

template <bool PRE, typename T>
static void launch_iters_p(int max_w, const cpk::CgArgs<T>& a, int count, dim3 g, hipStream_t st)
{
    if (max_w <= 9) launch_iters_w<9, PRE, T>(a, count, g, st);
    else if (max_w <= 12) launch_iters_w<12, PRE, T>(a, count, g, st);
    else if (max_w <= 16) launch_iters_w<16, PRE, T>(a, count, g, st);
    else launch_iters_w<0, PRE, T>(a, count, g, st);
}

template <typename T>
void launch_cg_iters(int max_w, int pre, const cpk::CgArgs<T>& a, int count, dim3 g, hipStream_t st)
{
    if (pre) launch_iters_p<true, T>(max_w, a, count, g, st);
    else launch_iters_p<false, T>(max_w, a, count, g, st);
}
template void launch_cg_iters<double>(int, int, const cpk::CgArgs<double>&, int, dim3, hipStream_t);
template void launch_cg_iters<float>(int, int, const cpk::CgArgs<float>&, int, dim3, hipStream_t);

// One launch per iteration (k_cg_iter_fused), iterations k0 .. k0+count-1.
template <int W, bool PRE, typename T>
static void launch_fused_w(const cpk::CgArgs<T>& a, int k0, int count, dim3 g, hipStream_t st)
{
    for (int j = 0; j < count; ++j)
        hipLaunchKernelGGL((cpk::k_cg_iter_fused<W, PRE, T>), g, dim3(cpk::kBlock), 0, st, a, k0 + j);
}

template <typename T>
static void launch_fused(int max_w, int pre, const cpk::CgArgs<T>& a, int k0, int count, dim3 g, hipStream_t st)
{
    if (pre) {
        if (max_w <= 9) launch_fused_w<9, true, T>(a, k0, count, g, st);
        else if (max_w <= 12) launch_fused_w<12, true, T>(a, k0, count, g, st);
        else if (max_w <= 16) launch_fused_w<16, true, T>(a, k0, count, g, st);
        else launch_fused_w<0, true, T>(a, k0, count, g, st);
    } else {
        if (max_w <= 9) launch_fused_w<9, false, T>(a, k0, count, g, st);
        else if (max_w <= 12) launch_fused_w<12, false, T>(a, k0, count, g, st);
        else if (max_w <= 16) launch_fused_w<16, false, T>(a, k0, count, g, st);
        else launch_fused_w<0, false, T>(a, k0, count, g, st);
    }
}

static bool use_fused(const cp_problem* P) { return P->nbn <= cpk::kFusedMaxBlocks; }

// The iteration kernel the solver uses, `count` times: the fused form (its
// scalar step included) on grids up to kFusedMaxBlocks, else the iteration
// kernel alone without its separate one-block step.
template <typename T>
static void launch_iter_kernels_only(int max_w, int pre, const cpk::CgArgs<T>& a, int count, dim3 g, hipStream_t st)
{
    if ((int)g.x <= cpk::kFusedMaxBlocks) {
        launch_fused<T>(max_w, pre, a, 0, count, g, st);
        return;
    }
    for (int j = 0; j < count; ++j) {
        if (max_w <= 9) {
            if (pre) hipLaunchKernelGGL((cpk::k_cg_iter<9, true, T>), g, dim3(cpk::kBlock), 0, st, a);
            else hipLaunchKernelGGL((cpk::k_cg_iter<9, false, T>), g, dim3(cpk::kBlock), 0, st, a);
        } else {
            if (pre) hipLaunchKernelGGL((cpk::k_cg_iter<0, true, T>), g, dim3(cpk::kBlock), 0, st, a);
            else hipLaunchKernelGGL((cpk::k_cg_iter<0, false, T>), g, dim3(cpk::kBlock), 0, st, a);
        }
    }
}

template <int W, bool PRE, typename T>
static void launch_init2_w(const cpk::CgArgs<T>& a, dim3 g, hipStream_t st)
{
    hipLaunchKernelGGL((cpk::k_cg_init2<W, PRE, T>), g, dim3(cpk::kBlock), 0, st, a);
    hipLaunchKernelGGL((cpk::k_cg_start<T>), dim3(1), dim3(cpk::kBlock), 0, st, a);
}

template <typename T>
static void launch_init2(int max_w, int pre, const cpk::CgArgs<T>& a, dim3 g, hipStream_t st)
{
    if (pre) {
        if (max_w <= 9) launch_init2_w<9, true, T>(a, g, st);
        else launch_init2_w<0, true, T>(a, g, st);
    } else {
        if (max_w <= 9) launch_init2_w<9, false, T>(a, g, st);
        else launch_init2_w<0, false, T>(a, g, st);
    }
}

// Runs a CG solve whose setup kernels have been launched: chunks of
// iteration launches, polling the done flag between chunks.
template <typename T>
static int run_cg_iterations(cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc, double tol2, double bnorm2,
                             long long* iters, int* status, double* ms)
{
    hipStream_t st = P->stream;
    const dim3 g(P->nbn);
    CP_HIP(hipEventRecord(P->ev[6], st));
    long long launched = 0;
    int chunk = (int)std::min<long long>(4096, std::max<long long>(16, P->last_cg_iters / 2));
    double g_prev = -1.;
    long long it_prev = 0;
    const bool fused = use_fused(P);
    for (;;) {
        if (fused) launch_fused<T>(P->max_w, P->precond, a, (int)launched, chunk, g, st);
        else launch_cg_iters<T>(P->max_w, P->precond, a, chunk, g, st);
        CP_HIP(hipGetLastError());
        launched += chunk;
        // fused launches keep the scalars in two slots; the last launch's is current
        const cpk::CgScal* cur = fused ? dsc + ((launched - 1) & 1) : dsc;
        CP_HIP(hipMemcpyAsync(P->h_cg, cur, sizeof(cpk::CgScal), hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
        const cpk::CgScal h = *P->h_cg;
        if (h.done) break;
        // predict the remaining iterations from the observed decay of (r, r):
        // since the last observation, or since the start at the first one
        // (instead of doubling the chunk there); iterations launched after
        // convergence return at once, so the chunk sizes never change results
        const double target = tol2 * (bnorm2 > 0. ? bnorm2 : h.bnorm2);
        int next = std::min(4096, chunk * 2);
        const double r0 = g_prev > 0. ? g_prev : h.rho0;
        const long long i0 = g_prev > 0. ? it_prev : 0;
        if (r0 > 0. && h.rho > 0. && h.rho < r0 && target > 0. && h.iters > i0) {
            const double rate = std::log(r0 / h.rho) / double(h.iters - i0);
            const double rem = std::log(h.rho / target) / rate;
            if (rem > 0 && rem < 1e7) next = (int)std::min(4096., std::max(8., rem * 1.05 + 4.));
        }
        g_prev = h.rho;
        it_prev = h.iters;
        chunk = next;
        if (launched > (long long)h.maxit + 8192) break;  // the kernel stops itself at maxit
    }
    CP_HIP(hipEventRecord(P->ev[7], st));
    CP_HIP(hipEventSynchronize(P->ev[7]));
    float t = 0.f;
    CP_HIP(hipEventElapsedTime(&t, P->ev[6], P->ev[7]));
    *ms = t;
    *iters = P->h_cg->iters;
    *status = P->h_cg->status;
    return CP_OK;
}

// One masked CG solve of K u = b from the current u (FP64).
static int cg_fp64(cp_problem* P, double tol, int maxit, long long* iters, double* relres, double* ms)
{
    const double bn = host_bnorm2(P);
    auto a = P->cg64_args();
    if (int s = enqueue_cg_init(P, CP_FP64, bn, tol, maxit)) return s;
    int status = 0;
    if (int s = P->precond == 2 ? run_mg_cg<double>(P, a, P->sc64, tol * tol, bn, iters, &status, ms)
                                : run_cg_iterations<double>(P, a, P->sc64, tol * tol, bn, iters, &status, ms))
        return s;
    const cpk::CgScal h = *P->h_cg;
    if (relres) *relres = std::sqrt(h.rho / (h.bnorm2 > 0 ? h.bnorm2 : 1.));
    P->last_cg_iters = *iters;
    if (*iters > 0) P->last_cg_ms_per_iter = *ms / double(*iters);
    if (status == 2) return fail(CP_ERR_NAN, "CG: NaN in the residual");
    if (status == 1) return failf(CP_ERR_NOT_CONVERGED, "CG: %d iterations without convergence", maxit);
    return CP_OK;
}

cpk::ResArgs residual_args(cp_problem* P)
{
    cpk::ResArgs ra{};
    ra.incPtr = P->inc_ptr;
    ra.inc = P->inc;
    ra.conn = P->conn;
    ra.xy = P->xy;
    ra.ecnu = P->ecnu;
    ra.u = P->u;
    ra.b = P->b;
    ra.mask = P->mask;
    ra.r64 = P->res64;
    ra.r32 = P->b32;
    ra.sc = P->refs;
    ra.partials = P->partials;
    ra.nn = P->nn;
    ra.nblocks = P->nbn;
    ra.plane = P->M.plane;
    ra.h = P->M.h;
    return ra;
}

// CG setup + r_0 / w_0 kernels (FP64 from the current u, or FP32 from zero
// on the refinement residual).
int enqueue_cg_init(cp_problem* P, int precision, double bnorm2, double tol, int maxit)
{
    const dim3 g(P->nbn), b(cpk::kBlock);
    hipStream_t st = P->stream;
    if (precision == CP_FP64) {
        auto a = P->cg64_args();
        hipLaunchKernelGGL(cpk::k_cg_setup, dim3(1), dim3(1), 0, st, P->sc64, bnorm2, tol * tol, maxit);
        hipLaunchKernelGGL(cpk::k_cg_init1<double>, g, b, 0, st, a, 0);
        if (P->precond != 2) launch_init2<double>(P->max_w, P->precond, a, g, st);
    } else {
        auto a = P->cg32_args();
        hipLaunchKernelGGL(cpk::k_cg_setup, dim3(1), dim3(1), 0, st, P->sc32, bnorm2, tol * tol, maxit);
        hipLaunchKernelGGL(cpk::k_cg_init1<float>, g, b, 0, st, a, 1);
        if (P->precond != 2) launch_init2<float>(P->max_w, P->precond, a, g, st);
    }
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int enqueue_residual(cp_problem* P)
{
    if (int s = launch_residual(P, residual_args(P), P->stream)) return s;
    // the 4-node residual also sums the magnitude of |K| |u| + |b|
    if (P->npe == 4)
        hipLaunchKernelGGL(cpk::k_residual_total<2>, dim3(1), dim3(cpk::kBlock), 0, P->stream, P->refs,
                           (const double*)P->partials, P->nbn);
    else
        hipLaunchKernelGGL(cpk::k_residual_total<1>, dim3(1), dim3(cpk::kBlock), 0, P->stream, P->refs,
                           (const double*)P->partials, P->nbn);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int refine_decide(Refine& R, double rn, double mag2, double tol, double inner_tol, double* itol)
{
    if (rn != rn) return REF_NAN;
    if (R.ref_norm == 0.) R.ref_norm = rn;
    const double eps = 2.220446049250313e-16;
    const double floor2 = kRefFloor * kRefFloor * eps * eps * mag2;
    if (rn == 0. || rn <= tol * tol * R.ref_norm || rn <= floor2) return REF_DONE;
    if (R.prev > 0. && rn > 0.25 * R.prev && ++R.stalled >= 3) return REF_STALLED;
    R.prev = rn;
    // ask one FP32 solve for the reduction still needed (half of it, for
    // margin), but never for less than inner_tol: FP32 cannot go further
    const double need = 0.5 * tol * std::sqrt(R.ref_norm / rn);
    *itol = std::max(inner_tol, std::min(need, 0.5));
    return REF_SOLVE;
}

int enqueue_correction(cp_problem* P)
{
    hipLaunchKernelGGL(cpk::k_add_correction, dim3(P->nbn), dim3(cpk::kBlock), 0, P->stream, P->u,
                       (const float2*)P->x32, P->nn, P->nbn);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int enqueue_plastic(cp_problem* P)
{
    hipLaunchKernelGGL(cpk::k_plastic_strains, dim3(P->nbe), dim3(cpk::kBlock), 0, P->stream, P->F, P->psi,
                       (const double2*)P->ecnu, P->M.E, P->M.nu, P->ne, P->nbe);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int enqueue_rm_reset(cp_problem* P)
{
    hipLaunchKernelGGL(cpk::k_rm_reset, dim3(1), dim3(1), 0, P->stream, P->rms);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

// FP32 CG on K32 d = r64 with FP64 iterative refinement of u.  loose: an
// inexact secant iteration's solve (tol above the final CG tolerance): once
// one FP32 solve was asked for the whole reduction still needed and met it,
// the refinement ends without an FP64 residual to confirm it (the level's
// final solve is always confirmed).
static int cg_mixed(cp_problem* P, double tol, int maxit, double inner_tol, long long* iters, long long* refs,
                    double* relres, double* ms, bool loose = false)
{
    hipStream_t st = P->stream;
    const double bn = host_bnorm2(P);
    auto a = P->cg32_args();
    *iters = 0;
    *refs = 0;
    *ms = 0.;
    Refine R;
    R.ref_norm = bn;
    for (int k = 0; k < 200; ++k) {
        if (int s = enqueue_residual(P)) return s;
        CP_HIP(hipMemcpyAsync(P->h_ref, P->refs, sizeof(cpk::RefScal), hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
        const double rn = P->h_ref->rnorm2;
        ++*refs;
        double itol = 0.;
        const int act = refine_decide(R, rn, P->h_ref->mag2, tol, inner_tol, &itol);
        if (act == REF_NAN) return fail(CP_ERR_NAN, "refinement: NaN residual");
        if (relres) *relres = R.ref_norm > 0 ? std::sqrt(rn / R.ref_norm) : 0.;
        if (act == REF_DONE) return CP_OK;
        if (act == REF_STALLED)
            return failf(CP_ERR_NOT_CONVERGED,
                         "refinement stalled at relative residual %.3e (rounding floor of the FP64 residual %.3e)",
                         std::sqrt(rn / R.ref_norm),
                         kRefFloor * 2.220446049250313e-16 * std::sqrt(P->h_ref->mag2 / R.ref_norm));
        if (int s = enqueue_cg_init(P, CP_MIXED, rn, itol, maxit)) return s;
        long long it = 0;
        int status = 0;
        double t = 0.;
        if (int s = P->precond == 2 ? run_mg_cg<float>(P, a, P->sc32, itol * itol, rn, &it, &status, &t, true)
                                    : run_cg_iterations<float>(P, a, P->sc32, itol * itol, rn, &it, &status, &t))
            return s;
        *iters += it;
        *ms += t;
        P->last_cg_iters = it;
        if (it > 0) P->last_cg_ms_per_iter = t / double(it);
        if (status == 2) return fail(CP_ERR_NAN, "inner CG: NaN");
        if (int s = enqueue_correction(P)) return s;
        const cpk::CgScal& h = *P->h_cg;
        if (loose && itol > inner_tol && status == 0 && h.done && h.rho0 > 0. && h.rho <= itol * itol * h.rho0) {
            if (relres) *relres = std::sqrt(rn * (h.rho / h.rho0) / R.ref_norm);
            return CP_OK;
        }
    }
    return fail(CP_ERR_NOT_CONVERGED, "refinement: 200 passes without convergence");
}

int launch_return_map(cp_problem* P, int k, int aitken, int fixed_outer, int max_outer, int elastic, double tol,
                      int solved_tight)
{
    cpk::RmArgs a{};
    a.conn = P->conn;
    a.xy = P->xy;
    a.u = P->u;
    a.ecnu = P->ecnu;
    a.psi = P->psi;
    a.F = P->F;
    a.G = P->G;
    a.rcur = P->rcur;
    a.rprev = P->rprev;
    a.sc = P->rms;
    a.partials = P->partials;
    a.M = P->M;
    a.ne = P->ne;
    a.nblocks = P->nbe;
    a.k = k;
    a.aitken = aitken;
    a.fixed_outer = fixed_outer;
    a.max_outer = max_outer;
    a.elastic = elastic;
    a.solved_tight = solved_tight;
    a.tol = tol;
    const dim3 g(P->nbe), b(cpk::kBlock);
    if (P->npe == 8) hipLaunchKernelGGL(cpk::k_return_map<8>, g, b, 0, P->stream, a);
    else hipLaunchKernelGGL(cpk::k_return_map<4>, g, b, 0, P->stream, a);
    hipLaunchKernelGGL(cpk::k_rm_decide, dim3(1), b, 0, P->stream, a);
    hipLaunchKernelGGL(cpk::k_apply_update, g, b, 0, P->stream, P->psi, P->ecnu, (const double*)P->G,
                       (const double*)P->rcur, P->rprev, (const cpk::RmScal*)P->rms, P->M.E, P->M.nu, P->ne, P->nbe,
                       aitken);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int set_load(cp_problem* P, double factor, int reset)
{
    P->load = factor;
    hipLaunchKernelGGL(cpk::k_set_load, dim3(P->nbn), dim3(cpk::kBlock), 0, P->stream, P->b, P->u, P->rp0d, P->u0d,
                       P->mask, factor, reset, P->nn);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

int reset_state(cp_problem* P)
{
    hipLaunchKernelGGL(cpk::k_init_state, dim3(P->nbe), dim3(cpk::kBlock), 0, P->stream, P->psi, P->ecnu, P->rprev,
                       P->M.E, P->M.nu, P->ne);
    CP_HIP(hipGetLastError());
    CP_HIP(hipMemcpyAsync(P->psi_prev, P->psi, sizeof(double) * P->ne, hipMemcpyDeviceToDevice, P->stream));
    CP_HIP(hipMemsetAsync(P->u_prev, 0, sizeof(double2) * P->nn, P->stream));
    P->lam_prev = 0.;
    P->solved = false;
    return set_load(P, P->load, 1);
}

// Next load level; with the predictor, extrapolate u and log(psi) from the
// last two converged levels (the level before the first one is the
// unloaded state u = 0, psi = 1 at factor 0).
int advance_load(cp_problem* P, double factor, int predictor)
{
    if (P->solved && predictor && P->load != P->lam_prev) {
        const double ratio = (factor - P->load) / (P->load - P->lam_prev);
        hipLaunchKernelGGL(cpk::k_predict_u, dim3(P->nbn), dim3(cpk::kBlock), 0, P->stream, P->u, P->u_prev, ratio,
                           P->nn);
        if (predictor >= 2)
            hipLaunchKernelGGL(cpk::k_predict_psi, dim3(P->nbe), dim3(cpk::kBlock), 0, P->stream, P->psi,
                               P->psi_prev, P->ecnu, ratio, P->M.E, P->M.nu, P->ne);
        else
            CP_HIP(hipMemcpyAsync(P->psi_prev, P->psi, sizeof(double) * P->ne, hipMemcpyDeviceToDevice, P->stream));
        CP_HIP(hipGetLastError());
        P->lam_prev = P->load;
    } else if (P->solved) {
        CP_HIP(hipMemcpyAsync(P->u_prev, P->u, sizeof(double2) * P->nn, hipMemcpyDeviceToDevice, P->stream));
        CP_HIP(hipMemcpyAsync(P->psi_prev, P->psi, sizeof(double) * P->ne, hipMemcpyDeviceToDevice, P->stream));
        P->lam_prev = P->load;
    }
    P->solved = false;
    return set_load(P, factor, 0);
}

// One load level: the secant loop of solveCUDA (PlasticitySolver.cu:487-503)
// with caps, optional Aitken steps and the mixed-precision linear solve.
int solve_level(cp_problem* P, const cp_options* opt, int elastic, cp_stats* st)
{
    cp_stats S{};
    const double t0 = now_s();
    const int prec = opt->precision;
    if (prec != CP_FP64 && prec != CP_MIXED) return fail(CP_ERR_INVALID, "unknown precision");
    if (int s = problem_alloc_solver(P, prec)) return s;
    if (opt->precond < 0 || opt->precond > 2) return fail(CP_ERR_INVALID, "precond must be 0, 1 or 2");
    P->precond = opt->precond;
    hipStream_t stream = P->stream;
    if (!opt->warm)
        if (int s = reset_state(P)) return s;
    hipLaunchKernelGGL(cpk::k_rm_reset, dim3(1), dim3(1), 0, stream, P->rms);
    const int max_outer = opt->fixed_outer > 0 ? opt->fixed_outer : (elastic ? 1 : std::max(1, opt->max_outer));
    // multigrid CG needs tens of iterations: a cap keeps a stagnating solve
    // (FP32 on a badly conditioned level) from running on
    const int maxit = opt->cg_maxit > 0 ? opt->cg_maxit : (P->precond == 2 ? 2000 : 1000000000);
    const double inner = opt->inner_tol > 0 ? opt->inner_tol : 1e-3;
    int rc = CP_OK;
    double prev_rel = 1.0;
    std::string lin_msg;
    bool lin_last = false;
    for (int k = 0; k < max_outer; ++k) {
        CP_HIP(hipEventRecord(P->ev[0], stream));
        // multigrid: a fresh hierarchy at the first secant iteration of the
        // level, later ones may reuse the coarse levels
        if (int s = launch_assemble(P, prec, k == 0)) return s;
        CP_HIP(hipEventRecord(P->ev[1], stream));
        long long it = 0, refs = 0;
        double relres = 0., cgms = 0.;
        // inexact secant iterations: the linear solve only needs to be as
        // accurate as the current secant residual (cg_forcing > 0)
        double ctol = opt->cg_tol;
        if (opt->cg_forcing > 0. && !elastic && opt->fixed_outer <= 0)
            ctol = std::max(opt->cg_tol, opt->cg_forcing * prev_rel);
        const int tight = ctol <= opt->cg_tol;
        int s = prec == CP_FP64 ? cg_fp64(P, ctol, maxit, &it, &relres, &cgms)
                                : cg_mixed(P, ctol, maxit, inner, &it, &refs, &relres, &cgms, !tight);
        S.cg_iters += it;
        S.refinements += refs;
        S.cg_seconds += cgms * 1e-3;
        if (s == CP_ERR_NAN) {
            rc = s;
            S.status = CP_ERR_NAN;
            S.outer_iters = k + 1;
            break;
        }
        if (s && s != CP_ERR_NOT_CONVERGED) return s;
        // a linear solve that stopped at its cap (cg_maxit, or a stalled
        // refinement) is not a tight one: the level cannot exit on it, and a
        // level whose last iteration ends on one is not converged
        const bool lin_ok = s == CP_OK;
        if (!lin_ok) lin_msg = cp_last_error();
        CP_HIP(hipEventRecord(P->ev[2], stream));
        if (int q = launch_return_map(P, k, opt->aitken, opt->fixed_outer, max_outer, elastic, opt->tol,
                                      tight && lin_ok))
            return q;
        CP_HIP(hipMemcpyAsync(P->h_rm, P->rms, sizeof(cpk::RmScal), hipMemcpyDeviceToHost, stream));
        CP_HIP(hipEventRecord(P->ev[3], stream));
        CP_HIP(hipStreamSynchronize(stream));
        float ta = 0.f, tr = 0.f;
        CP_HIP(hipEventElapsedTime(&ta, P->ev[0], P->ev[1]));
        CP_HIP(hipEventElapsedTime(&tr, P->ev[2], P->ev[3]));
        S.asm_seconds += ta * 1e-3;
        S.rm_seconds += tr * 1e-3;
        const cpk::RmScal h = *P->h_rm;
        S.outer_iters = k + 1;
        S.rel_err = h.rel_err;
        // once the secant residual is below tol, the next solve is a tight one
        prev_rel = h.rel_err <= opt->tol ? 0. : h.rel_err;
        if (h.status == 2) {
            rc = fail(CP_ERR_NAN, "NaN in the secant update");
            S.status = CP_ERR_NAN;
            break;
        }
        if (h.last) {
            S.status = h.converged && lin_ok ? CP_OK : CP_ERR_NOT_CONVERGED;
            lin_last = !lin_ok;
            break;
        }
    }
    if (!elastic && S.status != CP_ERR_NAN) {
        hipLaunchKernelGGL(cpk::k_plastic_strains, dim3(P->nbe), dim3(cpk::kBlock), 0, stream, P->F, P->psi,
                           (const double2*)P->ecnu, P->M.E, P->M.nu, P->ne, P->nbe);
        CP_HIP(hipGetLastError());
    }
    CP_HIP(hipStreamSynchronize(stream));
    S.seconds = now_s() - t0;
    P->solved = (S.status == CP_OK);
    if (st) *st = S;
    if (rc) return rc;
    if (S.status == CP_ERR_NOT_CONVERGED && lin_last)
        return failf(CP_ERR_NOT_CONVERGED, "the level's last linear solve did not converge (%s)", lin_msg.c_str());
    if (S.status == CP_ERR_NOT_CONVERGED) return fail(CP_ERR_NOT_CONVERGED, "secant loop hit its iteration cap");
    return CP_OK;
}

}  // namespace cp

// ====================================================================== C-ABI

extern "C" {


int cp_abi_version(void) { return CP_ABI_VERSION; }

int cp_material_f(const cp_material* mat, double eps, double* sigma)
{
    if (!mat || !sigma) return cp::fail(CP_ERR_INVALID, "cp_material_f: null argument");
    if (mat->law < CP_LAW_LINEAR || mat->law > CP_LAW_SWIFT) return cp::fail(CP_ERR_INVALID, "unknown material law");
    const cpk::MatParams M{mat->E, mat->nu, mat->h, mat->sigmaT, mat->law, mat->plane, mat->K_T, mat->A, mat->m, mat->n};
    *sigma = cpk::mat_f(M, eps);
    return CP_OK;
}

int cp_ctx_create(int device, cp_ctx** out)
{
    if (!out) return cp::fail(CP_ERR_INVALID, "cp_ctx_create: null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return cp::fail(CP_ERR_NO_DEVICE, "no HIP device visible");
    }
    if (device < 0 || device >= n) return cp::failf(CP_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    CP_HIP(hipSetDevice(device));
    cp_ctx* c = new cp_ctx();
    c->device = device;
    if (hipGetDeviceProperties(&c->prop, device) != hipSuccess) {
        delete c;
        return cp::fail(CP_ERR_DEVICE, "hipGetDeviceProperties failed");
    }
    // this device's copy of the staging array (device `device` is current)
    const hipError_t e = hipGetSymbolAddress(&c->scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !c->scratch) {
        delete c;
        return cp::failf(CP_ERR_DEVICE, "hipGetSymbolAddress of the staging array failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return CP_OK;
}

int cp_ctx_destroy(cp_ctx* ctx)
{
    delete ctx;
    return CP_OK;
}

int cp_device_info(cp_ctx* ctx, char* name, int name_len, size_t* total_mem, int* cu_count)
{
    if (!ctx) return cp::fail(CP_ERR_INVALID, "null ctx");
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", ctx->prop.name, ctx->prop.gcnArchName);
    }
    if (total_mem) *total_mem = ctx->prop.totalGlobalMem;
    if (cu_count) *cu_count = ctx->prop.multiProcessorCount;
    return CP_OK;
}

int cp_problem_create(cp_ctx* ctx, int nn, int ne, const double* x, const double* y, const int* conn,
                      const double* rp, const double* u0, const unsigned char* mask, const cp_material* mat,
                      cp_problem** out)
{
    if (!out) return cp::fail(CP_ERR_INVALID, "null out");
    if (ctx) CP_HIP(hipSetDevice(ctx->device));
    return cp::problem_create(ctx, nn, ne, x, y, conn, rp, u0, mask, mat, nullptr, out);
}

int cp_problem_create_n(cp_ctx* ctx, int npe, int nn, int ne, const double* x, const double* y, const int* conn,
                        const double* rp, const double* u0, const unsigned char* mask, const cp_material* mat,
                        cp_problem** out)
{
    if (!out) return cp::fail(CP_ERR_INVALID, "null out");
    if (ctx) CP_HIP(hipSetDevice(ctx->device));
    return cp::problem_create(ctx, nn, ne, x, y, conn, rp, u0, mask, mat, nullptr, out, npe);
}

int cp_problem_destroy(cp_problem* P)
{
    if (!P) return CP_OK;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    for (auto& e : P->ev)
        if (e) (void)hipEventDestroy(e);
    cp::mg_destroy(P);
    if (P->h_cg) (void)hipHostFree(P->h_cg);
    if (P->common.base) (void)hipFree(P->common.base);
    if (P->a64.base) (void)hipFree(P->a64.base);
    if (P->amx.base) (void)hipFree(P->amx.base);
    if (P->aq8.base) (void)hipFree(P->aq8.base);
    if (P->own_stream && P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
    return CP_OK;
}

int cp_problem_bytes(int nn, int ne, long long nnz, int precision, size_t* bytes)
{
    if (!bytes || nn < 1 || ne < 1) return cp::fail(CP_ERR_INVALID, "cp_problem_bytes: bad arguments");
    // stored blocks ~ true blocks (nnz/4) plus slice padding; 4 incidences per element
    const long long nblk = nnz / 4 + (long long)nn / 8 + 64 * 16;
    const int nb = blocks_for(std::max(nn, ne));
    cp::Layout L = cp::problem_layout(nn, ne, nblk, 4LL * ne, (nn + 63) / 64, nb);
    *bytes = L.common + (precision == CP_FP64 ? L.fp64 : L.mixed);
    return CP_OK;
}

int cp_problem_csr_nnz(cp_problem* P, long long* nnz)
{
    if (!P || !nnz) return cp::fail(CP_ERR_INVALID, "null argument");
    *nnz = 4LL * P->nblk;
    return CP_OK;
}

int cp_problem_get_csr(cp_problem* P, int* rows, int* cols)
{
    if (!P || !rows || !cols) return cp::fail(CP_ERR_INVALID, "null argument");
    cp::Adjacency adj;
    adj.ptr = P->adj_ptr;
    adj.node = P->adj_node;
    cp::export_csr(adj, P->nn, rows, cols);
    return CP_OK;
}

int cp_internal_force_noalloc(cp_problem* P, double* f_host)
{
    if (!P || !f_host) return cp::fail(CP_ERR_INVALID, "null argument");
    if ((size_t)P->nn * sizeof(double2) > kScratchBytes)
        return cp::failf(CP_ERR_INVALID,
                         "cp_internal_force_noalloc: %d nodes exceed the %zu-byte device staging (at most %zu nodes)",
                         P->nn, kScratchBytes, kScratchBytes / sizeof(double2));
    if (!P->ctx || !P->ctx->scratch) return cp::fail(CP_ERR_STATE, "cp_internal_force_noalloc: no staging address");
    // every entry of the staged force is written by the kernel: nothing to clear
    double2* f_dev = static_cast<double2*>(P->ctx->scratch);
    cpk::ResArgs ra{};
    ra.incPtr = P->inc_ptr;
    ra.inc = P->inc;
    ra.conn = P->conn;
    ra.xy = P->xy;
    ra.ecnu = P->ecnu;
    ra.u = P->u;
    ra.r64 = f_dev;
    ra.nn = P->nn;
    ra.nblocks = P->nbn;
    ra.plane = P->M.plane;
    ra.h = P->M.h;
    if (int s = cp::launch_residual(P, ra, P->stream)) return s;
    CP_HIP(hipMemcpyAsync(f_host, f_dev, sizeof(double2) * P->nn, hipMemcpyDeviceToHost, P->stream));
    CP_HIP(hipStreamSynchronize(P->stream));
    return CP_OK;
}

int cp_problem_get_values(cp_problem* P, double* vals)
{
    if (!P || !vals) return cp::fail(CP_ERR_INVALID, "null argument");
    if (!P->have_fp64_values) return cp::fail(CP_ERR_STATE, "no FP64 assembly has run yet");
    std::vector<double2> a(P->nblk_stored), b(P->nblk_stored);
    std::vector<int2> sl(P->nslices);
    CP_HIP(hipMemcpyAsync(a.data(), P->va, sizeof(double2) * P->nblk_stored, hipMemcpyDeviceToHost, P->stream));
    CP_HIP(hipMemcpyAsync(b.data(), P->vb, sizeof(double2) * P->nblk_stored, hipMemcpyDeviceToHost, P->stream));
    CP_HIP(hipMemcpyAsync(sl.data(), P->slice, sizeof(int2) * P->nslices, hipMemcpyDeviceToHost, P->stream));
    CP_HIP(hipStreamSynchronize(P->stream));
    long long row0 = 0;
    for (int n = 0; n < P->nn; ++n) {
        const int deg = P->adj_ptr[n + 1] - P->adj_ptr[n];
        const long long row1 = row0 + 2 * deg;
        const size_t base = (size_t)sl[n >> 6].x + (n & 63);
        for (int k = 0; k < deg; ++k) {
            const size_t q = base + (size_t)k * cpk::kWave;
            vals[row0 + 2 * k] = a[q].x;
            vals[row0 + 2 * k + 1] = a[q].y;
            vals[row1 + 2 * k] = b[q].x;
            vals[row1 + 2 * k + 1] = b[q].y;
        }
        row0 = row1 + 2 * deg;
    }
    return CP_OK;
}

int cp_set_load(cp_problem* P, double factor)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    if (int s = cp::set_load(P, factor, 0)) return s;
    CP_HIP(hipStreamSynchronize(P->stream));
    return CP_OK;
}

int cp_advance_load(cp_problem* P, double factor, int predictor)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    if (int s = cp::advance_load(P, factor, predictor)) return s;
    CP_HIP(hipStreamSynchronize(P->stream));
    return CP_OK;
}

int cp_reset_state(cp_problem* P)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    if (int s = cp::reset_state(P)) return s;
    CP_HIP(hipStreamSynchronize(P->stream));
    return CP_OK;
}

int cp_assemble(cp_problem* P, int precision)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    if (int s = cp::launch_assemble(P, precision)) return s;
    CP_HIP(hipStreamSynchronize(P->stream));
    return CP_OK;
}

int cp_cg(cp_problem* P, double tol, int maxit, int precision, double inner_tol, long long* iters, double* relres)
{
    if (!P || !iters) return cp::fail(CP_ERR_INVALID, "null argument");
    if (int s = cp::problem_alloc_solver(P, precision)) return s;
    double ms = 0.;
    if (maxit <= 0) maxit = 1000000000;
    if (precision == CP_FP64) return cp::cg_fp64(P, tol, maxit, iters, relres, &ms);
    long long refs = 0;
    return cp::cg_mixed(P, tol, maxit, inner_tol > 0 ? inner_tol : 1e-3, iters, &refs, relres, &ms);
}

int cp_return_map(cp_problem* P, double* rel_err)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    hipLaunchKernelGGL(cpk::k_rm_reset, dim3(1), dim3(1), 0, P->stream, P->rms);
    if (int s = cp::launch_return_map(P, 0, 0, 0, 1, 1, 0., 1)) return s;
    CP_HIP(hipMemcpyAsync(P->h_rm, P->rms, sizeof(cpk::RmScal), hipMemcpyDeviceToHost, P->stream));
    CP_HIP(hipStreamSynchronize(P->stream));
    if (rel_err) *rel_err = P->h_rm->rel_err;
    return P->h_rm->status == 2 ? cp::fail(CP_ERR_NAN, "NaN in the secant update") : CP_OK;
}

int cp_plastic_strains(cp_problem* P)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    hipLaunchKernelGGL(cpk::k_plastic_strains, dim3(P->nbe), dim3(cpk::kBlock), 0, P->stream, P->F, P->psi,
                       (const double2*)P->ecnu, P->M.E, P->M.nu, P->ne, P->nbe);
    CP_HIP(hipGetLastError());
    CP_HIP(hipStreamSynchronize(P->stream));
    return CP_OK;
}

int cp_internal_force(cp_problem* P, double* f_host)
{
    if (!P || !f_host) return cp::fail(CP_ERR_INVALID, "null argument");
    if (int s = cp::problem_alloc_solver(P, CP_MIXED)) return s;
    cpk::ResArgs ra{};
    ra.incPtr = P->inc_ptr;
    ra.inc = P->inc;
    ra.conn = P->conn;
    ra.xy = P->xy;
    ra.ecnu = P->ecnu;
    ra.u = P->u;
    ra.r64 = P->res64;
    ra.nn = P->nn;
    ra.nblocks = P->nbn;
    ra.plane = P->M.plane;
    ra.h = P->M.h;
    if (int s = cp::launch_residual(P, ra, P->stream)) return s;
    CP_HIP(hipMemcpyAsync(f_host, P->res64, sizeof(double2) * P->nn, hipMemcpyDeviceToHost, P->stream));
    CP_HIP(hipStreamSynchronize(P->stream));
    return CP_OK;
}

static cp_options defaults_if_null(const cp_options* o)
{
    cp_options d{};
    d.cg_tol = 1e-7;
    d.cg_maxit = 0;
    d.tol = 1e-5;
    d.max_outer = 1000;
    d.precision = CP_FP64;
    d.inner_tol = 1e-3;
    return o ? *o : d;
}

int cp_solve(cp_problem* P, const cp_options* opt, cp_stats* stats)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    cp_options o = defaults_if_null(opt);
    return cp::solve_level(P, &o, 0, stats);
}

int cp_solve_elastic(cp_problem* P, const cp_options* opt, cp_stats* stats)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    cp_options o = defaults_if_null(opt);
    return cp::solve_level(P, &o, 1, stats);
}

int cp_solve_steps(cp_problem* P, int nsteps, double final_factor, const cp_options* opt, cp_stats* per_step)
{
    if (!P || nsteps < 1) return cp::fail(CP_ERR_INVALID, "cp_solve_steps: bad arguments");
    cp_options o = defaults_if_null(opt);
    const bool warm0 = o.warm != 0;
    for (int n = 1; n <= nsteps; ++n) {
        const double lam = final_factor * double(n) / double(nsteps);
        o.warm = (n > 1 || warm0) ? 1 : 0;
        if (o.warm) {
            if (int s = cp::advance_load(P, lam, o.predictor)) return s;
        } else {
            P->load = lam;
        }
        cp_stats st{};
        int s = cp::solve_level(P, &o, 0, &st);
        if (per_step) per_step[n - 1] = st;
        if (s) return s;
    }
    return CP_OK;
}

int cp_download(cp_problem* P, int field, double* host)
{
    if (!P || !host) return cp::fail(CP_ERR_INVALID, "null argument");
    hipStream_t st = P->stream;
    if (field == CP_F_U || field == CP_F_RHS) {
        CP_HIP(hipMemcpyAsync(host, field == CP_F_U ? P->u : P->b, sizeof(double2) * P->nn, hipMemcpyDeviceToHost, st));
    } else if (field == CP_F_EC || field == CP_F_NUC) {
        std::vector<double2> t(P->ne);
        CP_HIP(hipMemcpyAsync(t.data(), P->ecnu, sizeof(double2) * P->ne, hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
        for (int e = 0; e < P->ne; ++e) host[e] = field == CP_F_EC ? t[e].x : t[e].y;
        return CP_OK;
    } else if (field == CP_F_PSI) {
        CP_HIP(hipMemcpyAsync(host, P->psi, sizeof(double) * P->ne, hipMemcpyDeviceToHost, st));
    } else if (field >= 0 && field < CP_NFIELDS) {
        CP_HIP(hipMemcpyAsync(host, P->F + (size_t)field * P->ne, sizeof(double) * P->ne, hipMemcpyDeviceToHost, st));
    } else {
        return cp::failf(CP_ERR_INVALID, "unknown field %d", field);
    }
    CP_HIP(hipStreamSynchronize(st));
    return CP_OK;
}

int cp_upload_state(cp_problem* P, const double* u, const double* psi)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    hipStream_t st = P->stream;
    if (u) CP_HIP(hipMemcpyAsync(P->u, u, sizeof(double2) * P->nn, hipMemcpyHostToDevice, st));
    std::vector<double2> en;
    if (psi) {
        en.resize(P->ne);
        for (int e = 0; e < P->ne; ++e) {
            if (!(psi[e] > 0)) return cp::fail(CP_ERR_INVALID, "psi must be positive");
            en[e] = make_double2(P->M.E / psi[e], 0.5 * (1. - (1. - 2. * P->M.nu) / psi[e]));
        }
        CP_HIP(hipMemcpyAsync(P->psi, psi, sizeof(double) * P->ne, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(P->ecnu, en.data(), sizeof(double2) * P->ne, hipMemcpyHostToDevice, st));
    }
    CP_HIP(hipStreamSynchronize(st));
    return CP_OK;
}

int cp_last_cg_kernel_ms(cp_problem* P, double* ms)
{
    if (!P || !ms) return cp::fail(CP_ERR_INVALID, "null argument");
    *ms = P->last_cg_ms_per_iter;
    return CP_OK;
}

int cp_problem_layout(cp_problem* P, long long* nblk_stored, long long* nblk, int* max_width)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    if (nblk_stored) *nblk_stored = P->nblk_stored;
    if (nblk) *nblk = P->nblk;
    if (max_width) *max_width = P->max_w;
    return CP_OK;
}

int cp_bench_spmv(cp_problem* P, int precision, int reps, double* ms, double* bytes_stored, double* bytes_algo)
{
    if (!P || reps < 1) return cp::fail(CP_ERR_INVALID, "bad arguments");
    if (int s = cp::launch_assemble(P, precision)) return s;
    hipStream_t st = P->stream;
    const dim3 g(P->nbn), b(cpk::kBlock);
    const double s = precision == CP_FP64 ? 8. : 4.;
    CP_HIP(hipEventRecord(P->ev[4], st));
    if (precision == CP_FP64) {
        auto a = P->cg64_args();
        for (int i = 0; i < reps; ++i)
            hipLaunchKernelGGL(cpk::k_spmv<double>, g, b, 0, st, a, (const double2*)P->u, P->w64[i & 1]);
    } else {
        auto a = P->cg32_args();
        CP_HIP(hipMemsetAsync(P->x32, 0, sizeof(float2) * P->nn, st));
        for (int i = 0; i < reps; ++i)
            hipLaunchKernelGGL(cpk::k_spmv<float>, g, b, 0, st, a, (const float2*)P->x32, P->w32[i & 1]);
    }
    CP_HIP(hipGetLastError());
    CP_HIP(hipEventRecord(P->ev[5], st));
    CP_HIP(hipEventSynchronize(P->ev[5]));
    float t = 0.f;
    CP_HIP(hipEventElapsedTime(&t, P->ev[4], P->ev[5]));
    if (ms) *ms = t / reps;
    const double vec = 2. * s * P->nn;   // x read once, y written once
    if (bytes_stored) *bytes_stored = P->nblk_stored * (4. * s + 4.) + 8. * P->nslices + 2. * vec;
    if (bytes_algo) *bytes_algo = P->nblk * (4. * s + 4.) + 8. * P->nslices + 2. * vec;
    return CP_OK;
}

}  // extern "C"

namespace {
// Plain device copy, four float4 per thread, for the achievable HBM
// bandwidth that the bandwidth-bound kernels are compared with.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i + 256 * k < n) v[k] = src[i + 256 * k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i + 256 * k < n) dst[i + 256 * k] = v[k];
}

// Read-only pass over the same buffer: every element is loaded and summed;
// the store happens only for a sum the zeroed input never gives, so the
// loads stay and nothing is written.
__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i + 256 * k < n) {
            const float4 v = src[i + 256 * k];
            acc += (v.x + v.y) + (v.z + v.w);
        }
    if (acc == 1.0f) dst[i] = make_float4(acc, acc, acc, acc);
}
}  // namespace

extern "C" {

int cp_bench_copy(cp_ctx* ctx, size_t bytes, int reps, double* ms, double* ms_read)
{
    if (!ctx || reps < 1 || bytes < 16 || bytes % 16 != 0 || !ms)
        return cp::fail(CP_ERR_INVALID, "cp_bench_copy: bad arguments (bytes a positive multiple of 16)");
    CP_HIP(hipSetDevice(ctx->device));
    const size_t n = bytes / 16;
    if ((n + 1023) / 1024 > 0x7fffffff) return cp::fail(CP_ERR_INVALID, "cp_bench_copy: too large");
    float4 *src = nullptr, *dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    int status = CP_OK;
    float t = 0.f;
    if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess ||
        hipMemsetAsync(src, 0, bytes, st) != hipSuccess) {
        status = cp::fail(CP_ERR_NOMEM, "cp_bench_copy: allocation failed");
    } else {
        const dim3 g((unsigned)((n + 1023) / 1024));
        hipLaunchKernelGGL(k_copy, g, dim3(256), 0, st, src, dst, n);   // warm-up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, g, dim3(256), 0, st, src, dst, n);
        (void)hipEventRecord(e1, st);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_read, g, dim3(256), 0, st, src, dst, n);
        (void)hipEventRecord(e2, st);
        float t2 = 0.f;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e2) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess || hipEventElapsedTime(&t2, e1, e2) != hipSuccess) {
            status = cp::fail(CP_ERR_DEVICE, "cp_bench_copy: launch failed");
        } else {
            *ms = t / reps;
            if (ms_read) *ms_read = t2 / reps;
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e2) (void)hipEventDestroy(e2);
    if (st) (void)hipStreamDestroy(st);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return status;
}

int cp_bench_assemble(cp_problem* P, int precision, int reps, double* ms)
{
    if (!P || reps < 1 || (precision != CP_FP64 && precision != CP_MIXED))
        return cp::fail(CP_ERR_INVALID, "bad arguments");
    const int pre = P->precond;
    P->precond = 0;   // the element stiffness and its assembly only
    if (int s = cp::launch_assemble(P, precision)) {
        P->precond = pre;
        return s;
    }
    CP_HIP(hipEventRecord(P->ev[4], P->stream));
    for (int i = 0; i < reps; ++i)
        if (int s = cp::launch_assemble(P, precision)) {
            P->precond = pre;
            return s;
        }
    CP_HIP(hipEventRecord(P->ev[5], P->stream));
    CP_HIP(hipEventSynchronize(P->ev[5]));
    P->precond = pre;
    float t = 0.f;
    CP_HIP(hipEventElapsedTime(&t, P->ev[4], P->ev[5]));
    if (ms) *ms = t / reps;
    return CP_OK;
}

int cp_bench_cg(cp_problem* P, int precision, int precond, int reps, double* ms, double* bytes_algo)
{
    if (!P || reps < 1 || (precision != CP_FP64 && precision != CP_MIXED) || precond < 0 || precond > 1)
        return cp::fail(CP_ERR_INVALID, "bad arguments");
    P->precond = precond ? 1 : 0;
    if (int s = cp::launch_assemble(P, precision)) return s;
    hipStream_t st = P->stream;
    const dim3 g(P->nbn), b(cpk::kBlock);
    const double s = precision == CP_FP64 ? 8. : 4.;
    if (precision == CP_FP64) {
        auto a = P->cg64_args();
        // tol 0 and a huge cap keep every launch live
        hipLaunchKernelGGL(cpk::k_cg_setup, dim3(1), dim3(1), 0, st, P->sc64, 1.0, 0.0, 2000000000);
        hipLaunchKernelGGL(cpk::k_cg_init1<double>, g, b, 0, st, a, 0);
        cp::launch_init2<double>(P->max_w, P->precond, a, g, st);
        CP_HIP(hipEventRecord(P->ev[4], st));
        cp::launch_iter_kernels_only<double>(P->max_w, P->precond, a, reps, g, st);
    } else {
        auto a = P->cg32_args();
        // right-hand side of the inner solve: the FP64 residual of the current u
        cpk::ResArgs ra{};
        ra.incPtr = P->inc_ptr;
        ra.inc = P->inc;
        ra.conn = P->conn;
        ra.xy = P->xy;
        ra.ecnu = P->ecnu;
        ra.u = P->u;
        ra.b = P->b;
        ra.mask = P->mask;
        ra.r64 = P->res64;
        ra.r32 = P->b32;
        ra.nn = P->nn;
        ra.nblocks = P->nbn;
        ra.plane = P->M.plane;
        ra.h = P->M.h;
        if (int s = cp::launch_residual(P, ra, st)) return s;
        hipLaunchKernelGGL(cpk::k_cg_setup, dim3(1), dim3(1), 0, st, P->sc32, 1.0, 0.0, 2000000000);
        hipLaunchKernelGGL(cpk::k_cg_init1<float>, g, b, 0, st, a, 1);
        cp::launch_init2<float>(P->max_w, P->precond, a, g, st);
        CP_HIP(hipEventRecord(P->ev[4], st));
        cp::launch_iter_kernels_only<float>(P->max_w, P->precond, a, reps, g, st);
    }
    CP_HIP(hipGetLastError());
    CP_HIP(hipEventRecord(P->ev[5], st));
    CP_HIP(hipEventSynchronize(P->ev[5]));
    float t = 0.f;
    CP_HIP(hipEventElapsedTime(&t, P->ev[4], P->ev[5]));
    if (ms) *ms = t / reps;
    if (P->nbn <= cpk::kFusedMaxBlocks) {
        // every timed launch must have iterated (tolerance 0: only a NaN stops)
        const cpk::CgScal* last = (precision == CP_FP64 ? P->sc64 : P->sc32) + ((reps - 1) & 1);
        CP_HIP(hipMemcpy(P->h_cg, last, sizeof(cpk::CgScal), hipMemcpyDeviceToHost));
        if (P->h_cg->done) return cp::fail(CP_ERR_NAN, "cp_bench_cg: the iteration stopped early");
    }
    // matrix once + 5 vector reads (r, w, s_prev, p, x) + 5 writes + mask
    // (+ the block-Jacobi inverse, 4 values per node, when preconditioned)
    if (bytes_algo)
        *bytes_algo = P->nblk * (4. * s + 4.) + 8. * P->nslices + 10. * 2. * s * P->nn + 2. * P->nn +
                      (P->precond ? 4. * s * P->nn : 0.);
    return CP_OK;
}

}  // extern "C"
