// cp_mg.hip -- host side of the geometric multigrid preconditioner
// (precond = 2): lattice set-up, level hierarchy, set-up after each
// assembly, the V-cycle and CG preconditioned by it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <type_traits>
#include <vector>

#include "cp_mg.h"
#include "cp_problem.h"

namespace cp {

struct MgLevel {
    int n1 = 0, n2 = 0, nn = 0, nb = 0, nslices = 0;
    int co1 = 1, co2 = 1;    // directions halved from this level to the next
    int seg = 0;             // 1: segment smoother along direction 1, else 2x2 block Jacobi
    int2* slice = nullptr;   // levels >= 1 (level 0 uses the problem's matrix)
    int* col = nullptr;
    uchar2* mask = nullptr;
    double2 *va = nullptr, *vb = nullptr;
    float4* vf = nullptr;
    double4* minv64 = nullptr;
    float4* minv32 = nullptr;
    float4* vsym32 = nullptr;   // launched coarse levels of an open lattice: FP32 symmetric lattice layout
    double* lam = nullptr;   // [0] FP64 operator, [1] FP32 operator
    void* vec[2][6] = {};    // per precision: b, xa, xb, da, db, res
};

struct MgState {
    int nlev = 0;
    int per1 = 0;            // the lattice is periodic in direction 1
    std::vector<MgLevel> L;
    int* lg0 = nullptr;      // node -> lattice index of level 0 (nullptr: identity)
    int* node0 = nullptr;    // lattice index -> node of level 0 (nullptr: identity)
    signed char* slot0 = nullptr;   // level 0: node x 9 lattice offsets -> position in its row
    // level 0 in FP32 also in the symmetric lattice layout (nodes numbered
    // lattice point by lattice point, open lattice, 4-node quads)
    bool sym_ok = false;
    float4* vsym32 = nullptr;
    // an 8-node mesh: level 0 is the mesh itself, level 1 its corner
    // lattice; the transfer between them (cp_mg.h, MgGrid)
    bool q8 = false;
    int2* q8_par = nullptr;
    int* q8_ch = nullptr;
    double* dense = nullptr; // coarsest operator, inverted in place
    int nd = 0;
    bool alloc[2] = {false, false};
    std::vector<void*> blocks;
    // one CG iteration (update, V-cycle, dots, step) and the start of a
    // solve (clear, V-cycle, dots, scalars) captured per precision
    hipGraphExec_t iter[2] = {nullptr, nullptr};
    hipGraphExec_t start[2] = {nullptr, nullptr};
    bool use_graphs = true;
    // levels lt.. (at most kTailThreads nodes each) run in one workgroup
    int lt = 0;
    // coarse levels rebuilt every `rebuild` set-ups (1: every assembly; a
    // plate ramp runs 15 % faster at 8 than at 1 with 3 % more CG iterations);
    // set-ups since the last rebuild, per precision
    int rebuild = 8;
    int since[2] = {0, 0};
    bool built[2] = {false, false};
    // level 0's symmetric layout written by the assembly itself (the next
    // set-up skips k_mg_sym_build)
    bool sym_assembled = false;
    void* tail[2] = {nullptr, nullptr};
    // the tail image (cp_mg.h, TailImage) per precision
    cpk::TailImage tail_img[2] = {};

    template <typename X>
    int get(X** p, size_t count)
    {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(X)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(CP_ERR_NOMEM, "multigrid: device allocation failed");
        }
        blocks.push_back(q);
        *p = reinterpret_cast<X*>(q);
        return CP_OK;
    }
    ~MgState()
    {
        for (hipGraphExec_t& e : iter)
            if (e) (void)hipGraphExecDestroy(e);
        for (hipGraphExec_t& e : start)
            if (e) (void)hipGraphExecDestroy(e);
        for (void* q : blocks) (void)hipFree(q);
    }
};

void mg_destroy(cp_problem* P)
{
    delete P->mg;
    P->mg = nullptr;
}

static int blocks_of(int n) { return (n + cpk::kBlock - 1) / cpk::kBlock; }

constexpr double kSegFrac = 0.05;   // share of stretched cells that selects the segment smoother
// Largest coarse segment level smoothed with Chebyshev degree 2 in FP64.
constexpr int kDeg2Nodes = 65536;
// Smallest finest segment level smoothed with Chebyshev degree 2.
constexpr int kDeg2Finest = 1 << 20;

// Per-precision value and vector arrays, allocated on first use.
static int mg_alloc(cp_problem* P, int prec)
{
    MgState& G = *P->mg;
    const int pi = prec == CP_FP64 ? 0 : 1;
    if (G.alloc[pi]) return CP_OK;
    for (int l = 0; l < G.nlev; ++l) {
        MgLevel& L = G.L[l];
        const size_t stored = (size_t)L.nslices * cpk::kWave * 9;
        if (l > 0) {
            if (prec == CP_FP64) {
                if (int s = G.get(&L.va, stored)) return s;
                if (int s = G.get(&L.vb, stored)) return s;
                if (int s = G.get(&L.minv64, L.nn)) return s;
            } else {
                if (int s = G.get(&L.vf, stored)) return s;
                if (int s = G.get(&L.minv32, L.nn)) return s;
                if (!G.per1 && l < G.lt)
                    if (int s = G.get(&L.vsym32, (size_t)L.nslices * cpk::kWave * cpk::kSymSlots)) return s;
            }
        }
        if (l == 0 && prec != CP_FP64 && G.sym_ok)
            if (int s = G.get(&G.vsym32, (size_t)L.nslices * cpk::kWave * cpk::kSymSlots)) return s;
        const size_t vb = prec == CP_FP64 ? sizeof(double2) : sizeof(float2);
        for (int k = 0; k < 6; ++k) {
            if (l == 0 && k == 0) continue;   // level 0's right-hand side is the CG residual
            char* q = nullptr;
            if (int s = G.get(&q, vb * L.nn)) return s;
            L.vec[pi][k] = q;
        }
    }
    G.alloc[pi] = true;
    return CP_OK;
}

template <typename T>
static cpk::CgArgs<T> level_args(cp_problem* P, int l);

// LDS bytes of one tail level's six vectors (16-byte aligned).
static int tail_lds_bytes(int nn, size_t vbytes) { return (int)((6 * (size_t)nn * vbytes + 15) / 16 * 16); }

// Device table of the coarse-tail levels for precision T (fixed addresses).
template <typename T>
static int mg_tail_table(cp_problem* P)
{
    using V = typename cpk::V2<T>::type;
    MgState& G = *P->mg;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    if (G.tail[pi]) return CP_OK;
    std::vector<cpk::TailLevel<T>> tab;
    int lds = 0;
    for (int l = G.lt; l < G.nlev; ++l) {
        MgLevel& L = G.L[l];
        cpk::TailLevel<T> t{};
        t.lds = lds;
        lds += tail_lds_bytes(L.nn, sizeof(V));
        t.a = level_args<T>(P, l);
        t.v.b = (const V*)L.vec[pi][0];
        t.v.xa = (V*)L.vec[pi][1];
        t.v.xb = (V*)L.vec[pi][2];
        t.v.da = (V*)L.vec[pi][3];
        t.v.db = (V*)L.vec[pi][4];
        t.v.res = (V*)L.vec[pi][5];
        t.v.lam = L.lam + pi * cpk::kLamSlots;
        if (l + 1 < G.nlev) {
            t.grid.n1f = L.n1;
            t.grid.n2f = L.n2;
            t.grid.n1c = G.L[l + 1].n1;
            t.grid.n2c = G.L[l + 1].n2;
            t.grid.co1 = L.co1;
            t.grid.co2 = L.co2;
            t.grid.per1 = G.per1;
        }
        t.mat_lds = t.minv_lds = t.mask_lds = -1;
        tab.push_back(t);
    }
    if (lds > cpk::kTailVecLds) return fail(CP_ERR_INVALID, "multigrid: coarse tail exceeds its LDS");
    // the rest of the LDS (the tail image): copies of every level's mask
    // (at most two bytes per node of the vectors' 96), then of the pivot
    // inverses, then of the rows, coarsest level first, where they fit; the
    // coarsest level is solved with its inverse and needs only its mask
    const int nt = (int)tab.size();
    const int budget = cpk::tail_lds_budget<T>(), img_lo = lds;
    for (int i = nt - 1; i >= 0; --i) {
        tab[i].mask_lds = lds;
        lds += cpk::tail_mask_bytes(tab[i].a.nn);
    }
    if (lds > budget) return fail(CP_ERR_INVALID, "multigrid: coarse tail masks exceed its LDS");
    for (int i = nt - 2; i >= 0; --i) {
        const int need = cpk::tail_minv_bytes<T>(tab[i].a.nn);
        if (lds + need > budget) break;
        tab[i].minv_lds = lds;
        lds += need;
    }
    for (int i = nt - 2; i >= 0; --i) {
        const int need = cpk::tail_rows_bytes<T>(tab[i].a.nn);
        if (lds + need > budget) break;
        tab[i].mat_lds = lds;
        lds += need;
    }
    for (auto& t : tab)
        t.where = t.minv_lds < 0 ? cpk::kTailGlobal : (t.mat_lds < 0 ? cpk::kTailMinvLds : cpk::kTailAllLds);
    if (nt > cpk::kTailMaxLevels) return fail(CP_ERR_INVALID, "multigrid: too many coarse tail levels");
    cpk::TailLevel<T>* d = nullptr;
    if (int s = G.get(&d, tab.size())) return s;
    CP_HIP(hipMemcpy(d, tab.data(), sizeof(cpk::TailLevel<T>) * tab.size(), hipMemcpyHostToDevice));
    float4* img = nullptr;
    if (int s = G.get(&img, (size_t)(lds - img_lo) / 16)) return s;
    G.tail[pi] = d;
    G.tail_img[pi] = cpk::TailImage{img, (lds - img_lo) / 16, img_lo};
    return CP_OK;
}

template <>
cpk::CgArgs<double> level_args<double>(cp_problem* P, int l)
{
    if (l == 0) return P->cg64_args();
    const MgLevel& L = P->mg->L[l];
    cpk::CgArgs<double> a{};
    a.col = L.col;
    a.slice = L.slice;
    a.va = L.va;
    a.vb = L.vb;
    a.mask = L.mask;
    a.minv = L.minv64;
    a.nn = L.nn;
    a.nblocks = L.nb;
    a.lat_n1 = L.n1;
    a.lat_n2 = L.n2;
    a.lat_per1 = P->mg->per1;
    return a;
}

template <>
cpk::CgArgs<float> level_args<float>(cp_problem* P, int l)
{
    if (l == 0) {
        cpk::CgArgs<float> a = P->cg32_args();
        const MgState& G = *P->mg;
        if (G.vsym32) {
            a.vsym = G.vsym32;
            a.lat_n1 = G.L[0].n1;
            a.lat_n2 = G.L[0].n2;
        }
        return a;
    }
    const MgLevel& L = P->mg->L[l];
    cpk::CgArgs<float> a{};
    a.col = L.col;
    a.slice = L.slice;
    a.va = L.vf;
    a.vb = nullptr;
    a.mask = L.mask;
    a.minv = L.minv32;
    a.nn = L.nn;
    a.nblocks = L.nb;
    a.lat_n1 = L.n1;
    a.lat_n2 = L.n2;
    a.lat_per1 = P->mg->per1;
    a.vsym = L.vsym32;
    return a;
}

template <typename T>
static typename cpk::V4<T>::type* level_minv(cp_problem* P, int l);
template <>
double4* level_minv<double>(cp_problem* P, int l)
{
    return l == 0 ? P->minv64 : P->mg->L[l].minv64;
}
template <>
float4* level_minv<float>(cp_problem* P, int l)
{
    return l == 0 ? P->minv32 : P->mg->L[l].minv32;
}

// Coarse levels of at least kWideTileNodes nodes in the FP32 symmetric
// lattice layout smooth in the finest level's 16 x 32 tiles (row template
// kSymW), smaller ones in 16 x 16 tiles (kSymW16).
constexpr int kWideTileNodes = 1 << 18;
template <typename T>
static bool wide_tiles(const MgState& G, int l)
{
    return std::is_same<T, float>::value && l > 0 && G.L[l].vsym32 && G.L[l].nn >= kWideTileNodes;
}

template <typename T>
static cpk::SegGrid seg_grid(const MgState& G, int l)
{
    const MgLevel& L = G.L[l];
    cpk::SegGrid g{};
    g.n1 = L.n1;
    g.n2 = L.n2;
    g.per1 = G.per1;
    const int cols = l == 0 || wide_tiles<T>(G, l) ? cpk::seg_cols<0>() : cpk::seg_cols<cpk::kLatW>();
    g.nt2 = (L.n2 + cols - 1) / cols;
    g.ntiles = (L.n1 + cpk::kSeg - 1) / cpk::kSeg * g.nt2;
    g.node_of = l == 0 ? G.node0 : nullptr;
    g.slot_pos = l == 0 ? G.slot0 : nullptr;
    return g;
}

template <typename T>
static int mg_setup_t(cp_problem* P, int nlev)
{
    MgState& G = *P->mg;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    hipStream_t st = P->stream;
    if constexpr (std::is_same<T, float>::value)
        if (G.vsym32 && !G.sym_assembled)
            hipLaunchKernelGGL(cpk::k_mg_sym_build, dim3(P->nbn), dim3(cpk::kBlock), 0, st, P->cg32_args(),
                               (const signed char*)G.slot0, G.vsym32);
    G.sym_assembled = false;
    for (int l = 0; l < nlev; ++l) {
        const MgLevel& L = G.L[l];
        cpk::CgArgs<T> a = level_args<T>(P, l);
        if (l > 0) {
            const MgLevel& F = G.L[l - 1];
            cpk::GalArgs<T, T> g{};
            g.f = level_args<T>(P, l - 1);
            g.lg = l == 1 ? G.lg0 : nullptr;
            g.node_of = l == 1 ? G.node0 : nullptr;
            g.slot_pos = l == 1 ? G.slot0 : nullptr;
            g.n1f = F.n1;
            g.n2f = F.n2;
            g.co1 = F.co1;
            g.co2 = F.co2;
            g.per1 = G.per1;
            g.n1c = L.n1;
            g.n2c = L.n2;
            g.nnc = L.nn;
            g.nbc = blocks_of(9 * L.nn);
            g.slice_c = L.slice;
            if (pi == 0) {
                g.va_c = L.va;
                g.vb_c = L.vb;
            } else {
                g.vf_c = L.vf;
            }
            g.mask_c = L.mask;
            if (l == 1 && G.q8) {
                g.q8_par = G.q8_par;
                g.q8_ch = G.q8_ch;
                hipLaunchKernelGGL((cpk::k_mg_galerkin_q8<T, T>), dim3(g.nbc), dim3(cpk::kBlock), 0, st, g);
            } else {
                const int nt1 = (L.n1 + cpk::kGalT - 1) / cpk::kGalT, nt2 = (L.n2 + cpk::kGalT - 1) / cpk::kGalT;
                hipLaunchKernelGGL((cpk::k_mg_galerkin<T, T>), dim3(nt1 * nt2), dim3(cpk::kGalThreads), 0, st, g,
                                   nt2, nt1 * nt2);
            }
            if constexpr (std::is_same<T, float>::value)
                if (L.vsym32)
                    hipLaunchKernelGGL(cpk::k_mg_sym_from_lat, dim3(L.nb), dim3(cpk::kBlock), 0, st,
                                       (const float4*)L.vf, L.nn, L.vsym32);
            a = level_args<T>(P, l);
        }
        if (l < G.nlev - 1 && L.seg) {
            const int cols = (L.n1 + cpk::kSeg - 1) / cpk::kSeg * L.n2;
            hipLaunchKernelGGL(cpk::k_mg_seg_setup<T>, dim3(blocks_of(cols)), dim3(cpk::kBlock), 0, st, a,
                               seg_grid<T>(G, l), level_minv<T>(P, l), L.lam + pi * cpk::kLamSlots);
        } else if (l < G.nlev - 1) {
            hipLaunchKernelGGL(cpk::k_mg_smooth_setup<T>, dim3(L.nb), dim3(cpk::kBlock), 0, st, a,
                               level_minv<T>(P, l), P->partials);
            hipLaunchKernelGGL(cpk::k_mg_lam, dim3(1), dim3(cpk::kBlock), 0, st, (const double*)P->partials, L.nb,
                               L.lam + pi * cpk::kLamSlots);
        } else {
            hipLaunchKernelGGL(cpk::k_mg_dense<T>, dim3(L.nb), dim3(cpk::kBlock), 0, st, a, G.dense, G.nd);
            hipLaunchKernelGGL(cpk::k_mg_invert, dim3(1), dim3(1024), 0, st, G.dense, G.nd);
        }
    }
    // the tail levels changed: their LDS image
    if (nlev == G.nlev && G.tail[pi] && G.tail_img[pi].n16 > 0)
        hipLaunchKernelGGL(cpk::k_mg_tail_image<T>, dim3(G.nlev - G.lt), dim3(cpk::kBlock), 0, st,
                           (const cpk::TailLevel<T>*)G.tail[pi], (char*)G.tail_img[pi].img, G.tail_img[pi].lo);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

// Whether the coming set-up rebuilds the coarse levels (mg_setup).
static bool mg_rebuild_due(const MgState& G, int pi, int full)
{
    return full || !G.built[pi] || G.since[pi] + 1 >= G.rebuild;
}

int mg_assemble_direct(cp_problem* P, int precision, int full, const cpk::AsmArgs& a, int* done)
{
    *done = 0;
    if (!P->mg || precision == CP_FP64 || P->precond != 2) return CP_OK;
    MgState& G = *P->mg;
    if (!G.sym_ok || !G.vsym32 || !G.L[0].seg || G.L[0].n2 < 3 || mg_rebuild_due(G, 1, full)) return CP_OK;
    hipLaunchKernelGGL(cpk::k_assemble_sym, dim3(P->nbn), dim3(cpk::kBlock), 0, P->stream, a, G.L[0].n2, G.vsym32);
    CP_HIP(hipGetLastError());
    G.sym_assembled = true;
    *done = 1;
    return CP_OK;
}

int mg_setup(cp_problem* P, int precision, int full)
{
    if (!P->mg) return fail(CP_ERR_STATE, "multigrid needs the mesh lattice: call cp_problem_set_grid first");
    if (int s = mg_alloc(P, precision)) return s;
    if (int s = precision == CP_FP64 ? mg_tail_table<double>(P) : mg_tail_table<float>(P)) return s;
    MgState& G = *P->mg;
    const int pi = precision == CP_FP64 ? 0 : 1;
    // a reused coarse hierarchy stays a symmetric positive definite
    // preconditioner (each level's smoother is set up from that level's own
    // operator); only its quality depends on how far the moduli have moved
    const bool all = mg_rebuild_due(G, pi, full);
    if (all) {
        G.built[pi] = true;
        G.since[pi] = 0;
    } else {
        ++G.since[pi];
    }
    const int nl = all ? G.nlev : 1;
    return precision == CP_FP64 ? mg_setup_t<double>(P, nl) : mg_setup_t<float>(P, nl);
}

// ------------------------------------------------------------ V-cycle steps
//
// The V-cycle and the CG iteration are written once, against an emitter that
// takes each step with its key (its place in the iteration: section, level,
// step) and its launch size: Direct launches it at once on the problem's
// stream; BatchRec (serving) collects the steps of several problems and
// launches each key's steps of all of them together.

enum Sect : unsigned long long { SECT_UPDATE = 0, SECT_CHECK = 1, SECT_DOWN = 2, SECT_TAIL = 3, SECT_UP = 4,
                                 SECT_DOTS = 5, SECT_STEP = 6 };
constexpr unsigned long long step_key(Sect s, int level, int step)
{
    // the up-sweep runs from the coarse levels to the fine ones
    return (static_cast<unsigned long long>(s) << 40) |
           (static_cast<unsigned long long>(s == SECT_UP ? 65535 - level : level) << 8) | (unsigned)step;
}
constexpr int kStepRestrict = 16;   // restriction after a level's down steps

struct Direct {
    hipStream_t st;
    template <int W, int MODE, typename T>
    void smooth(unsigned long long, const cpk::StepOp<T>& o, int nblocks)
    {
        hipLaunchKernelGGL((cpk::k_mg_smooth<W, MODE, T>), dim3(nblocks), dim3(cpk::kBlock), 0, st, o.a, o.v, o.xi,
                           o.di, o.xo, o.dout, o.second, o.last);
    }
    template <int W, int MODE, typename T>
    void smooth_seg(unsigned long long, const cpk::StepOp<T>& o, int ntiles)
    {
        hipLaunchKernelGGL((cpk::k_mg_smooth_seg<W, MODE, T>), dim3(ntiles), dim3(cpk::kSeg * cpk::seg_cols<W>()), 0,
                           st, o.a, o.v, o.xi, o.di, o.xo, o.dout, o.second, o.last, o.g);
    }
    template <typename T>
    void restrict_(unsigned long long, const cpk::GridOp<T>& o, int nblocks)
    {
        hipLaunchKernelGGL(cpk::k_mg_restrict<T>, dim3(nblocks), dim3(cpk::kBlock), 0, st, o.gr, o.src, o.dst, o.mask,
                           o.nn, nblocks, o.sc);
    }
    template <typename T>
    void prolong(unsigned long long, const cpk::GridOp<T>& o, int nblocks)
    {
        hipLaunchKernelGGL(cpk::k_mg_prolong<T>, dim3(nblocks), dim3(cpk::kBlock), 0, st, o.gr, o.src, o.xin, o.dst,
                           o.mask, o.nn, nblocks, o.sc);
    }
    template <typename T>
    void tail(unsigned long long, const cpk::TailOp<T>& o)
    {
        hipLaunchKernelGGL(cpk::k_mg_tail<T>, dim3(1), dim3(cpk::kTailThreads), 0, st, o.lv, o.nl, o.Dinv, o.nd, o.sc,
                           o.im, o.rg, o.rres);
    }
    template <typename T>
    void update(unsigned long long, const cpk::CgOp<T>& o, int nblocks)
    {
        hipLaunchKernelGGL(cpk::k_mg_cg_update<T>, dim3(nblocks), dim3(cpk::kBlock), 0, st, o.a, o.u);
    }
    template <typename T>
    void check(unsigned long long, const cpk::CgOp<T>& o)
    {
        hipLaunchKernelGGL(cpk::k_mg_cg_check<T>, dim3(1), dim3(cpk::kBlock), 0, st, o.a);
    }
    template <int W, typename T>
    void dots(unsigned long long, const cpk::CgOp<T>& o, int nblocks)
    {
        hipLaunchKernelGGL((cpk::k_mg_cg_dots<W, T>), dim3(nblocks), dim3(cpk::kBlock), 0, st, o.a, o.u);
    }
    template <typename T>
    void step(unsigned long long, const cpk::CgOp<T>& o)
    {
        hipLaunchKernelGGL(cpk::k_cg_step<T>, dim3(1), dim3(cpk::kBlock), 0, st, o.a);
    }
};

template <int W, int MODE, typename T, typename E>
static void smooth(E& e, unsigned long long key, const cpk::CgArgs<T>& a, const cpk::MgVec<T>& v, const void* xi,
                   const void* di, void* xo, void* dout, int second, int last)
{
    using V = typename cpk::V2<T>::type;
    cpk::StepOp<T> o{};
    o.a = a;
    o.v = v;
    o.xi = (const V*)xi;
    o.di = (const V*)di;
    o.xo = (V*)xo;
    o.dout = (V*)dout;
    o.second = second;
    o.last = last;
    e.template smooth<W, MODE, T>(key, o, a.nblocks);
}

template <int W, typename T, typename E>
static void smooth_level_down(E& e, int l, const cpk::CgArgs<T>& a, const cpk::MgVec<T>& v)
{
    smooth<W, cpk::MG_PRE1, T>(e, step_key(SECT_DOWN, l, 0), a, v, nullptr, nullptr, v.xa, v.da, 0, 0);
    smooth<W, cpk::MG_CHEB, T>(e, step_key(SECT_DOWN, l, 1), a, v, v.xa, v.da, v.xb, v.db, 1, 0);
    smooth<W, cpk::MG_RESID, T>(e, step_key(SECT_DOWN, l, 2), a, v, v.xb, v.db, v.xa, nullptr, 0, 0);
}

template <int W, typename T, typename E>
static void smooth_level_up(E& e, int l, const cpk::CgArgs<T>& a, const cpk::MgVec<T>& v)
{
    smooth<W, cpk::MG_POST1, T>(e, step_key(SECT_UP, l, 1), a, v, v.xb, nullptr, nullptr, v.da, 0, 0);
    smooth<W, cpk::MG_CHEB, T>(e, step_key(SECT_UP, l, 2), a, v, v.xb, v.da, v.xa, v.db, 0, 0);
    smooth<W, cpk::MG_CHEB, T>(e, step_key(SECT_UP, l, 3), a, v, v.xa, v.db, v.xb, nullptr, 1, 1);
}

template <int W, int MODE, typename T, typename E>
static void smooth_seg(E& e, unsigned long long key, const cpk::CgArgs<T>& a, const cpk::MgVec<T>& v, const void* xi,
                       const void* di, void* xo, void* dout, int second, int last, const cpk::SegGrid& g)
{
    using V = typename cpk::V2<T>::type;
    cpk::StepOp<T> o{};
    o.a = a;
    o.v = v;
    o.xi = (const V*)xi;
    o.di = (const V*)di;
    o.xo = (V*)xo;
    o.dout = (V*)dout;
    o.second = second;
    o.last = last;
    o.g = g;
    e.template smooth_seg<W, MODE, T>(key, o, g.ntiles);
}

// The same steps with the segment smoother: x1 = c0 D^-1 b takes a launch
// of its own (D^-1 b at a neighbour is not local any more); on coarse
// levels that launch also restricts the finer level's residual into b
// (g.rres set).  Each step after the first writes the next iterate x + d
// beside d (the _X modes), so the step after it reads one vector at the
// neighbours, and the residual step writes the residual alone.
// Degree 2 stops the same recurrence one step earlier (the step-k
// coefficients do not depend on the degree): the pre-smoothed x then lands
// in xb, the prolongation writes xa, and the post-smoothing ends in xb.
// Buffers, degree 3: down x1 -> xa, x2 -> xb, x3 -> xa (d1 in da);
// up (from xb) x' -> xa, x'' -> res (consumed by then), x -> xb.
template <int W, typename T, typename E>
static void smooth_level_down_seg(E& e, int l, const cpk::CgArgs<T>& a, const cpk::MgVec<T>& v,
                                  const cpk::SegGrid& g, int deg)
{
    const auto k = [&](int s) { return step_key(SECT_DOWN, l, s); };
    if (g.rres) smooth_seg<W, cpk::SEG_RX0, T>(e, k(0), a, v, nullptr, nullptr, v.xa, nullptr, 0, 0, g);
    else smooth_seg<W, cpk::SEG_X0, T>(e, k(0), a, v, nullptr, nullptr, v.xa, nullptr, 0, 0, g);
    // d of the last pre-smoothing step is not stored: the residual step reads
    // only the combined iterate
    smooth_seg<W, cpk::SEG_FIRST_X, T>(e, k(1), a, v, v.xa, nullptr, v.xb, deg == 2 ? nullptr : v.da, 0, 0, g);
    if (deg == 2) {
        smooth<W, cpk::MG_RESID_X, T>(e, k(2), a, v, v.xb, nullptr, nullptr, nullptr, 0, 0);
        return;
    }
    smooth_seg<W, cpk::MG_CHEB_X, T>(e, k(2), a, v, v.xb, v.da, v.xa, nullptr, 1, 0, g);
    smooth<W, cpk::MG_RESID_X, T>(e, k(3), a, v, v.xa, nullptr, nullptr, nullptr, 0, 0);
}

template <int W, typename T, typename E>
static void smooth_level_up_seg(E& e, int l, const cpk::CgArgs<T>& a, const cpk::MgVec<T>& v, const cpk::SegGrid& g,
                                int deg)
{
    const auto k = [&](int s) { return step_key(SECT_UP, l, s); };
    if (deg == 2) {
        smooth_seg<W, cpk::MG_POST1_X, T>(e, k(1), a, v, v.xa, nullptr, v.db, v.da, 0, 0, g);
        smooth_seg<W, cpk::MG_CHEB_X, T>(e, k(2), a, v, v.db, v.da, v.xb, nullptr, 0, 1, g);
        return;
    }
    smooth_seg<W, cpk::MG_POST1_X, T>(e, k(1), a, v, v.xb, nullptr, v.xa, v.da, 0, 0, g);
    smooth_seg<W, cpk::MG_CHEB_X, T>(e, k(2), a, v, v.xa, v.da, v.res, v.db, 0, 0, g);
    smooth_seg<W, cpk::MG_CHEB_X, T>(e, k(3), a, v, v.res, v.db, v.xb, nullptr, 1, 1, g);
}

template <typename T, typename F>
static void by_width(int w, const F& f)
{
    if (w <= 9) f(std::integral_constant<int, 9>());
    else if (w <= 12) f(std::integral_constant<int, 12>());
    else if (w <= 16) f(std::integral_constant<int, 16>());
    else f(std::integral_constant<int, 0>());
}

// Row instantiation of level 0: the symmetric lattice layout where it is
// built (FP32), else the problem's row width.
template <typename T, typename F>
static void by_level0(const cp_problem* P, const F& f)
{
    if (std::is_same<T, float>::value && P->mg->vsym32) f(std::integral_constant<int, cpk::kSymW>());
    else by_width<T>(P->max_w, f);
}

// Row instantiation of level l: level 0's; on the coarse levels the
// symmetric lattice layout where it is built (FP32, open lattice), else the
// lattice rows.
template <typename T, typename F>
static void by_level(const cp_problem* P, int l, const F& f)
{
    if (l == 0) by_level0<T>(P, f);
    else if (wide_tiles<T>(*P->mg, l)) f(std::integral_constant<int, cpk::kSymW>());
    else if (std::is_same<T, float>::value && P->mg->L[l].vsym32) f(std::integral_constant<int, cpk::kSymW16>());
    else f(std::integral_constant<int, cpk::kLatW>());
}

// u = M r: one V-cycle; the result lands in level 0's xb.
template <typename T, typename E>
static void mg_vcycle(E& e, cp_problem* P, const typename cpk::V2<T>::type* rhs, const cpk::CgScal* sc)
{
    using V = typename cpk::V2<T>::type;
    MgState& G = *P->mg;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    auto vec = [&](int l) {
        MgLevel& L = G.L[l];
        cpk::MgVec<T> v{};
        v.b = l == 0 ? rhs : (const V*)L.vec[pi][0];
        v.xa = (V*)L.vec[pi][1];
        v.xb = (V*)L.vec[pi][2];
        v.da = (V*)L.vec[pi][3];
        v.db = (V*)L.vec[pi][4];
        v.res = (V*)L.vec[pi][5];
        v.lam = L.lam + pi * cpk::kLamSlots;
        v.sc = sc;
        return v;
    };
    auto grid = [&](int l) {
        cpk::MgGrid g{};
        g.lg = l == 0 ? G.lg0 : nullptr;
        g.node_of = l == 0 ? G.node0 : nullptr;
        g.n1f = G.L[l].n1;
        g.n2f = G.L[l].n2;
        g.n1c = G.L[l + 1].n1;
        g.n2c = G.L[l + 1].n2;
        g.co1 = G.L[l].co1;
        g.co2 = G.L[l].co2;
        g.per1 = G.per1;
        if (l == 0 && G.q8) {
            g.q8_par = G.q8_par;
            g.q8_ch = G.q8_ch;
        }
        return g;
    };
    const int last = G.lt;   // levels lt.. run in the one-workgroup tail
    // Chebyshev degree per launched segment level: 2 on the coarse levels
    // in FP32 (in FP64 from level 2, on levels of at most kDeg2Nodes nodes)
    // and on a finest level of at least kDeg2Finest nodes, 3 elsewhere.
    // Degree 2 saves two smoothing launches per level and V-cycle for
    // about a quarter more CG iterations: a gain where the launches are
    // latency-bound (small levels) or stream HBM (the 2M-node plate: 4 %
    // faster ramp), a loss on a cache-resident finest level (the 632x316
    // plate: 7 % slower ramp)
    auto degree = [&](int l) {
        if (!G.L[l].seg) return 3;
        if (l == 0) return G.L[0].nn >= kDeg2Finest ? 2 : 3;
        if (std::is_same<T, float>::value) return 2;
        return l >= 2 && G.L[l].nn <= kDeg2Nodes ? 2 : 3;
    };
    // a launched segment level below level 0 restricts its right-hand side
    // in its first smoothing launch
    auto fused_restrict = [&](int l) { return l > 0 && l < last && G.L[l].seg; };
    for (int l = 0; l < last; ++l) {
        const cpk::CgArgs<T> a = level_args<T>(P, l);
        const cpk::MgVec<T> v = vec(l);
        if (G.L[l].seg) {
            cpk::SegGrid sg = seg_grid<T>(G, l);
            if (fused_restrict(l)) {
                sg.rg = grid(l - 1);
                sg.rres = G.L[l - 1].vec[pi][5];
            }
            by_level<T>(P, l, [&](auto w) { smooth_level_down_seg<decltype(w)::value, T>(e, l, a, v, sg, degree(l)); });
        } else {
            by_level<T>(P, l, [&](auto w) { smooth_level_down<decltype(w)::value, T>(e, l, a, v); });
        }
        // (the level above the tail is restricted by the tail itself)
        if (fused_restrict(l + 1) || l + 1 == last) continue;
        const MgLevel& C = G.L[l + 1];
        cpk::GridOp<T> o{};
        o.gr = grid(l);
        o.src = (const V*)v.res;
        o.dst = (V*)C.vec[pi][0];
        o.mask = (const uchar2*)C.mask;
        o.nn = C.nn;
        o.sc = sc;
        e.template restrict_<T>(step_key(SECT_DOWN, l, kStepRestrict), o, C.nb);
    }
    {
        cpk::TailOp<T> o{};
        o.lv = (const cpk::TailLevel<T>*)G.tail[pi];
        o.nl = G.nlev - G.lt;
        o.Dinv = (const double*)G.dense;
        o.nd = G.nd;
        o.sc = sc;
        o.im = G.tail_img[pi];
        o.rg = grid(last - 1);
        o.rres = (const V*)G.L[last - 1].vec[pi][5];
        e.template tail<T>(step_key(SECT_TAIL, 0, 0), o);
    }
    for (int l = last - 1; l >= 0; --l) {
        const cpk::CgArgs<T> a = level_args<T>(P, l);
        const cpk::MgVec<T> v = vec(l);
        const MgLevel& L = G.L[l];
        const MgLevel& C = G.L[l + 1];
        // the pre-smoothed x is in xa (degree 3) or xb (degree 2); the
        // prolongation writes the other buffer
        const bool d2 = degree(l) == 2;
        cpk::GridOp<T> o{};
        o.gr = grid(l);
        o.src = (const V*)C.vec[pi][2];
        o.xin = (const V*)(d2 ? v.xb : v.xa);
        o.dst = d2 ? v.xa : v.xb;
        o.mask = (const uchar2*)a.mask;
        o.nn = L.nn;
        o.sc = sc;
        e.template prolong<T>(step_key(SECT_UP, l, 0), o, L.nb);
        if (L.seg) {
            const cpk::SegGrid sg = seg_grid<T>(G, l);
            by_level<T>(P, l, [&](auto w) { smooth_level_up_seg<decltype(w)::value, T>(e, l, a, v, sg, degree(l)); });
        } else {
            by_level<T>(P, l, [&](auto w) { smooth_level_up<decltype(w)::value, T>(e, l, a, v); });
        }
    }
}

// One iteration of CG preconditioned by the V-cycle: p, s, x, r updates,
// the stop test on the new r, u = M r, w = A u with the three dot products,
// scalars of the next step.
// (`a` holds the solve's vectors; level 0's row layouts come with it from
// level_args.)
template <typename T, typename E>
static void mg_cg_iteration(E& e, cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc,
                            const typename cpk::V2<T>::type* u)
{
    cpk::CgOp<T> o{};
    o.a = level_args<T>(P, 0);
    o.a.sc = a.sc;
    o.u = u;
    e.template update<T>(step_key(SECT_UPDATE, 0, 0), o, P->nbn);
    e.template check<T>(step_key(SECT_CHECK, 0, 0), o);
    mg_vcycle<T>(e, P, a.r[0], dsc);
    by_level0<T>(P, [&](auto w) { e.template dots<decltype(w)::value, T>(step_key(SECT_DOTS, 0, 0), o, P->nbn); });
    e.template step<T>(step_key(SECT_STEP, 0, 0), o);
}

// The first V-cycle of a solve: u_0 = M r_0, w_0 = A u_0 with the dots,
// and the scalars of the first step.
template <typename T>
static void mg_cg_start(cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc, const typename cpk::V2<T>::type* u)
{
    hipStream_t st = P->stream;
    const dim3 g(P->nbn), b(cpk::kBlock);
    hipLaunchKernelGGL(cpk::k_mg_cg_clear<T>, g, b, 0, st, a);
    Direct e{st};
    mg_vcycle<T>(e, P, a.r[0], dsc);
    cpk::CgArgs<T> a0 = level_args<T>(P, 0);
    a0.sc = a.sc;
    by_level0<T>(P, [&](auto w) {
        hipLaunchKernelGGL((cpk::k_mg_cg_dots<decltype(w)::value, T>), g, b, 0, st, a0, u);
    });
    hipLaunchKernelGGL(cpk::k_cg_start<T>, dim3(1), b, 0, st, a);
}

// ------------------------------------------------------------ lockstep batches
//
// BatchRec collects the steps of one CG iteration of every problem of a
// batch, grouped by (key, kernel): each group becomes one launch of the
// batched kernel (cp_mg.h) over the tiles {op, blk} of all its problems,
// in key order, so every problem's steps keep their order and each launch
// advances every problem that has that step.

// A batched kernel of cp_mg.h: its address, its workgroup size, and whether
// it takes tiles (ops, tiles, ntiles) or runs one workgroup per op (ops).
struct BatchKernel {
    const void* func;
    int threads;
    bool tiled;
};

template <int W, int MODE, typename T>
static BatchKernel bk_smooth() { return {reinterpret_cast<const void*>(&cpk::k_b_smooth<W, MODE, T>), cpk::kBlock, true}; }
template <int W, int MODE, typename T>
static BatchKernel bk_smooth_seg()
{
    return {reinterpret_cast<const void*>(&cpk::k_b_smooth_seg<W, MODE, T>), cpk::kSeg * cpk::seg_cols<W>(), true};
}
template <typename T>
static BatchKernel bk_restrict() { return {reinterpret_cast<const void*>(&cpk::k_b_restrict<T>), cpk::kBlock, true}; }
template <typename T>
static BatchKernel bk_prolong() { return {reinterpret_cast<const void*>(&cpk::k_b_prolong<T>), cpk::kBlock, true}; }
template <typename T>
static BatchKernel bk_tail() { return {reinterpret_cast<const void*>(&cpk::k_b_tail<T>), cpk::kTailThreads, false}; }
template <typename T>
static BatchKernel bk_update() { return {reinterpret_cast<const void*>(&cpk::k_b_cg_update<T>), cpk::kBlock, true}; }
template <typename T>
static BatchKernel bk_check() { return {reinterpret_cast<const void*>(&cpk::k_b_cg_check<T>), cpk::kBlock, false}; }
template <int W, typename T>
static BatchKernel bk_dots() { return {reinterpret_cast<const void*>(&cpk::k_b_cg_dots<W, T>), cpk::kBlock, true}; }
template <typename T>
static BatchKernel bk_step() { return {reinterpret_cast<const void*>(&cpk::k_b_cg_step<T>), cpk::kBlock, false}; }

struct BatchRec {
    struct Group {
        BatchKernel k{};
        size_t op_size = 0;
        std::vector<char> ops;
        std::vector<int> nblk;
    };
    // ordered by step key; within a key, the groups of different kernels
    // hold different problems and may run in any order
    std::map<std::pair<unsigned long long, const void*>, Group> groups;

    template <typename O>
    void add(unsigned long long key, const BatchKernel& k, const O& o, int nblk)
    {
        Group& g = groups[{key, k.func}];
        g.k = k;
        g.op_size = sizeof(O);
        const char* p = reinterpret_cast<const char*>(&o);
        g.ops.insert(g.ops.end(), p, p + sizeof(O));
        g.nblk.push_back(k.tiled ? nblk : 1);
    }
    template <int W, int MODE, typename T>
    void smooth(unsigned long long k, const cpk::StepOp<T>& o, int nb) { add(k, bk_smooth<W, MODE, T>(), o, nb); }
    template <int W, int MODE, typename T>
    void smooth_seg(unsigned long long k, const cpk::StepOp<T>& o, int nt) { add(k, bk_smooth_seg<W, MODE, T>(), o, nt); }
    template <typename T>
    void restrict_(unsigned long long k, const cpk::GridOp<T>& o, int nb) { add(k, bk_restrict<T>(), o, nb); }
    template <typename T>
    void prolong(unsigned long long k, const cpk::GridOp<T>& o, int nb) { add(k, bk_prolong<T>(), o, nb); }
    template <typename T>
    void tail(unsigned long long k, const cpk::TailOp<T>& o) { add(k, bk_tail<T>(), o, 1); }
    template <typename T>
    void update(unsigned long long k, const cpk::CgOp<T>& o, int nb) { add(k, bk_update<T>(), o, nb); }
    template <typename T>
    void check(unsigned long long k, const cpk::CgOp<T>& o) { add(k, bk_check<T>(), o, 1); }
    template <int W, typename T>
    void dots(unsigned long long k, const cpk::CgOp<T>& o, int nb) { add(k, bk_dots<W, T>(), o, nb); }
    template <typename T>
    void step(unsigned long long k, const cpk::CgOp<T>& o) { add(k, bk_step<T>(), o, 1); }
};

// A batch's launches (their tables in one device buffer) as a graph of one
// iteration, and the problems they were recorded for (rebuilt when that
// set changes).  The graph is built node by node, not captured: the serving
// loop's preparation thread makes device calls meanwhile, which would
// invalidate a capture.
struct cp_mg_batch {
    std::vector<long long> ids;   // the callers' names of the problems recorded
    int precision = -1;
    int nlaunch = 0;
    char* dev = nullptr;
    size_t dev_size = 0;
    hipEvent_t done = nullptr;    // the last run's launches have finished
    bool ran = false;
    hipGraphExec_t exec = nullptr;
    ~cp_mg_batch()
    {
        if (ran) (void)hipEventSynchronize(done);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (done) (void)hipEventDestroy(done);
        if (dev) (void)hipFree(dev);
    }
};

void mg_batch_destroy(cp_mg_batch* b) { delete b; }

template <typename T>
static int mg_batch_build(cp_mg_batch* B, cp_problem* const* probs, const long long* ids, int n)
{
    BatchRec rec;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    for (int i = 0; i < n; ++i) {
        cp_problem* P = probs[i];
        if (!P->mg || !P->mg->alloc[pi] || !P->mg->tail[pi])
            return fail(CP_ERR_STATE, "multigrid batch: a problem without its prepared hierarchy");
        cpk::CgArgs<T> a;
        if constexpr (std::is_same<T, double>::value) a = P->cg64_args();
        else a = P->cg32_args();
        if (a.nblocks != P->nbn) return fail(CP_ERR_STATE, "multigrid batch: inconsistent block count");
        using V = typename cpk::V2<T>::type;
        mg_cg_iteration<T>(rec, P, a, a.sc, (const V*)P->mg->L[0].vec[pi][2]);
    }
    // layout: per group its ops, then its tiles (tiled groups), 256-byte aligned
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t need = 0;
    for (const auto& kv : rec.groups) {
        const BatchRec::Group& g = kv.second;
        need += align(g.ops.size());
        if (g.k.tiled) {
            long long nt = 0;
            for (int b : g.nblk) nt += b;
            if (nt <= 0 || nt > (1LL << 30)) return fail(CP_ERR_INVALID, "multigrid batch: bad tile count");
            need += align((size_t)nt * sizeof(int2));
        }
    }
    std::vector<char> host(need, 0);
    struct Launch {
        BatchKernel k;
        size_t ops_off, tiles_off;
        int n;
    };
    std::vector<Launch> launches;
    size_t off = 0;
    for (const auto& kv : rec.groups) {
        const BatchRec::Group& g = kv.second;
        const int nops = (int)g.nblk.size();
        Launch L{g.k, off, 0, 0};
        std::memcpy(host.data() + off, g.ops.data(), g.ops.size());
        off += align(g.ops.size());
        if (g.k.tiled) {
            L.tiles_off = off;
            int2* t = reinterpret_cast<int2*>(host.data() + off);
            int nt = 0;
            for (int i = 0; i < nops; ++i)
                for (int b = 0; b < g.nblk[i]; ++b) t[nt++] = make_int2(i, b);
            L.n = nt;
            off += align((size_t)nt * sizeof(int2));
        } else {
            L.n = nops;
        }
        launches.push_back(L);
    }
    // the previous run may still read the old tables
    if (B->ran) CP_HIP(hipEventSynchronize(B->done));
    if (B->exec) {
        CP_HIP(hipGraphExecDestroy(B->exec));
        B->exec = nullptr;
    }
    if (need > B->dev_size) {
        if (B->dev) CP_HIP(hipFree(B->dev));
        B->dev = nullptr;
        B->dev_size = 0;
        if (hipMalloc(&B->dev, need) != hipSuccess) {
            (void)hipGetLastError();
            return fail(CP_ERR_NOMEM, "multigrid batch: device allocation failed");
        }
        B->dev_size = need;
    }
    CP_HIP(hipMemcpy(B->dev, host.data(), need, hipMemcpyHostToDevice));
    // one iteration as a chain of kernel nodes
    hipGraph_t graph = nullptr;
    CP_HIP(hipGraphCreate(&graph, 0));
    hipGraphNode_t prev = nullptr;
    int s = CP_OK;
    for (const Launch& L : launches) {
        const void* ops = B->dev + L.ops_off;
        const void* tiles = B->dev + L.tiles_off;
        int n = L.n;
        void* targs[3] = {(void*)&ops, (void*)&tiles, (void*)&n};
        hipKernelNodeParams kp{};
        kp.func = const_cast<void*>(L.k.func);
        kp.gridDim = dim3(L.n);
        kp.blockDim = dim3(L.k.threads);
        kp.sharedMemBytes = 0;
        kp.kernelParams = targs;
        kp.extra = nullptr;
        hipGraphNode_t node = nullptr;
        if (hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp) != hipSuccess) {
            s = fail(CP_ERR_DEVICE, "multigrid batch: graph node creation failed");
            break;
        }
        prev = node;
    }
    if (s == CP_OK && hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        B->exec = nullptr;
        s = fail(CP_ERR_DEVICE, "multigrid batch: graph instantiation failed");
    }
    (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (s != CP_OK) return s;
    B->nlaunch = (int)launches.size();
    B->ids.assign(ids, ids + n);
    B->precision = pi == 0 ? CP_FP64 : CP_MIXED;
    return CP_OK;
}

int mg_batch_run(cp_mg_batch** pb, cp_problem* const* probs, const long long* ids, int n, int precision, int iters,
                 hipStream_t st)
{
    if (n < 1 || iters < 1) return fail(CP_ERR_INVALID, "multigrid batch: empty");
    if (!*pb) {
        auto* b = new cp_mg_batch();
        if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
            delete b;
            return fail(CP_ERR_DEVICE, "multigrid batch: event creation failed");
        }
        *pb = b;
    }
    cp_mg_batch* B = *pb;
    const bool same = B->precision == precision && (int)B->ids.size() == n && std::equal(B->ids.begin(), B->ids.end(), ids);
    if (!same) {
        B->ids.clear();
        if (int s = precision == CP_FP64 ? mg_batch_build<double>(B, probs, ids, n)
                                         : mg_batch_build<float>(B, probs, ids, n))
            return s;
    }
    for (int j = 0; j < iters; ++j) CP_HIP(hipGraphLaunch(B->exec, st));
    CP_HIP(hipGetLastError());
    CP_HIP(hipEventRecord(B->done, st));
    B->ran = true;
    return CP_OK;
}

int mg_batch_launches(const cp_mg_batch* b) { return b ? b->nlaunch : 0; }

template <typename F>
static int capture(cp_problem* P, hipGraphExec_t* exec, const F& record)
{
    hipGraph_t graph = nullptr;
    CP_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
    record();
    const hipError_t e = hipStreamEndCapture(P->stream, &graph);
    if (e != hipSuccess) return failf(CP_ERR_DEVICE, "multigrid: graph capture failed: %s", hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
        *exec = nullptr;
        return failf(CP_ERR_DEVICE, "multigrid: graph instantiation failed: %s", hipGetErrorString(e2));
    }
    return CP_OK;
}

// The iteration and the start of a solve recorded once each as a HIP graph
// (all their arguments are fixed device addresses, the scalars live on the
// device) and replayed: about 50 small launches per V-cycle cost more on
// the host than on the device otherwise.
template <typename T>
static int mg_graph(cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc, const typename cpk::V2<T>::type* u,
                    hipGraphExec_t* start, hipGraphExec_t* iter)
{
    MgState& G = *P->mg;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    if (!G.iter[pi])
        if (int s = capture(P, &G.iter[pi], [&] {
                Direct e{P->stream};
                mg_cg_iteration<T>(e, P, a, dsc, u);
            }))
            return s;
    if (!G.start[pi])
        if (int s = capture(P, &G.start[pi], [&] { mg_cg_start<T>(P, a, dsc, u); })) return s;
    *iter = G.iter[pi];
    *start = G.start[pi];
    return CP_OK;
}

int mg_prepare(cp_problem* P, int precision)
{
    if (!P->mg) return fail(CP_ERR_STATE, "multigrid needs the mesh lattice: call cp_problem_set_grid first");
    if (int s = mg_alloc(P, precision)) return s;
    if (int s = precision == CP_FP64 ? mg_tail_table<double>(P) : mg_tail_table<float>(P)) return s;
    if (!P->mg->use_graphs) return CP_OK;
    hipGraphExec_t start = nullptr, iter = nullptr;
    if (precision == CP_FP64)
        return mg_graph<double>(P, P->cg64_args(), P->sc64, (const double2*)P->mg->L[0].vec[0][2], &start, &iter);
    return mg_graph<float>(P, P->cg32_args(), P->sc32, (const float2*)P->mg->L[0].vec[1][2], &start, &iter);
}

template <typename T>
static int mg_start_t(cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc, hipGraphExec_t* exec)
{
    using V = typename cpk::V2<T>::type;
    MgState& G = *P->mg;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    const V* u = (const V*)G.L[0].vec[pi][2];
    hipGraphExec_t start = nullptr;
    *exec = nullptr;
    if (G.use_graphs)
        if (int s = mg_graph<T>(P, a, dsc, u, &start, exec)) return s;
    if (start) CP_HIP(hipGraphLaunch(start, P->stream));
    else mg_cg_start<T>(P, a, dsc, u);
    CP_HIP(hipGetLastError());
    return CP_OK;
}

template <typename T>
static int mg_iterations_t(cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc, hipGraphExec_t exec, int n)
{
    using V = typename cpk::V2<T>::type;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    const V* u = (const V*)P->mg->L[0].vec[pi][2];
    Direct e{P->stream};
    for (int j = 0; j < n; ++j) {
        if (exec) CP_HIP(hipGraphLaunch(exec, P->stream));
        else mg_cg_iteration<T>(e, P, a, dsc, u);
    }
    CP_HIP(hipGetLastError());
    return CP_OK;
}

// Iterations after convergence still cost their (empty) launches, so the
// first chunk is the predicted count: from the reduction still needed when
// the solve starts from zero (the refinement passes, where (r_0, r_0) is
// bnorm2) and the rate of the previous solves, else the previous solve's
// count.
int mg_cg_first_chunk(const cp_problem* P, double tol2, bool zero_start)
{
    int chunk;
    if (zero_start && P->mg_lrate < 0. && tol2 > 0. && tol2 < 1.)
        chunk = (int)std::lround(std::log(tol2) / P->mg_lrate);
    else
        chunk = (int)P->last_cg_iters;
    return std::min(64, std::max(1, chunk));
}

void mg_cg_observe(cp_problem* P, const cpk::CgScal& h)
{
    if (h.iters > 0 && h.rho > 0. && h.rho0 > h.rho) {
        const double lr = std::log(h.rho / h.rho0) / double(h.iters);
        P->mg_lrate = P->mg_lrate < 0. ? 0.7 * P->mg_lrate + 0.3 * lr : lr;
    }
}

size_t mg_bytes_estimate(int nn, int precision)
{
    // per coarse node: 9 columns, the stencil values (FP64: two double2
    // planes; FP32: float4), pivot inverse, mask, six vectors; per finest
    // node: five vectors, the slot table and the lattice maps; the
    // coarsest inverse
    const bool f64 = precision == CP_FP64;
    const double coarse = 36. + (f64 ? 288. : 144.) + (f64 ? 32. : 16.) + 2. + 6. * (f64 ? 16. : 8.);
    const double fine = 5. * (f64 ? 16. : 8.) + 9. + 8.;
    return (size_t)(nn * (coarse + fine)) + (size_t)cpk::kMgDenseMax * cpk::kMgDenseMax * 8 + (1 << 20);
}

int mg_cg_enqueue_start(cp_problem* P, int precision)
{
    hipGraphExec_t exec = nullptr;
    return precision == CP_FP64 ? mg_start_t<double>(P, P->cg64_args(), P->sc64, &exec)
                                : mg_start_t<float>(P, P->cg32_args(), P->sc32, &exec);
}

int mg_cg_enqueue_iterations(cp_problem* P, int precision, int n)
{
    MgState& G = *P->mg;
    const int pi = precision == CP_FP64 ? 0 : 1;
    hipGraphExec_t exec = G.use_graphs ? G.iter[pi] : nullptr;
    return precision == CP_FP64 ? mg_iterations_t<double>(P, P->cg64_args(), P->sc64, exec, n)
                                : mg_iterations_t<float>(P, P->cg32_args(), P->sc32, exec, n);
}

// CG preconditioned by the V-cycle (Chronopoulos-Gear form, one reduction
// per iteration): the caller has launched the scalar set-up and r_0.
template <typename T>
int run_mg_cg(cp_problem* P, const cpk::CgArgs<T>& a, cpk::CgScal* dsc, double tol2, double bnorm2,
              long long* iters, int* status, double* ms, bool zero_start)
{
    hipStream_t st = P->stream;
    CP_HIP(hipEventRecord(P->ev[6], st));
    hipGraphExec_t exec = nullptr;
    if (int s = mg_start_t<T>(P, a, dsc, &exec)) return s;
    long long launched = 0;
    int chunk = mg_cg_first_chunk(P, tol2, zero_start && bnorm2 > 0.);
    double g_prev = -1.;
    long long it_prev = 0;
    for (;;) {
        if (int s = mg_iterations_t<T>(P, a, dsc, exec, chunk)) return s;
        launched += chunk;
        CP_HIP(hipMemcpyAsync(P->h_cg, dsc, sizeof(cpk::CgScal), hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
        const cpk::CgScal h = *P->h_cg;
        if (h.done) break;
        const double target = tol2 * (bnorm2 > 0. ? bnorm2 : h.bnorm2);
        int next = std::min(64, chunk * 2);
        const double r0 = g_prev > 0. ? g_prev : h.rho0;
        const long long i0 = g_prev > 0. ? it_prev : 0;
        if (r0 > 0. && h.rho > 0. && h.rho < r0 && target > 0. && h.iters > i0) {
            const double rate = std::log(r0 / h.rho) / double(h.iters - i0);
            const double rem = std::log(h.rho / target) / rate;
            if (rem > 0 && rem < 1e7) next = (int)std::min(64., std::max(1., std::ceil(rem)));
        }
        g_prev = h.rho;
        it_prev = h.iters;
        chunk = next;
        if (launched > (long long)h.maxit + 256) break;
    }
    CP_HIP(hipEventRecord(P->ev[7], st));
    CP_HIP(hipEventSynchronize(P->ev[7]));
    float t = 0.f;
    CP_HIP(hipEventElapsedTime(&t, P->ev[6], P->ev[7]));
    *ms = t;
    *iters = P->h_cg->iters;
    *status = P->h_cg->status;
    mg_cg_observe(P, *P->h_cg);
    return CP_OK;
}
template int run_mg_cg<double>(cp_problem*, const cpk::CgArgs<double>&, cpk::CgScal*, double, double, long long*,
                               int*, double*, bool);
template int run_mg_cg<float>(cp_problem*, const cpk::CgArgs<float>&, cpk::CgScal*, double, double, long long*,
                              int*, double*, bool);

template <typename T>
static int mg_bench_t(cp_problem* P, int reps, double* ms_vcycle, double* ms_smooth, double* bytes_smooth)
{
    using V = typename cpk::V2<T>::type;
    MgState& G = *P->mg;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    hipStream_t st = P->stream;
    const cpk::CgArgs<T> a = level_args<T>(P, 0);
    const dim3 g(P->nbn), b(cpk::kBlock);
    // right-hand side: the residual of the current state (the FP32 solve
    // works on the FP64 residual, as in the refinement loop)
    if (pi == 1)
        if (int s = enqueue_residual(P)) return s;
    hipLaunchKernelGGL(cpk::k_cg_init1<T>, g, b, 0, st, a, pi);
    MgLevel& L = G.L[0];
    cpk::MgVec<T> v{};
    v.b = a.r[0];
    v.xa = (V*)L.vec[pi][1];
    v.xb = (V*)L.vec[pi][2];
    v.da = (V*)L.vec[pi][3];
    v.db = (V*)L.vec[pi][4];
    v.res = (V*)L.vec[pi][5];
    v.lam = L.lam + pi * cpk::kLamSlots;
    const bool seg = L.seg != 0;
    const cpk::SegGrid sg = seg_grid<T>(G, 0);
    Direct e{st};
    const unsigned long long k0 = 0;
    by_level0<T>(P, [&](auto w) {
        if (seg) {
            smooth_seg<decltype(w)::value, cpk::SEG_X0, T>(e, k0, a, v, nullptr, nullptr, v.xa, nullptr, 0, 0, sg);
            smooth_seg<decltype(w)::value, cpk::SEG_FIRST, T>(e, k0, a, v, v.xa, nullptr, nullptr, v.da, 0, 0, sg);
        } else {
            smooth<decltype(w)::value, cpk::MG_PRE1, T>(e, k0, a, v, nullptr, nullptr, v.xa, v.da, 0, 0);
        }
    });
    CP_HIP(hipEventRecord(P->ev[4], st));
    // the Chebyshev step the V-cycle runs: on a segment level the one that
    // reads the combined iterate (x + d, as its previous step writes it)
    for (int i = 0; i < reps; ++i)
        by_level0<T>(P, [&](auto w) {
            if (seg) smooth_seg<decltype(w)::value, cpk::MG_CHEB_X, T>(e, k0, a, v, v.xa, v.da, v.xb, v.db, 1, 0, sg);
            else smooth<decltype(w)::value, cpk::MG_CHEB, T>(e, k0, a, v, v.xa, v.da, v.xb, v.db, 1, 0);
        });
    CP_HIP(hipEventRecord(P->ev[5], st));
    for (int i = 0; i < reps; ++i) mg_vcycle<T>(e, P, a.r[0], nullptr);
    CP_HIP(hipEventRecord(P->ev[6], st));
    CP_HIP(hipGetLastError());
    CP_HIP(hipEventSynchronize(P->ev[6]));
    float t1 = 0.f, t2 = 0.f;
    CP_HIP(hipEventElapsedTime(&t1, P->ev[4], P->ev[5]));
    CP_HIP(hipEventElapsedTime(&t2, P->ev[5], P->ev[6]));
    if (ms_smooth) *ms_smooth = t1 / reps;
    if (ms_vcycle) *ms_vcycle = t2 / reps;
    const double s = std::is_same<T, double>::value ? 8. : 4.;
    // matrix + columns, slices (the symmetric lattice layout: its five
    // blocks per node); x, d, b read, x, d written (2 values per node
    // each), the block-diagonal inverse (4 values per node), the mask
    const double mat = G.vsym32 && pi == 1 ? (double)G.L[0].nslices * cpk::kWave * cpk::kSymSlots * 4. * s
                                           : P->nblk_stored * (4. * s + 4.) + 8. * P->nslices;
    if (bytes_smooth) *bytes_smooth = mat + (5. * 2. * s + 4. * s) * P->nn + 2. * P->nn;
    return CP_OK;
}

// The multigrid CG's operator product alone: `reps` launches of its
// w = mask (A u) kernel with the three dot products (level 0's layout: the
// symmetric lattice layout in FP32 where it is built), u a V-cycle result.
template <typename T>
static int mg_bench_spmv_t(cp_problem* P, int reps, double* ms, double* bytes)
{
    using V = typename cpk::V2<T>::type;
    MgState& G = *P->mg;
    const int pi = std::is_same<T, double>::value ? 0 : 1;
    hipStream_t st = P->stream;
    cpk::CgArgs<T> a0;
    if constexpr (std::is_same<T, double>::value) a0 = P->cg64_args();
    else a0 = P->cg32_args();
    hipLaunchKernelGGL(cpk::k_cg_setup, dim3(1), dim3(1), 0, st, a0.sc, 1.0, 0.0, 1 << 30);
    cpk::CgOp<T> o{};
    o.a = level_args<T>(P, 0);
    o.a.sc = a0.sc;
    o.u = (const V*)G.L[0].vec[pi][2];
    Direct e{st};
    CP_HIP(hipEventRecord(P->ev[4], st));
    for (int i = 0; i < reps; ++i)
        by_level0<T>(P, [&](auto w) { e.template dots<decltype(w)::value, T>(step_key(SECT_DOTS, 0, 0), o, P->nbn); });
    CP_HIP(hipEventRecord(P->ev[5], st));
    CP_HIP(hipGetLastError());
    CP_HIP(hipEventSynchronize(P->ev[5]));
    float t = 0.f;
    CP_HIP(hipEventElapsedTime(&t, P->ev[4], P->ev[5]));
    if (ms) *ms = t / reps;
    const double s = std::is_same<T, double>::value ? 8. : 4.;
    const double mat = G.vsym32 && pi == 1 ? (double)G.L[0].nslices * cpk::kWave * cpk::kSymSlots * 4. * s
                                           : P->nblk_stored * (4. * s + 4.) + 8. * P->nslices;
    // u and r read, w written (2 values per node each), the mask, the partials
    if (bytes) *bytes = mat + 3. * 2. * s * P->nn + 2. * P->nn + 24. * P->nbn;
    return CP_OK;
}

}  // namespace cp

extern "C" {

int cp_bench_mg(cp_problem* P, int precision, int reps, double* ms_vcycle, double* ms_smooth, double* bytes_smooth)
{
    if (!P || reps < 1 || (precision != CP_FP64 && precision != CP_MIXED)) return cp::fail(CP_ERR_INVALID, "bad arguments");
    if (!P->mg) return cp::fail(CP_ERR_STATE, "cp_bench_mg: no multigrid lattice set");
    P->precond = 2;
    if (int s = cp::launch_assemble(P, precision)) return s;
    return precision == CP_FP64 ? cp::mg_bench_t<double>(P, reps, ms_vcycle, ms_smooth, bytes_smooth)
                                : cp::mg_bench_t<float>(P, reps, ms_vcycle, ms_smooth, bytes_smooth);
}

int cp_bench_mg_spmv(cp_problem* P, int precision, int reps, double* ms, double* bytes)
{
    if (!P || reps < 1 || (precision != CP_FP64 && precision != CP_MIXED)) return cp::fail(CP_ERR_INVALID, "bad arguments");
    if (!P->mg) return cp::fail(CP_ERR_STATE, "cp_bench_mg_spmv: no multigrid lattice set");
    P->precond = 2;
    if (int s = cp::launch_assemble(P, precision)) return s;
    return precision == CP_FP64 ? cp::mg_bench_spmv_t<double>(P, reps, ms, bytes)
                                : cp::mg_bench_spmv_t<float>(P, reps, ms, bytes);
}

namespace cp {

// The lattice of a structured mesh and the level hierarchy on it; per1: the
// lattice is periodic in direction 1 (the rays of a ring or of a plate with
// a hole: ray n1 - 1 neighbours ray 0).
// An 8-node mesh over its corner lattice (lg: node -> lattice point, -1 for
// the mid-side nodes): every element's corners are lattice neighbours, each
// mid-side node lies on the edge of two of them; the transfer tables of
// MgGrid (parents per fine node, children per lattice point).
static int q8_transfer(const cp_problem* P, int n1, int per1, int n2, const std::vector<int>& lg,
                       std::vector<int2>& par, std::vector<int>& ch)
{
    const int nn = P->nn, ne = P->ne, nl = n1 * n2;
    if (P->conn_host.size() != 8 * (size_t)ne) return fail(CP_ERR_STATE, "cp_problem_set_grid: no 8-node connectivity");
    par.assign(nn, make_int2(-1, -1));
    for (int n = 0; n < nn; ++n)
        if (lg[n] >= 0) par[n] = make_int2(lg[n], -1);
    for (int e = 0; e < ne; ++e) {
        const int* c = P->conn_host.data() + 8 * (size_t)e;
        for (int k = 0; k < 4; ++k) {
            const int a = lg[c[k]], b = lg[c[(k + 1) % 4]], m = c[4 + k];
            if (a < 0 || b < 0) return failf(CP_ERR_INVALID, "cp_problem_set_grid: corner of element %d is not on the lattice", e);
            if (lg[m] >= 0) return failf(CP_ERR_INVALID, "cp_problem_set_grid: mid-side node %d is on the corner lattice", m);
            const int o1 = cpk::mg_offset(n1, per1, a / n2, b / n2), o2 = b % n2 - a % n2;
            if (std::abs(o1) + std::abs(o2) != 1)
                return failf(CP_ERR_INVALID, "cp_problem_set_grid: the edge of element %d is not a lattice step", e);
            const int2 pq = make_int2(std::min(a, b), std::max(a, b));
            if (par[m].x < 0) par[m] = pq;
            else if (par[m].x != pq.x || par[m].y != pq.y)
                return failf(CP_ERR_INVALID, "cp_problem_set_grid: mid-side node %d on two edges", m);
        }
    }
    for (int n = 0; n < nn; ++n)
        if (par[n].x < 0) return failf(CP_ERR_INVALID, "cp_problem_set_grid: node %d is in no element", n);
    ch.assign((size_t)nl * cpk::kQ8Children, -1);
    for (int n = 0; n < nn; ++n) {
        if (par[n].y < 0) {
            ch[(size_t)par[n].x * cpk::kQ8Children] = n;
            continue;
        }
        for (int p : {par[n].x, par[n].y}) {
            int j = 1;
            while (j < cpk::kQ8Children && ch[(size_t)p * cpk::kQ8Children + j] >= 0) ++j;
            if (j == cpk::kQ8Children)
                return failf(CP_ERR_INVALID, "cp_problem_set_grid: lattice point %d has more than 4 edges", p);
            ch[(size_t)p * cpk::kQ8Children + j] = n;
        }
    }
    return CP_OK;
}

static int set_grid(cp_problem* P, int n1, int n2, const int* node_of, int per1)
{
    if (!P || n1 < 2 || n2 < 2 || (per1 && n1 < 3)) return fail(CP_ERR_INVALID, "cp_problem_set_grid: bad arguments");
    const bool q8 = P->npe == 8;
    const int nn = P->nn;
    // 4-node elements: the nodes are the lattice; 8-node elements: the
    // corners are, and level 0 is the mesh itself above it
    const long long nl = (long long)n1 * n2;
    if (q8 ? (nl > nn || !node_of) : nl != nn)
        return failf(CP_ERR_INVALID, "cp_problem_set_grid: %d x %d lattice for %d nodes%s", n1, n2, nn,
                     q8 ? " (8-node elements: the corner lattice and its node_of)" : "");
    std::vector<int> lg(nn, -1);
    if (node_of) {
        for (int l = 0; l < nl; ++l) {
            const int n = node_of[l];
            if (n < 0 || n >= nn || lg[n] >= 0) return fail(CP_ERR_INVALID, "cp_problem_set_grid: node_of is not a permutation");
            lg[n] = l;
        }
    } else {
        for (int n = 0; n < nn; ++n) lg[n] = n;
    }
    std::vector<int2> q8par;
    std::vector<int> q8ch;
    if (q8) {
        if (int s = q8_transfer(P, n1, per1, n2, lg, q8par, q8ch)) return s;
    } else {
        // every coupling of the matrix must stay within one lattice step
        for (int n = 0; n < nn; ++n) {
            const int a1 = lg[n] / n2, a2 = lg[n] % n2;
            for (int q = P->adj_ptr[n]; q < P->adj_ptr[n + 1]; ++q) {
                const int m = P->adj_node[q];
                const int b1 = lg[m] / n2, b2 = lg[m] % n2;
                if (std::abs(cpk::mg_offset(n1, per1, a1, b1)) > 1 || std::abs(a2 - b2) > 1)
                    return failf(CP_ERR_INVALID, "cp_problem_set_grid: nodes %d and %d share an element but are not lattice neighbours",
                                 n, m);
            }
        }
    }
    CP_HIP(hipStreamSynchronize(P->stream));
    mg_destroy(P);
    auto* G = new MgState();
    P->mg = G;
    G->per1 = per1;
    G->q8 = q8;
    // mean lattice edge lengths along the two directions: a stretched mesh
    // couples strongly across its short edges, so only that direction is
    // halved until the two are within a factor 2 (semi-coarsening)
    double len1 = 0., len2 = 0.;
    std::vector<double2> pos(nl);   // lattice point coordinates in lattice order
    {
        std::vector<double2> xy(nn);
        CP_HIP(hipMemcpy(xy.data(), P->xy, sizeof(double2) * nn, hipMemcpyDeviceToHost));
        for (int n = 0; n < nn; ++n)
            if (lg[n] >= 0) pos[lg[n]] = xy[n];
        long long c1n = 0, c2n = 0;
        for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n2; ++i2) {
                const double2 p = pos[i1 * n2 + i2];
                if (i1 + 1 < n1 || per1) {
                    const double2 q = pos[((i1 + 1) % n1) * n2 + i2];
                    len1 += std::hypot(q.x - p.x, q.y - p.y);
                    ++c1n;
                }
                if (i2 + 1 < n2) {
                    const double2 q = pos[i1 * n2 + i2 + 1];
                    len2 += std::hypot(q.x - p.x, q.y - p.y);
                    ++c2n;
                }
            }
        len1 /= std::max<long long>(c1n, 1);
        len2 /= std::max<long long>(c2n, 1);
    }
    // an 8-node mesh's own level above its corner lattice (no lattice of
    // its own: 0 x 0, the point smoother on its rows)
    if (q8) {
        MgLevel L;
        L.nn = nn;
        L.nb = blocks_of(nn);
        L.nslices = (nn + cpk::kWave - 1) / cpk::kWave;
        L.co1 = L.co2 = 0;
        G->L.push_back(L);
    }
    const int first = q8 ? 1 : 0;   // the first lattice level
    // levels until the coarsest has at most kMgDenseMax DOFs
    int c1 = n1, c2 = n2;
    for (;;) {
        MgLevel L;
        L.n1 = c1;
        L.n2 = c2;
        L.nn = c1 * c2;
        L.nb = blocks_of(L.nn);
        L.nslices = (L.nn + cpk::kWave - 1) / cpk::kWave;
        const bool done = 2 * c1 * c2 <= cpk::kMgDenseMax;
        // a periodic direction halves only to an even count of at least 3 points
        const bool can1 = per1 ? cpk::mg_per_halves(c1) : c1 > 2;
        L.co1 = !(len1 > 2. * len2) && can1;
        L.co2 = !(len2 > 2. * len1) && c2 > 2;
        if (!L.co1 && !L.co2) L.co1 = can1, L.co2 = c2 > 2;
        G->L.push_back(L);
        if (done) break;
        const int d1 = cpk::mg_coarse_n(c1, L.co1, per1), d2 = cpk::mg_coarse_n(c2, L.co2);
        if (d1 == c1 && d2 == c2) break;
        if (d1 != c1) len1 *= 2.;
        if (d2 != c2) len2 *= 2.;
        c1 = d1;
        c2 = d2;
    }
    G->nlev = (int)G->L.size();
    G->lt = G->nlev - 1;
    // the tail: the coarsest levels of at most kTailThreads nodes whose
    // vectors fit its LDS in FP64
    {
        int lds = tail_lds_bytes(G->L[G->lt].nn, sizeof(double2));
        while (G->lt > 1 && G->L[G->lt - 1].nn <= cpk::kTailThreads &&
               lds + tail_lds_bytes(G->L[G->lt - 1].nn, sizeof(double2)) <= cpk::kTailVecLds) {
            --G->lt;
            lds += tail_lds_bytes(G->L[G->lt].nn, sizeof(double2));
        }
    }
    // smoother of each level: fully coarsened levels (outside the one-
    // workgroup tail) where at least kSegFrac of the cells are stretched
    // along direction 1 (mean direction-1 edge below half the direction-2
    // edge) use the segment smoother; coarse nodes sit on their fine
    // positions, so the cells of every level are measured on the mesh
    for (int l = first; l < G->lt; ++l) {
        MgLevel& L = G->L[l];
        long long cells = 0, stretched = 0;
        for (int i1 = 0; i1 + 1 < L.n1; ++i1)
            for (int i2 = 0; i2 + 1 < L.n2; ++i2) {
                auto at = [&](int a1, int a2) { return pos[a1 * L.n2 + a2]; };
                auto len = [](double2 p, double2 q) { return std::hypot(q.x - p.x, q.y - p.y); };
                const double h1 = 0.5 * (len(at(i1, i2), at(i1 + 1, i2)) + len(at(i1, i2 + 1), at(i1 + 1, i2 + 1)));
                const double h2 = 0.5 * (len(at(i1, i2), at(i1, i2 + 1)) + len(at(i1 + 1, i2), at(i1 + 1, i2 + 1)));
                ++cells;
                stretched += h1 < 0.5 * h2;
            }
        // (a periodic direction of one segment would couple that segment to
        // itself across the seam, outside its tridiagonal solve)
        L.seg = L.co1 && L.co2 && cells > 0 && stretched >= kSegFrac * cells && !(per1 && L.n1 <= cpk::kSeg);
        if (l + 1 < G->nlev) {
            const MgLevel& C = G->L[l + 1];
            std::vector<double2> pc((size_t)C.nn);
            for (int K1 = 0; K1 < C.n1; ++K1)
                for (int K2 = 0; K2 < C.n2; ++K2)
                    pc[(size_t)K1 * C.n2 + K2] =
                        pos[cpk::mg_pos(L.n1, L.co1, K1, per1) * L.n2 + cpk::mg_pos(L.n2, L.co2, K2)];
            pos.swap(pc);
        }
    }
    if (G->nlev < 2 || 2 * G->L.back().nn > cpk::kMgDenseMax) {
        mg_destroy(P);
        return fail(CP_ERR_INVALID, "cp_problem_set_grid: lattice too small or too thin for multigrid");
    }
    int s = CP_OK;
    for (int l = 0; l < G->nlev && !s; ++l) {
        MgLevel& L = G->L[l];
        if ((s = G->get(&L.lam, 2 * cpk::kLamSlots))) break;
        if (l == 0) continue;
        if ((s = G->get(&L.mask, L.nn))) break;
        if ((s = G->get(&L.slice, L.nslices))) break;
        if ((s = G->get(&L.col, (size_t)L.nslices * cpk::kWave * 9))) break;
        std::vector<int2> hs(L.nslices);
        std::vector<int> hc((size_t)L.nslices * cpk::kWave * 9, 0);
        for (int sl = 0; sl < L.nslices; ++sl) {
            hs[sl] = make_int2(sl * cpk::kWave * 9, 9);
            for (int lane = 0; lane < cpk::kWave; ++lane) {
                const int C = sl * cpk::kWave + lane;
                if (C >= L.nn) continue;
                const int K1 = C / L.n2, K2 = C % L.n2;
                for (int d1 = -1; d1 <= 1; ++d1)
                    for (int d2 = -1; d2 <= 1; ++d2) {
                        const int D1 = cpk::mg_step(L.n1, per1, K1, d1), D2 = K2 + d2;
                        const bool in = D1 >= 0 && D2 >= 0 && D2 < L.n2;
                        hc[(size_t)hs[sl].x + (size_t)cpk::mg_slot(d1, d2) * cpk::kWave + lane] = in ? D1 * L.n2 + D2 : C;
                    }
            }
        }
        if (hipMemcpy(L.slice, hs.data(), sizeof(int2) * hs.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(L.col, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice) != hipSuccess)
            s = fail(CP_ERR_DEVICE, "cp_problem_set_grid: upload failed");
    }
    if (!s && q8) {
        if (!(s = G->get(&G->q8_par, nn)) && !(s = G->get(&G->q8_ch, q8ch.size())))
            if (hipMemcpy(G->q8_par, q8par.data(), sizeof(int2) * nn, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(G->q8_ch, q8ch.data(), sizeof(int) * q8ch.size(), hipMemcpyHostToDevice) != hipSuccess)
                s = fail(CP_ERR_DEVICE, "cp_problem_set_grid: upload failed");
    }
    // level 0 rows: position of each lattice offset in the node's row (the
    // row lists the node itself, then its neighbours, in adjacency order)
    if (!s && !q8) {
        std::vector<signed char> sp((size_t)nn * 9, -1);
        for (int n = 0; n < nn; ++n) {
            const int a1 = lg[n] / n2, a2 = lg[n] % n2;
            for (int q = P->adj_ptr[n]; q < P->adj_ptr[n + 1]; ++q) {
                const int m = P->adj_node[q];
                const int e1 = cpk::mg_offset(n1, per1, a1, lg[m] / n2), e2 = lg[m] % n2 - a2;
                sp[(size_t)n * 9 + cpk::mg_slot(e1, e2)] = (signed char)(q - P->adj_ptr[n]);
            }
        }
        if (!(s = G->get(&G->slot0, sp.size())))
            if (hipMemcpy(G->slot0, sp.data(), sp.size(), hipMemcpyHostToDevice) != hipSuccess)
                s = fail(CP_ERR_DEVICE, "cp_problem_set_grid: upload failed");
    }
    bool identity = !q8;
    for (int n = 0; n < nn && identity; ++n) identity = lg[n] == n;
    G->sym_ok = identity && !per1 && P->npe == 4;
    if (!s && !identity && !q8) {
        std::vector<int> no(nn);
        for (int n = 0; n < nn; ++n) no[lg[n]] = n;
        if (!(s = G->get(&G->lg0, nn)) && !(s = G->get(&G->node0, nn)))
            if (hipMemcpy(G->lg0, lg.data(), sizeof(int) * nn, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(G->node0, no.data(), sizeof(int) * nn, hipMemcpyHostToDevice) != hipSuccess)
                s = fail(CP_ERR_DEVICE, "cp_problem_set_grid: upload failed");
    }
    if (!s) {
        G->nd = 2 * G->L.back().nn;
        s = G->get(&G->dense, (size_t)G->nd * G->nd);
    }
    if (s) {
        mg_destroy(P);
        return s;
    }
    return CP_OK;
}

}  // namespace cp

int cp_problem_set_grid(cp_problem* P, int n1, int n2, const int* node_of)
{
    return cp::set_grid(P, n1, n2, node_of, 0);
}

int cp_problem_set_grid_periodic(cp_problem* P, int n1, int n2, const int* node_of)
{
    return cp::set_grid(P, n1, n2, node_of, 1);
}

int cp_problem_set_graphs(cp_problem* P, int enable)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    if (!P->mg) return cp::fail(CP_ERR_STATE, "cp_problem_set_graphs: no multigrid lattice set");
    P->mg->use_graphs = enable != 0;
    return CP_OK;
}

int cp_problem_mg_levels(cp_problem* P, int* levels, int* dims)
{
    if (!P) return cp::fail(CP_ERR_INVALID, "null problem");
    if (!P->mg) {
        if (levels) *levels = 0;
        return CP_OK;
    }
    if (levels) *levels = P->mg->nlev;
    if (dims)
        for (int l = 0; l < P->mg->nlev; ++l) {
            dims[2 * l] = P->mg->L[l].n1;
            dims[2 * l + 1] = P->mg->L[l].n2;
        }
    return CP_OK;
}

int cp_problem_set_mg_rebuild(cp_problem* P, int every)
{
    if (!P || every < 1) return cp::fail(CP_ERR_INVALID, "cp_problem_set_mg_rebuild: every must be >= 1");
    if (!P->mg) return cp::fail(CP_ERR_STATE, "cp_problem_set_mg_rebuild: no multigrid lattice set");
    P->mg->rebuild = every;
    return CP_OK;
}

int cp_problem_mg_level0_layout(cp_problem* P, int* sym)
{
    if (!P || !sym) return cp::fail(CP_ERR_INVALID, "null argument");
    if (!P->mg) return cp::fail(CP_ERR_STATE, "cp_problem_mg_level0_layout: no multigrid lattice set");
    *sym = P->mg->sym_ok ? 1 : 0;
    return CP_OK;
}

int cp_problem_mg_smoothers(cp_problem* P, int* kinds)
{
    if (!P || !kinds) return cp::fail(CP_ERR_INVALID, "cp_problem_mg_smoothers: bad arguments");
    if (!P->mg) return cp::fail(CP_ERR_STATE, "cp_problem_mg_smoothers: no multigrid lattice set");
    for (int l = 0; l < P->mg->nlev; ++l) kinds[l] = P->mg->L[l].seg;
    return CP_OK;
}

}  // extern "C"
