// cp_mg.h -- gfx950 device code of the geometric multigrid preconditioner.
//
// Structured quad meshes carry a logical grid: node = node_of[i1 * n2 + i2]
// on an n1 x n2 lattice, every matrix coupling within one lattice step.  Each
// coarser level halves the lattice directions chosen for it (coarse k sits
// on fine min(2k, n-1); a direction of 2 or fewer points is kept; on
// stretched meshes only the direction of the short element edges, which
// couples strongly, is halved) with bilinear interpolation P, and carries the Galerkin operator P^T A P of the level
// above, formed on the device as a 3x3-node stencil of 2x2 blocks in the same
// sliced-ELL layout as the fine matrix (slot 0 the node itself, slots 1..8
// the neighbours in raster order; slots outside the lattice point at the
// node with zero values).  Fixed DOFs are removed with the mask: P is zero
// on them, and a coarse DOF that only interpolates to fixed DOFs gets a unit
// diagonal and mask 0.
//
// Smoothing: Chebyshev iteration of degree 3 on D^-1 A over
// [lam/30, 1.1 lam], D the 2x2 block diagonal with lam the Gershgorin bound
// of |D^-1 A| (no eigenvalue iteration), or, on levels whose cells are
// stretched along direction 1, D the block-tridiagonal part along segments
// of that direction with lam = 4 (segment smoother below).  Pre- and
// post-smoothing are the same polynomial, so the V-cycle is a symmetric
// operator and can precondition CG.
// The coarsest level (at most kMgDenseMax DOFs) is solved with an explicit
// inverse formed by Gauss-Jordan elimination in one workgroup.
//
// All sums run in a fixed order: results do not depend on scheduling.
#pragma once

#include <type_traits>

#include "cp_kernels.h"

namespace cpk {

constexpr int kMgDenseMax = 128;   // DOFs of the coarsest level (its inverse stays in registers)
constexpr int kQ8Children = 5;     // fine children of a corner-lattice point: itself and its edges' mid-side nodes
constexpr double kMgLowFrac = 1. / 30.;

// 1-D coarsening: coarse point count of n fine points (co = 0: the
// direction is kept, as when a level is semi-coarsened along the other one).
// A periodic direction (per = 1: point n - 1 neighbours point 0) is halved
// only while it stays even with at least 3 coarse points, so that its two
// neighbours of a point are distinct points: coarse K sits on fine 2K.
__host__ __device__ inline bool mg_per_halves(int n) { return n % 2 == 0 && n >= 6; }
__host__ __device__ inline int mg_coarse_n(int n, int co = 1, int per = 0)
{
    if (per) return co ? n / 2 : n;
    return (n <= 2 || !co) ? n : n / 2 + 1;
}

// Coarse parents of fine point f of an n-point direction and twice their
// bilinear weights, with integer selects only (no branches, so the code
// stays straight-line inside divergent regions).  A direction not coarsened,
// or an open one of at most 2 points: f itself.  Otherwise an even
// f sits on coarse point f/2; an odd f lies between f/2 and f/2 + 1 (the
// latter wrapping to 0 where the direction is periodic), except the last
// point of an open direction of even length, which is coarse point n/2.
// The parents come as q0, q1 with twice-weights c0, c1: 2 and 0 where f has
// one parent (q1 repeats it), 1 and 1 where it has two.
__device__ __forceinline__ void mg_parents_sel(int n, int co, int f, int per, int& q0, int& q1, int& c0, int& c1)
{
    const int h = f >> 1;
    const bool odd = (f & 1) != 0, pc = per && co, same = !pc && (n <= 2 || !co);
    const bool one = same || !odd || (!pc && f == n - 1);
    const int k1 = same ? f : (!odd ? h : n / 2);
    q0 = one ? k1 : h;
    q1 = one ? k1 : ((pc && h + 1 == n / 2) ? 0 : h + 1);
    c0 = one ? 2 : 1;
    c1 = one ? 0 : 1;
}

// Twice the weight (0, 1 or 2) of coarse point K in the interpolation to
// fine point f.
__device__ __forceinline__ int mg_weight2(int n, int co, int f, int K, int per = 0)
{
    int q0, q1, c0, c1;
    mg_parents_sel(n, co, f, per, q0, q1, c0, c1);
    return (K == q0 ? c0 : 0) + (K == q1 ? c1 : 0);
}

// Weight of coarse point K in the interpolation to fine point f (0 if none).
__device__ __forceinline__ double mg_weight(int n, int co, int f, int K, int per = 0)
{
    return 0.5 * (double)mg_weight2(n, co, f, K, per);
}

// Fine position of coarse point K.
__host__ __device__ inline int mg_pos(int n, int co, int K, int per = 0)
{
    if (per) return co ? 2 * K : K;
    return (n <= 2 || !co) ? K : (2 * K < n - 1 ? 2 * K : n - 1);
}

// Point f + d of an n-point direction, wrapped when it is periodic; -1
// outside a direction that is not.
__host__ __device__ inline int mg_step(int n, int per, int f, int d)
{
    const int q = f + d;
    if (per) return q < 0 ? q + n : (q >= n ? q - n : q);
    return (q < 0 || q >= n) ? -1 : q;
}

// Offset of point g from point f in [-n/2, n/2) along a periodic direction.
__host__ __device__ inline int mg_offset(int n, int per, int f, int g)
{
    int e = g - f;
    if (per) {
        if (e > n / 2) e -= n;
        else if (e < -(n - 1) / 2) e += n;
    }
    return e;
}

// Stencil slot of the coarse neighbour at offset (d1, d2), |d| <= 1.
__host__ __device__ inline int mg_slot(int d1, int d2)
{
    const int idx = (d1 + 1) * 3 + (d2 + 1);
    return idx < 4 ? idx + 1 : (idx == 4 ? 0 : idx);
}

template <typename T> struct ChebCoef { T c0, a1, b1, a2, b2; };

// A level's lambda and, after it, the Chebyshev coefficients it gives
// (kLamSlots doubles per precision): written together at set-up, read by
// every smoothing step instead of formed again in each thread.
constexpr int kLamSlots = 8;

// Coefficients of the first three steps of the Chebyshev three-term
// recurrence on [lam/30, 1.1 lam]; step k's coefficients do not depend on
// the degree, so degree 2 uses c0, a1, b1 only.
template <typename T>
__device__ __forceinline__ ChebCoef<T> mg_cheb(double lam)
{
    const double hi = 1.1 * lam, lo = kMgLowFrac * lam;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    const double sigma = theta / delta;
    const double r0 = 1. / sigma;
    const double r1 = 1. / (2. * sigma - r0);
    const double r2 = 1. / (2. * sigma - r1);
    return ChebCoef<T>{(T)(1. / theta), (T)(r1 * r0), (T)(2. * r1 / delta), (T)(r2 * r1), (T)(2. * r2 / delta)};
}

// Sets a level's lambda and its coefficients (one thread).
__device__ __forceinline__ void mg_set_lam(double* lam, double value)
{
    const ChebCoef<double> c = mg_cheb<double>(value);
    lam[0] = value;
    lam[1] = c.c0;
    lam[2] = c.a1;
    lam[3] = c.b1;
    lam[4] = c.a2;
    lam[5] = c.b2;
}

// The coefficients stored by mg_set_lam, in precision T.
template <typename T>
__device__ __forceinline__ ChebCoef<T> mg_cheb_of(const double* lam)
{
    return ChebCoef<T>{(T)lam[1], (T)lam[2], (T)lam[3], (T)lam[4], (T)lam[5]};
}


__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ double block_max(double v)
{
    __shared__ double sh[4];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) v = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    __syncthreads();
    return v;
}

// ------------------------------------------------------------ set-up kernels

template <typename TS, typename TD>
struct GalArgs {
    CgArgs<TS> f;          // fine level: matrix, mask
    const int* lg;         // fine node -> lattice index (nullptr: identity)
    const int* node_of;    // fine lattice index -> node (nullptr: identity)
    const signed char* slot_pos;   // fine node x 9 lattice offsets -> row position, -1 if absent
                                   // (nullptr: the coarse-level layout, position = slot)
    int n1f, n2f;
    int co1, co2;          // directions coarsened to the next level
    int per1;              // direction 1 periodic
    int n1c, n2c, nnc, nbc;
    const int2* slice_c;
    double2* va_c;         // FP64 destination planes
    double2* vb_c;
    float4* vf_c;          // FP32 destination
    uchar2* mask_c;
    // 8-node fine level (k_mg_galerkin_q8): the corner lattice's transfer
    const int2* q8_par;    // fine node -> its coarse parents (one for a corner, two for a mid-side node)
    const int* q8_ch;      // coarse point -> its fine children (kQ8Children, -1 padded)
};

template <typename T>
__device__ __forceinline__ void load_block_d(const CgArgs<T>& a, size_t q, double& m00, double& m01, double& m10,
                                             double& m11)
{
    T b00, b01, b10, b11;
    load_block(a, q, b00, b01, b10, b11);
    m00 = b00;
    m01 = b01;
    m10 = b10;
    m11 = b11;
}

// Galerkin product in tiles of kGalT x kGalT coarse lattice points: the
// workgroup first copies the rows of every fine point its coarse points
// interpolate to (at most kGalR x kGalR points, 9 blocks each in lattice
// offset order) into LDS, each block with the fine DOFs the mask fixes
// zeroed (its row by the point's mask, its columns by the neighbour's).
// One thread per (coarse point C, stencil slot t) then forms
// (P^T A P)(C, D_t), D_t the coarse neighbour of slot t: the sum over the
// fine children f of C and g of D_t that are lattice neighbours of
// w_f w_g A_fg, in lattice order, from LDS.  A_fg is found through the fine
// row's slot table (offset (g - f) -> storage position; level 0 carries
// one, coarser levels store slots in offset order).  Each fine row is read
// from memory once per tile instead of once per (coarse point, slot) that
// needs it.
constexpr int kGalT = 8;
constexpr int kGalThreads = 9 * kGalT * kGalT;
constexpr int kGalR = 2 * kGalT + 1;

// Children of coarse point K in an n-point direction: the fine points at
// offsets -1, 0, 1 from its position, their interpolation weights (0: not a
// child) and their place in a tile region whose first fine point (before
// wrapping) is b.  Fixed-size, fully unrolled, so the arrays stay in
// registers.
__device__ __forceinline__ void mg_children_local(int n, int co, int per, int K, int b, int (&f)[3], double (&w)[3],
                                                  int (&loc)[3])
{
    const int p = mg_pos(n, co, K, per);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int q = mg_step(n, per, p, a - 1);
        f[a] = q;
        w[a] = q < 0 ? 0. : mg_weight(n, co, q, K, per);
        loc[a] = p + a - 1 - b;
    }
}

// Fine region of the coarse points klo..khi of one direction: its first
// point b (may be negative along a periodic direction: wrapped on use) and
// its point count.
__device__ __forceinline__ void mg_gal_region(int n, int co, int per, int klo, int khi, int& b, int& r)
{
    b = mg_pos(n, co, klo, per) - 1;
    int e = mg_pos(n, co, khi, per) + 1;
    if (!per) {
        b = b < 0 ? 0 : b;
        e = e > n - 1 ? n - 1 : e;
    }
    r = e - b + 1;
}

template <typename TS, typename TD>
__global__ __launch_bounds__(kGalThreads) void k_mg_galerkin(GalArgs<TS, TD> g, int nt2, int ntiles)
{
    __shared__ TS blk[kGalR * kGalR * 9 * 4];
    const int tile = xcd_block(ntiles);
    const int T1 = tile / nt2, T2 = tile - T1 * nt2;
    const int K1lo = T1 * kGalT, K2lo = T2 * kGalT;
    const int K1hi = min(K1lo + kGalT, g.n1c) - 1, K2hi = min(K2lo + kGalT, g.n2c) - 1;
    int b1, R1, b2, R2;
    mg_gal_region(g.n1f, g.co1, g.per1, K1lo, K1hi, b1, R1);
    mg_gal_region(g.n2f, g.co2, 0, K2lo, K2hi, b2, R2);
    for (int i = threadIdx.x; i < R1 * R2 * 9; i += kGalThreads) {
        const int pt = i / 9, so = i - 9 * pt;
        const int l1 = pt / R2, l2 = pt - l1 * R2;
        int q1 = b1 + l1;
        if (g.per1) q1 = q1 < 0 ? q1 + g.n1f : (q1 >= g.n1f ? q1 - g.n1f : q1);
        const int q2 = b2 + l2;
        TS m00 = 0, m01 = 0, m10 = 0, m11 = 0;
        if (q1 >= 0 && q1 < g.n1f && q2 >= 0 && q2 < g.n2f) {
            const int lf = q1 * g.n2f + q2;
            const int fn = g.node_of ? g.node_of[lf] : lf;
            const int pos = g.slot_pos ? g.slot_pos[(size_t)fn * 9 + so] : so;
            if (pos >= 0) {
                const size_t q = (size_t)g.f.slice[fn >> 6].x + (fn & 63) + (size_t)pos * kWave;
                const uchar2 mf = g.f.mask[fn], mg = g.f.mask[g.f.col[q]];
                load_block(g.f, q, m00, m01, m10, m11);
                if (!(mf.x && mg.x)) m00 = 0;
                if (!(mf.x && mg.y)) m01 = 0;
                if (!(mf.y && mg.x)) m10 = 0;
                if (!(mf.y && mg.y)) m11 = 0;
            }
        }
        TS* d = blk + (size_t)i * 4;
        d[0] = m00;
        d[1] = m01;
        d[2] = m10;
        d[3] = m11;
    }
    __syncthreads();
    const int cl = threadIdx.x / 9, t = threadIdx.x - 9 * cl;
    const int K1 = K1lo + cl / kGalT, K2 = K2lo + cl % kGalT;
    if (K1 > K1hi || K2 > K2hi) return;
    const int C = K1 * g.n2c + K2;
    // slot t -> coarse offset (d1, d2)
    const int idx = t == 0 ? 4 : (t <= 4 ? t - 1 : t);
    const int d1 = idx / 3 - 1, d2 = idx % 3 - 1;
    const int D1 = mg_step(g.n1c, g.per1, K1, d1), D2 = K2 + d2;
    double acc0 = 0., acc1 = 0., acc2 = 0., acc3 = 0.;
    if (D1 >= 0 && D2 >= 0 && D2 < g.n2c) {
        int fa[3], fb[3], ga[3], gb[3], la[3], lb[3], lx[3];
        double wfa[3], wfb[3], wga[3], wgb[3];
        mg_children_local(g.n1f, g.co1, g.per1, K1, b1, fa, wfa, la);
        mg_children_local(g.n2f, g.co2, 0, K2, b2, fb, wfb, lb);
        mg_children_local(g.n1f, g.co1, g.per1, D1, b1, ga, wga, lx);
        mg_children_local(g.n2f, g.co2, 0, D2, b2, gb, wgb, lx);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (wfa[i] == 0.) continue;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (wfb[j] == 0.) continue;
                const TS* row = blk + (size_t)(la[i] * R2 + lb[j]) * 36;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (wga[k] == 0.) continue;
                    const int e1 = mg_offset(g.n1f, g.per1, fa[i], ga[k]);
                    if (e1 < -1 || e1 > 1) continue;
#pragma unroll
                    for (int l = 0; l < 3; ++l) {
                        if (wgb[l] == 0.) continue;
                        const int e2 = gb[l] - fb[j];
                        if (e2 < -1 || e2 > 1) continue;
                        const TS* m = row + 4 * mg_slot(e1, e2);
                        const double ww = wfa[i] * wfb[j] * wga[k] * wgb[l];
                        acc0 += ww * (double)m[0];
                        acc1 += ww * (double)m[1];
                        acc2 += ww * (double)m[2];
                        acc3 += ww * (double)m[3];
                    }
                }
            }
        }
    }
    if (t == 0) {
        uchar2 mc;
        mc.x = acc0 > 0. ? 1 : 0;
        mc.y = acc3 > 0. ? 1 : 0;
        if (!mc.x) acc0 = 1.;
        if (!mc.y) acc3 = 1.;
        g.mask_c[C] = mc;
    }
    const size_t q = (size_t)g.slice_c[C >> 6].x + (C & 63) + (size_t)t * kWave;
    if (g.va_c) {
        g.va_c[q] = make_double2(acc0, acc1);
        g.vb_c[q] = make_double2(acc2, acc3);
    } else {
        g.vf_c[q] = make_float4((float)acc0, (float)acc1, (float)acc2, (float)acc3);
    }
}

// Block-Jacobi inverse (as k_block_jacobi) and the per-block maximum of the
// Gershgorin row sums of |D^-1 A|.
// Galerkin operator of the corner lattice under an 8-node fine level:
// A_c(K, D) = sum over K's children f and D's children g of
// w_f A_fg w_g, the weights of the transfer the V-cycle uses (a corner is
// its own child with weight 1, each mid-side node of an edge a child of the
// edge's two corners with weight 1/2).  One thread per coarse point and
// stencil slot, as k_mg_galerkin; the fine row of each child is walked
// through its column indices and every column's parents are looked up.
template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_mg_galerkin_q8(GalArgs<TS, TD> g)
{
    const int id = xcd_block(g.nbc) * kBlock + threadIdx.x;
    if (id >= 9 * g.nnc) return;
    const int C = id / 9, t = id - 9 * C;
    const int K1 = C / g.n2c, K2 = C - K1 * g.n2c;
    const int idx = t == 0 ? 4 : (t <= 4 ? t - 1 : t);
    const int d1 = idx / 3 - 1, d2 = idx % 3 - 1;
    const int D1 = mg_step(g.n1c, g.per1, K1, d1), D2 = K2 + d2;
    double acc0 = 0., acc1 = 0., acc2 = 0., acc3 = 0.;
    if (D1 >= 0 && D2 >= 0 && D2 < g.n2c) {
        const int D = D1 * g.n2c + D2;
        for (int j = 0; j < kQ8Children; ++j) {
            const int f = g.q8_ch[(size_t)C * kQ8Children + j];
            if (f < 0) continue;
            const double wf = j == 0 ? 1. : 0.5;
            const uchar2 mf = g.f.mask[f];
            const int2 sl = g.f.slice[f >> 6];
            const size_t base = (size_t)sl.x + (f & 63);
            for (int k = 0; k < sl.y; ++k) {
                const size_t q = base + (size_t)k * kWave;
                const int gn = g.f.col[q];
                const int2 pp = g.q8_par[gn];
                const double wg = pp.y < 0 ? (pp.x == D ? 1. : 0.) : ((pp.x == D || pp.y == D) ? 0.5 : 0.);
                if (wg == 0.) continue;   // (padding blocks of the row are zero)
                const uchar2 mg = g.f.mask[gn];
                double m00, m01, m10, m11;
                load_block_d(g.f, q, m00, m01, m10, m11);
                const double ww = wf * wg;
                acc0 += (mf.x && mg.x) ? ww * m00 : 0.;
                acc1 += (mf.x && mg.y) ? ww * m01 : 0.;
                acc2 += (mf.y && mg.x) ? ww * m10 : 0.;
                acc3 += (mf.y && mg.y) ? ww * m11 : 0.;
            }
        }
    }
    if (t == 0) {
        uchar2 mc;
        mc.x = acc0 > 0. ? 1 : 0;
        mc.y = acc3 > 0. ? 1 : 0;
        if (!mc.x) acc0 = 1.;
        if (!mc.y) acc3 = 1.;
        g.mask_c[C] = mc;
    }
    const size_t q = (size_t)g.slice_c[C >> 6].x + (C & 63) + (size_t)t * kWave;
    if (g.va_c) {
        g.va_c[q] = make_double2(acc0, acc1);
        g.vb_c[q] = make_double2(acc2, acc3);
    } else {
        g.vf_c[q] = make_float4((float)acc0, (float)acc1, (float)acc2, (float)acc3);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_smooth_setup(CgArgs<T> a, typename V4<T>::type* minv,
                                                            double* partials)
{
    const int lb = xcd_block(a.nblocks);
    const int n = lb * kBlock + threadIdx.x;
    double v = 0.;
    if (n < a.nn) {
        const int2 sl = a.slice[n >> 6];
        const size_t base = (size_t)sl.x + (n & 63);
        double d00, d01, d10, d11;
        load_block_d(a, base, d00, d01, d10, d11);
        const double s01 = 0.5 * (d01 + d10);
        const uchar2 mk = a.mask[n];
        double i00 = 0., i01 = 0., i11 = 0.;
        if (mk.x && mk.y) {
            const double det = d00 * d11 - s01 * s01;
            i00 = d11 / det;
            i01 = -s01 / det;
            i11 = d00 / det;
        } else if (mk.x) {
            i00 = 1. / d00;
        } else if (mk.y) {
            i11 = 1. / d11;
        }
        minv[n] = typename V4<T>::type{(T)i00, (T)i01, (T)i11, (T)0};
        double rx = 0., ry = 0.;
        for (int k = 0; k < sl.y; ++k) {
            const size_t q = base + (size_t)k * kWave;
            const int m = a.col[q];
            double m00, m01, m10, m11;
            load_block_d(a, q, m00, m01, m10, m11);
            const uchar2 mm = a.mask[m];
            m00 = mm.x ? m00 : 0.;
            m10 = mm.x ? m10 : 0.;
            m01 = mm.y ? m01 : 0.;
            m11 = mm.y ? m11 : 0.;
            rx += fabs(i00 * m00 + i01 * m10) + fabs(i00 * m01 + i01 * m11);
            ry += fabs(i01 * m00 + i11 * m10) + fabs(i01 * m01 + i11 * m11);
        }
        v = fmax(rx, ry);
    }
    v = block_max(v);
    if (threadIdx.x == 0) partials[lb] = v;
}

static __global__ __launch_bounds__(kBlock) void k_mg_lam(const double* partials, int nblocks, double* lam)
{
    double v = 0.;
    for (int j = threadIdx.x; j < nblocks; j += kBlock) v = fmax(v, partials[j]);
    v = block_max(v);
    if (threadIdx.x == 0) mg_set_lam(lam, v > 0. ? v : 1.);
}

// Dense copy of a level's operator (nd = 2 nn DOFs), one node's two rows per thread.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_dense(CgArgs<T> a, double* D, int nd)
{
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    double* r0 = D + (size_t)(2 * n) * nd;
    double* r1 = r0 + nd;
    for (int j = 0; j < nd; ++j) r0[j] = r1[j] = 0.;
    const int2 sl = a.slice[n >> 6];
    const size_t base = (size_t)sl.x + (n & 63);
    for (int k = 0; k < sl.y; ++k) {
        const size_t q = base + (size_t)k * kWave;
        const int m = a.col[q];
        double m00, m01, m10, m11;
        load_block_d(a, q, m00, m01, m10, m11);
        r0[2 * m] += m00;
        r0[2 * m + 1] += m01;
        r1[2 * m] += m10;
        r1[2 * m + 1] += m11;
    }
}

// In-place inverse of the SPD matrix D (nd x nd, nd <= kMgDenseMax) by
// Gauss-Jordan elimination without pivoting, in one workgroup of 1024
// threads: thread (j = t % 128, i0 = t / 128) keeps D[i][j] for the rows
// i = i0, i0 + 8, ... in registers; each step broadcasts the pivot row and
// column through LDS.  The update is written without branches (selects),
// so the LDS reads of a step are issued together.
static __global__ __launch_bounds__(1024) void k_mg_invert(double* D, int nd)
{
    constexpr int G = 8, R = kMgDenseMax / G;
    __shared__ double colk[kMgDenseMax];
    __shared__ double rowk[kMgDenseMax];
    const int j = threadIdx.x % kMgDenseMax, i0 = threadIdx.x / kMgDenseMax;
    const bool active = j < nd;
    if (threadIdx.x < kMgDenseMax) colk[threadIdx.x] = rowk[threadIdx.x] = 0.;
    double v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + G * r;
        v[r] = (active && i < nd) ? D[(size_t)i * nd + j] : 0.;
    }
    __syncthreads();
    for (int k = 0; k < nd; ++k) {
        // column k and row k of the current matrix to LDS
        if (j == k) {
#pragma unroll
            for (int r = 0; r < R; ++r) colk[i0 + G * r] = v[r];
        }
        if (active && i0 == k % G) {
            const int rk = k / G;
            double t = 0.;
#pragma unroll
            for (int r = 0; r < R; ++r) t = r == rk ? v[r] : t;
            rowk[j] = t;
        }
        __syncthreads();
        double c[R];
#pragma unroll
        for (int r = 0; r < R; ++r) c[r] = colk[i0 + G * r];
        const double dk = colk[k];
        const double p = dk != 0. ? 1. / dk : 0.;
        const double rkj = rowk[j] * p;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = i0 + G * r;
            double nv = v[r] - c[r] * rkj;              // the rest
            nv = j == k ? -c[r] * p : nv;               // column k
            nv = i == k ? (j == k ? p : rkj) : nv;      // row k: scaled; pivot: 1 / d
            v[r] = (active && i < nd) ? nv : v[r];
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + G * r;
        if (active && i < nd) D[(size_t)i * nd + j] = v[r];
    }
}

// ------------------------------------------------------------ V-cycle kernels

template <typename T>
struct MgVec {
    using V = typename V2<T>::type;
    const V* b;      // right-hand side of this level
    V* xa;           // iterate buffers (ping-pong)
    V* xb;
    V* da;           // Chebyshev directions (ping-pong)
    V* db;
    V* res;          // residual handed to the coarser level
    const double* lam;
    const CgScal* sc;   // CG scalars: nothing to do once the solve is done
};

enum { MG_PRE1 = 0, MG_CHEB = 1, MG_RESID = 2, MG_POST1 = 3 };
// Steps whose input iterate comes combined (the previous step wrote x + d
// as well as d, so a neighbour's iterate is one load, not two): the
// residual step MG_RESID_X, and the segment steps MG_POST1_X, MG_CHEB_X,
// SEG_FIRST_X.  The additions are the same ones, made once by the writer
// instead of by every reader, so the results are bitwise the same.
enum { MG_RESID_X = 7, MG_POST1_X = 8, MG_CHEB_X = 9, SEG_FIRST_X = 10 };

// Coarse levels store lattice node n's row in slice n >> 6 (9 slots of 64
// lanes each) with slot k holding the neighbour at a fixed lattice offset
// (itself where the lattice ends, with zero values), so the row's
// addresses and neighbours follow from (i1, i2) without loading col / slice.
__device__ __forceinline__ int lat_neighbour(int k, int i1, int i2, int n1, int n2, int n, int per1 = 0)
{
    const int idx = k == 0 ? 4 : (k <= 4 ? k - 1 : k);
    const int d1 = idx / 3 - 1, d2 = idx % 3 - 1;
    const int j1 = mg_step(n1, per1, i1, d1), j2 = i2 + d2;
    return (j1 >= 0 && j2 >= 0 && j2 < n2) ? j1 * n2 + j2 : n;
}

__device__ __forceinline__ size_t lat_base(int n) { return (size_t)(n >> 6) * (9 * kWave) + (n & 63); }

template <typename T, typename ZF>
__device__ __forceinline__ void row_product_lat(const CgArgs<T>& a, int n, int i1, int i2, const ZF& zfun, T& yx,
                                                T& yy)
{
    const size_t base = lat_base(n);
    T m[9][4];
#pragma unroll
    for (int k = 0; k < 9; ++k) load_block(a, base + (size_t)k * kWave, m[k][0], m[k][1], m[k][2], m[k][3]);
    yx = 0;
    yy = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        T zx, zy;
        zfun(lat_neighbour(k, i1, i2, a.lat_n1, a.lat_n2, n, a.lat_per1), zx, zy);
        yx += m[k][0] * zx + m[k][1] * zy;
        yy += m[k][2] * zx + m[k][3] * zy;
    }
}

// Row width template value of the coarse levels: lattice addressing.
constexpr int kLatW = -1;

// Row width template value of level 0 in the symmetric lattice layout.
// Where the nodes are numbered lattice point by lattice point (rectangles,
// quarter plates and rings) the finest operator is also stored as five
// blocks per node: its diagonal block (made symmetric) and its couplings to
// the four "forward" neighbours at lattice offsets (0, 1), (1, -1), (1, 0),
// (1, 1), in slots 0..4 of 64-node slices (slot k of node n at
// (n >> 6) * 5 * 64 + k * 64 + (n & 63)); a neighbour outside the lattice
// has a zero block.  The coupling to a backward neighbour m = n - offset is
// the transpose of m's forward block: A is symmetric, so a row reads 80
// bytes of its own and the rest from rows the neighbouring lanes and tiles
// read anyway, instead of 9 blocks and 9 column indices (180 bytes).
constexpr int kSymW = -2;
// The same layout on a launched coarse level (the 16-column tiles of the
// coarse levels' segment smoother).
constexpr int kSymW16 = -3;
template <int W> constexpr bool sym_rows() { return W == kSymW || W == kSymW16; }
constexpr int kSymSlots = 5;

__host__ __device__ constexpr int sym_d1(int k) { return k == 0 || k == 1 ? 0 : 1; }
__host__ __device__ constexpr int sym_d2(int k) { return k == 0 ? 0 : (k == 1 ? 1 : k - 3); }
__device__ __forceinline__ size_t sym_base(int n) { return (size_t)(n >> 6) * (kSymSlots * kWave) + (n & 63); }

__device__ __forceinline__ void load_sym(const CgArgs<float>& a, size_t q, float (&m)[4])
{
    const nf4 u = reinterpret_cast<const nf4*>(a.vsym)[q];
    m[0] = u.x; m[1] = u.y; m[2] = u.z; m[3] = u.w;
}
// (the layout is built for FP32 only; the FP64 form exists for the templates)
__device__ __forceinline__ void load_sym(const CgArgs<double>&, size_t, double (&m)[4]) { m[0] = m[1] = m[2] = m[3] = 0.; }

// y = A z at node n of level 0 in the symmetric lattice layout.  Own blocks
// whose neighbour lies outside the lattice are zero, and so are the
// forward blocks a backward neighbour index reaches across a lattice column
// (its own forward neighbour would lie outside too), so only indices
// outside [0, nn) need a guard.
template <typename T, typename ZF>
__device__ __forceinline__ void row_product_sym(const CgArgs<T>& a, int n, const ZF& zfun, T& yx, T& yy)
{
    const int n2 = a.lat_n2;
    T m[kSymSlots][4], t[kSymSlots][4];
    int fw[kSymSlots], bw[kSymSlots];
    const size_t base = sym_base(n);
#pragma unroll
    for (int k = 0; k < kSymSlots; ++k) {
        load_sym(a, base + (size_t)k * kWave, m[k]);
        const int off = sym_d1(k) * n2 + sym_d2(k);
        fw[k] = n + off < a.nn ? n + off : n;
        bw[k] = n - off;
        if (k > 0 && bw[k] >= 0) {
            load_sym(a, sym_base(bw[k]) + (size_t)k * kWave, t[k]);
        } else {
            t[k][0] = t[k][1] = t[k][2] = t[k][3] = 0;
            bw[k] = n;
        }
    }
    yx = 0;
    yy = 0;
#pragma unroll
    for (int k = 0; k < kSymSlots; ++k) {
        T zx, zy;
        zfun(fw[k], zx, zy);
        yx += m[k][0] * zx + m[k][1] * zy;
        yy += m[k][2] * zx + m[k][3] * zy;
        if (k > 0) {
            zfun(bw[k], zx, zy);
            yx += t[k][0] * zx + t[k][2] * zy;
            yy += t[k][1] * zx + t[k][3] * zy;
        }
    }
}

// Level 0's operator in the symmetric lattice layout from its sliced-ELL
// rows (slot_pos: node x 9 lattice offsets -> row position, -1 if none).
__global__ __launch_bounds__(kBlock) void k_mg_sym_build(CgArgs<float> a, const signed char* slot_pos, float4* vs)
{
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    const int2 sl = a.slice[n >> 6];
    const size_t row = (size_t)sl.x + (n & 63), base = sym_base(n);
#pragma unroll
    for (int k = 0; k < kSymSlots; ++k) {
        const int pos = slot_pos[(size_t)n * 9 + mg_slot(sym_d1(k), sym_d2(k))];
        float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
        if (pos >= 0) {
            float m00, m01, m10, m11;
            load_block(a, row + (size_t)pos * kWave, m00, m01, m10, m11);
            if (k == 0) m01 = m10 = 0.5f * (m01 + m10);
            b = make_float4(m00, m01, m10, m11);
        }
        vs[base + (size_t)k * kWave] = b;
    }
}

// The same layout for a coarse level from its lattice rows (FP32 slots of
// fixed offsets), levels of an open lattice outside the tail.
__global__ __launch_bounds__(kBlock) void k_mg_sym_from_lat(const float4* lat, int nn, float4* vs)
{
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= nn) return;
    const size_t base = sym_base(n);
#pragma unroll
    for (int k = 0; k < kSymSlots; ++k) {
        float4 b = lat[lat_base(n) + (size_t)mg_slot(sym_d1(k), sym_d2(k)) * kWave];
        if (k == 0) b.y = b.z = 0.5f * (b.y + b.z);
        vs[base + (size_t)k * kWave] = b;
    }
}

// Level 0's operator assembled straight into the symmetric lattice layout,
// for the set-ups that keep the coarse levels (no Galerkin product needs
// the full rows): the bits k_assemble<W, false> followed by k_mg_sym_build
// store, on a lattice numbered node by node with n2 >= 3.  Element node m
// of an incidence of node n lies at offset m - n: 0 (slot 0), 1, n2 - 1, n2
// or n2 + 1 (the forward slots 1..4), or below 0 (a backward coupling, which
// that neighbour stores).  Each slot sums the same incidences in the same
// order as k_assemble; the diagonal block also goes, unsymmetrised, to
// position 0 of the node's FP32 row, where the segment set-up reads it.
// Registers held to three waves per SIMD (168; an 8-byte spill), which
// measured faster than the two waves the unconstrained 170 allow.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_assemble_sym(AsmArgs a, int n2,
                                                                                               float4* vs)
{
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    double acc[kSymSlots][4];
#pragma unroll
    for (int k = 0; k < kSymSlots; ++k) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.;
    for (int q = a.incPtr[n]; q < a.incPtr[n + 1]; ++q) {
        const int2 ic = a.inc[q];
        int sl[4], i;
        incidence_slots((unsigned)ic.y, sl, i);
        const int4 c = a.conn[ic.x];
        double blk[4][4];
        element_row_blocks(a, c, a.ecnu[ic.x], i, blk);
        const int cn[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int d = cn[j] - n;
            const int k = d == 0 ? 0 : (d == 1 ? 1 : (d == n2 - 1 ? 2 : (d == n2 ? 3 : (d == n2 + 1 ? 4 : -1))));
#pragma unroll
            for (int s = 0; s < kSymSlots; ++s) {
                if (s == k) {
                    acc[s][0] += blk[j][0] * a.h;
                    acc[s][1] += blk[j][1] * a.h;
                    acc[s][2] += blk[j][2] * a.h;
                    acc[s][3] += blk[j][3] * a.h;
                }
            }
        }
    }
    const size_t base = sym_base(n);
#pragma unroll
    for (int k = 0; k < kSymSlots; ++k) {
        float4 b = make_float4((float)acc[k][0], (float)acc[k][1], (float)acc[k][2], (float)acc[k][3]);
        if (k == 0) {
            a.vf[(size_t)a.slice[n >> 6].x + (n & 63)] = b;
            b.y = b.z = 0.5f * (b.y + b.z);
        }
        vs[base + (size_t)k * kWave] = b;
    }
}

// The couplings of level-0 node n to its previous (lo) and next (up)
// neighbour along direction 1 in the symmetric lattice layout: the next
// one is n's slot 3, the previous one the transpose of slot 3 of n - n2.
template <typename T>
__device__ __forceinline__ void sym_seg_couplings(const CgArgs<T>& a, int n, bool lo, bool up, T (&bl)[4], T (&bu)[4])
{
    if (lo) {
        T t[4];
        load_sym(a, sym_base(n - a.lat_n2) + 3 * (size_t)kWave, t);
        bl[0] = t[0];
        bl[1] = t[2];
        bl[2] = t[1];
        bl[3] = t[3];
    }
    if (up) load_sym(a, sym_base(n) + 3 * (size_t)kWave, bu);
}

// Row product of a multigrid level: lattice addressing on coarse levels
// (W == kLatW), else the sliced-ELL row of width W.
template <int W, typename T, typename ZF>
__device__ __forceinline__ void row_apply(const CgArgs<T>& a, int n, const ZF& zfun, T& yx, T& yy)
{
    if constexpr (W == kLatW) {
        const int i1 = n / a.lat_n2;
        row_product_lat<T>(a, n, i1, n - i1 * a.lat_n2, zfun, yx, yy);
    } else if constexpr (sym_rows<W>()) {
        row_product_sym<T>(a, n, zfun, yx, yy);
    } else {
        row_product<W, T>(a, n, zfun, yx, yy);
    }
}

// One smoothing step of a level.  The iterate at the neighbours is formed
// on the fly from the previous step's buffers, so each step is one launch:
//   MG_PRE1  x1 = c0 D^-1 b                    -> xo = x1, do = a1 x1 + b1 D^-1 (b - A x1)
//   MG_CHEB  x = xi + di                       -> xo = x (or x + d if last), do = a di + c D^-1 (b - A x)
//   MG_RESID x = xi + di                       -> xo = x, res = b - A x
//   MG_POST1 x = xi                            -> do = c0 D^-1 (b - A x)
//   MG_RESID_X x = xi (the previous step wrote x + d) -> res = b - A x
template <int W, int MODE, typename T>
__device__ __forceinline__ void mg_smooth_node(const CgArgs<T>& a, const MgVec<T>& v, const typename V2<T>::type* xi,
                                               const typename V2<T>::type* di, typename V2<T>::type* xo,
                                               typename V2<T>::type* dout, int second, int last, int n)
{
    using V = typename V2<T>::type;
    const ChebCoef<T> c = mg_cheb_of<T>(v.lam);
    const T ca = second ? c.a2 : c.a1, cb = second ? c.b2 : c.b1;
    T yx, yy;
    row_apply<W, T>(a, n,
                      [&](int m, T& zx, T& zy) {
                          if constexpr (MODE == MG_PRE1) {
                              const V bm = v.b[m];
                              apply_minv<T>(a.minv[m], bm.x, bm.y, zx, zy);
                              zx *= c.c0;
                              zy *= c.c0;
                          } else if constexpr (MODE == MG_POST1 || MODE == MG_RESID_X) {
                              const V xm = xi[m];
                              zx = xm.x;
                              zy = xm.y;
                          } else {
                              const V xm = xi[m], dm = di[m];
                              zx = xm.x + dm.x;
                              zy = xm.y + dm.y;
                          }
                      },
                      yx, yy);
    const uchar2 mk = a.mask[n];
    const V bn = v.b[n];
    const T rx = mk.x ? bn.x - yx : (T)0, ry = mk.y ? bn.y - yy : (T)0;
    if constexpr (MODE == MG_RESID_X) {
        v.res[n] = V{rx, ry};
        return;
    }
    const typename V4<T>::type mi = a.minv[n];
    T gx, gy;
    apply_minv<T>(mi, rx, ry, gx, gy);
    if constexpr (MODE == MG_PRE1) {
        T x1x, x1y;
        apply_minv<T>(mi, bn.x, bn.y, x1x, x1y);
        x1x *= c.c0;
        x1y *= c.c0;
        xo[n] = V{x1x, x1y};
        dout[n] = V{c.a1 * x1x + c.b1 * gx, c.a1 * x1y + c.b1 * gy};
    } else if constexpr (MODE == MG_POST1) {
        dout[n] = V{c.c0 * gx, c.c0 * gy};
    } else if constexpr (MODE == MG_CHEB) {
        const V xn = xi[n], dn = di[n];
        const V x = V{xn.x + dn.x, xn.y + dn.y};
        const V d = V{ca * dn.x + cb * gx, ca * dn.y + cb * gy};
        if (last) {
            xo[n] = V{x.x + d.x, x.y + d.y};
        } else {
            xo[n] = x;
            dout[n] = d;
        }
    } else {
        const V xn = xi[n], dn = di[n];
        xo[n] = V{xn.x + dn.x, xn.y + dn.y};
        v.res[n] = V{rx, ry};
    }
}

template <int W, int MODE, typename T>
__global__ __launch_bounds__(kBlock) void k_mg_smooth(CgArgs<T> a, MgVec<T> v, const typename V2<T>::type* xi,
                                                      const typename V2<T>::type* di, typename V2<T>::type* xo,
                                                      typename V2<T>::type* dout, int second, int last)
{
    if (v.sc && v.sc->done) return;
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    mg_smooth_node<W, MODE, T>(a, v, xi, di, xo, dout, second, last, n);
}

struct MgGrid {
    const int* lg;        // fine node -> lattice index (nullptr: identity)
    const int* node_of;   // fine lattice index -> node (nullptr: identity)
    int n1f, n2f, n1c, n2c;
    int co1, co2;         // directions coarsened to the next level
    int per1;             // direction 1 periodic
    // an 8-node fine level above the corner lattice: its transfer tables
    // (nullptr: a lattice level); see k_mg_galerkin_q8
    const int2* q8_par;
    const int* q8_ch;
};

// Coarse right-hand side: b_c(C) = mask_c * sum over C's fine children f of
// w_f * res_f (fine residuals are already zero on fixed DOFs).
// Restriction to coarse lattice point (K1, K2): the fine residuals of its
// children weighted by their interpolation weights.  All nine positions
// around the point are read (those off the lattice at a clamped position,
// with weight 0), so the loads issue together instead of one behind each
// branch; a zero-weight term adds an exact zero, so the sums are those of
// the children alone (they start at +0, and a rounded sum is never -0).
template <typename T>
__device__ __forceinline__ typename V2<T>::type mg_restrict_lat(const MgGrid& gr, const typename V2<T>::type* res,
                                                                uchar2 mc, int K1, int K2)
{
    using V = typename V2<T>::type;
    const int p1 = mg_pos(gr.n1f, gr.co1, K1, gr.per1), p2 = mg_pos(gr.n2f, gr.co2, K2);
    // twice the weights along each direction (integers, so no branches)
    int r1[3], c2[3], e1[3], e2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int f1 = mg_step(gr.n1f, gr.per1, p1, a - 1), f2 = p2 + a - 1;
        r1[a] = f1 < 0 ? p1 : f1;
        c2[a] = (f2 < 0 || f2 >= gr.n2f) ? p2 : f2;
        const int g1 = mg_weight2(gr.n1f, gr.co1, r1[a], K1, gr.per1), g2 = mg_weight2(gr.n2f, gr.co2, c2[a], K2);
        e1[a] = f1 < 0 ? 0 : g1;
        e2[a] = c2[a] != f2 ? 0 : g2;
    }
    V r[3][3];
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2) {
            const int lf = r1[a1] * gr.n2f + c2[a2];
            r[a1][a2] = res[gr.node_of ? gr.node_of[lf] : lf];
        }
    double sx = 0., sy = 0.;
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2) {
            // the product of the two weights, exactly
            const double w = 0.25 * (double)(e1[a1] * e2[a2]);
            sx += w * (double)r[a1][a2].x;
            sy += w * (double)r[a1][a2].y;
        }
    return V{mc.x ? (T)sx : (T)0, mc.y ? (T)sy : (T)0};
}

template <typename T>
__device__ __forceinline__ typename V2<T>::type mg_restrict_value(const MgGrid& gr, const typename V2<T>::type* res,
                                                                  uchar2 mc, int C)
{
    using V = typename V2<T>::type;
    double sx = 0., sy = 0.;
    if (gr.q8_ch) {
        // the corner itself with weight 1, the mid-side nodes of its edges 1/2
#pragma unroll
        for (int j = 0; j < kQ8Children; ++j) {
            const int f = gr.q8_ch[(size_t)C * kQ8Children + j];
            if (f < 0) continue;
            const V r = res[f];
            const double w = j == 0 ? 1. : 0.5;
            sx += w * (double)r.x;
            sy += w * (double)r.y;
        }
        return V{mc.x ? (T)sx : (T)0, mc.y ? (T)sy : (T)0};
    }
    const int K1 = C / gr.n2c;
    return mg_restrict_lat<T>(gr, res, mc, K1, C - K1 * gr.n2c);
}

template <typename T>
__device__ __forceinline__ void mg_restrict_node(const MgGrid& gr, const typename V2<T>::type* res,
                                                 typename V2<T>::type* bc, const uchar2* mask_c, int C)
{
    bc[C] = mg_restrict_value<T>(gr, res, mask_c[C], C);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_restrict(MgGrid gr, const typename V2<T>::type* res,
                                                        typename V2<T>::type* bc, const uchar2* mask_c, int nnc,
                                                        int nbc, const CgScal* sc)
{
    if (sc && sc->done) return;
    const int C = xcd_block(nbc) * kBlock + threadIdx.x;
    if (C >= nnc) return;
    mg_restrict_node<T>(gr, res, bc, mask_c, C);
}

// Prolongation to fine node n at lattice point (f1, f2) of an n1f x n2f lattice.
template <typename T>
__device__ __forceinline__ void mg_prolong_lat(const MgGrid& gr, const typename V2<T>::type* xc,
                                               const typename V2<T>::type* xin, typename V2<T>::type* xout,
                                               const uchar2* mask_f, int n, int f1, int f2)
{
    using V = typename V2<T>::type;
    // a lone parent is read twice, the second time with weight 0 (an exact
    // zero term), so the four loads issue together
    int Q1[2], Q2[2], E1[2], E2[2];
    mg_parents_sel(gr.n1f, gr.co1, f1, gr.per1, Q1[0], Q1[1], E1[0], E1[1]);
    mg_parents_sel(gr.n2f, gr.co2, f2, 0, Q2[0], Q2[1], E2[0], E2[1]);
    V c[2][2];
#pragma unroll
    for (int b1 = 0; b1 < 2; ++b1)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) c[b1][b2] = xc[Q1[b1] * gr.n2c + Q2[b2]];
    double sx = 0., sy = 0.;
#pragma unroll
    for (int b1 = 0; b1 < 2; ++b1)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) {
            // the product of the two weights, exactly
            const double w = 0.25 * (double)(E1[b1] * E2[b2]);
            sx += w * (double)c[b1][b2].x;
            sy += w * (double)c[b1][b2].y;
        }
    const uchar2 mk = mask_f[n];
    const V x = xin[n];
    xout[n] = V{x.x + (mk.x ? (T)sx : (T)0), x.y + (mk.y ? (T)sy : (T)0)};
}

// Coarse-grid correction: x_f = xin_f + mask_f * sum over f's parents of w * xc.
template <typename T>
__device__ __forceinline__ void mg_prolong_node(const MgGrid& gr, const typename V2<T>::type* xc,
                                                const typename V2<T>::type* xin, typename V2<T>::type* xout,
                                                const uchar2* mask_f, int n)
{
    using V = typename V2<T>::type;
    if (gr.q8_par) {
        const int2 pp = gr.q8_par[n];
        const V c1 = xc[pp.x];
        double sx = c1.x, sy = c1.y;
        if (pp.y >= 0) {
            const V c2 = xc[pp.y];
            sx = 0.5 * (double)c1.x + 0.5 * (double)c2.x;
            sy = 0.5 * (double)c1.y + 0.5 * (double)c2.y;
        }
        const uchar2 mk = mask_f[n];
        const V x = xin[n];
        xout[n] = V{x.x + (mk.x ? (T)sx : (T)0), x.y + (mk.y ? (T)sy : (T)0)};
        return;
    }
    const int lf = gr.lg ? gr.lg[n] : n;
    const int f1 = lf / gr.n2f;
    mg_prolong_lat<T>(gr, xc, xin, xout, mask_f, n, f1, lf - f1 * gr.n2f);
}


template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_prolong(MgGrid gr, const typename V2<T>::type* xc,
                                                       const typename V2<T>::type* xin, typename V2<T>::type* xout,
                                                       const uchar2* mask_f, int nnf, int nbf, const CgScal* sc)
{
    if (sc && sc->done) return;
    const int n = xcd_block(nbf) * kBlock + threadIdx.x;
    if (n >= nnf) return;
    mg_prolong_node<T>(gr, xc, xin, xout, mask_f, n);
}

// ------------------------------------------------------------ segment smoother
//
// Where elements are stretched (short edges along lattice direction 1) the
// matrix couples strongly along that direction and a point smoother leaves
// errors that are smooth across the weak direction untouched.  On such
// levels D is the block-tridiagonal part of A along segments of kSeg nodes
// of direction 1 (segment s of lattice column i2: nodes i1 = s kSeg ..
// s kSeg + kSeg - 1), so one application of D^-1 solves kSeg coupled nodes
// at once.  Per node the set-up stores the inverse of the pivot block of
// the block-LU factorisation of its segment,
//   Dt_0 = D_0,  Dt_k = D_k - L_k Dt_{k-1}^-1 L_k^T   (L_k = A(k, k-1)),
// in the minv slot, and the solve D z = rho is
//   y_k = Dt_k^-1 rho_k - (Dt_k^-1 L_k) y_{k-1},   z_k = y_k - (Dt_k^-1 A(k, k+1)) z_{k+1}.
// The segments coloured by (s mod 2, i2 mod 2) are uncoupled within a
// colour, so lambda_max(D^-1 A) <= 4: the Chebyshev interval is
// [4/30, 4.4] with no eigenvalue estimate.
constexpr int kSeg = 16;           // segment length (direction 1) and tile height

// Lane (lane - D) / (lane + D) of the caller's row of 16 lanes (DPP
// row_shr / row_shl): the segment scans keep each segment in one row.  A
// lane with no such source lane keeps its own value (the scans never use it).
template <int D>
__device__ __forceinline__ int row_up_i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x110 | D, 0xf, 0xf, false); }
template <int D>
__device__ __forceinline__ int row_down_i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x100 | D, 0xf, 0xf, false); }
template <int D>
__device__ __forceinline__ float row_up(float v) { return __int_as_float(row_up_i<D>(__float_as_int(v))); }
template <int D>
__device__ __forceinline__ float row_down(float v) { return __int_as_float(row_down_i<D>(__float_as_int(v))); }
template <int D>
__device__ __forceinline__ double row_up(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = row_up_i<D>((int)(b & 0xffffffffLL)), hi = row_up_i<D>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int D>
__device__ __forceinline__ double row_down(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = row_down_i<D>((int)(b & 0xffffffffLL)), hi = row_down_i<D>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One step of each segment scan, D = 1, 2, 4, 8 (k: the lane's place in its
// segment):
//   forward:  (v, M) <- (v + M v', M M') from lane k - D, where k >= D
//   backward: (v, N) <- (v + N v'', N N'') from lane k + D, where k + D < 16
template <int D, typename T>
__device__ __forceinline__ void seg_scan_up(int k, T& v0, T& v1, T& m00, T& m01, T& m10, T& m11)
{
    const T pv0 = row_up<D>(v0), pv1 = row_up<D>(v1);
    const T p00 = row_up<D>(m00), p01 = row_up<D>(m01), p10 = row_up<D>(m10), p11 = row_up<D>(m11);
    if (k >= D) {
        const T n0 = v0 + m00 * pv0 + m01 * pv1, n1 = v1 + m10 * pv0 + m11 * pv1;
        const T a00 = m00 * p00 + m01 * p10, a01 = m00 * p01 + m01 * p11;
        const T a10 = m10 * p00 + m11 * p10, a11 = m10 * p01 + m11 * p11;
        v0 = n0;
        v1 = n1;
        m00 = a00;
        m01 = a01;
        m10 = a10;
        m11 = a11;
    }
}
template <int D, typename T>
__device__ __forceinline__ void seg_scan_down(int k, T& v0, T& v1, T& n00, T& n01, T& n10, T& n11)
{
    const T pw0 = row_down<D>(v0), pw1 = row_down<D>(v1);
    const T p00 = row_down<D>(n00), p01 = row_down<D>(n01), p10 = row_down<D>(n10), p11 = row_down<D>(n11);
    if (k + D < 16) {
        const T w0 = v0 + n00 * pw0 + n01 * pw1, w1 = v1 + n10 * pw0 + n11 * pw1;
        const T a00 = n00 * p00 + n01 * p10, a01 = n00 * p01 + n01 * p11;
        const T a10 = n10 * p00 + n11 * p10, a11 = n10 * p01 + n11 * p11;
        v0 = w0;
        v1 = w1;
        n00 = a00;
        n01 = a01;
        n10 = a10;
        n11 = a11;
    }
}
// Tile width (direction 2) of a smoothing tile: 32 columns (512 threads) on
// the finest level, where the wider rows of consecutive nodes stream the
// matrix better; 16 (256 threads) on the coarse levels, which need the
// workgroups for parallelism.
template <int W> constexpr int seg_cols() { return (W == kLatW || W == kSymW16) ? 16 : 32; }
constexpr double kSegLam = 4.;

struct SegGrid {
    int n1, n2;              // lattice of the level
    int per1;                // direction 1 periodic
    int nt2, ntiles;         // tiles along direction 2, tiles in all
    const int* node_of;      // lattice index -> node (level 0 only; nullptr: identity)
    const signed char* slot_pos;   // level 0: node x 9 lattice offsets -> row position
    MgGrid rg;               // SEG_RX0: the restriction from the finer level
    const void* rres;        // SEG_RX0: the finer level's residual
};

// Row positions of the couplings to the previous and next node of the
// segment, packed into the unused fourth component of the node's minv entry
// (bit pattern only; kSegNone: no such coupling).
constexpr unsigned kSegNone = 0xff;
template <typename T> __device__ __forceinline__ T seg_pack(unsigned b);
template <> __device__ __forceinline__ float seg_pack<float>(unsigned b) { return __uint_as_float(b); }
template <> __device__ __forceinline__ double seg_pack<double>(unsigned b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned seg_unpack(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned seg_unpack(double v) { return (unsigned)__double_as_longlong(v); }

__device__ __forceinline__ int seg_node(const SegGrid& g, int i1, int i2)
{
    const int li = i1 * g.n2 + i2;
    return g.node_of ? g.node_of[li] : li;
}

// Row position of the coupling of node n to its lattice neighbour at
// (d1, d2) (always present for lattice neighbours along direction 1).
__device__ __forceinline__ int seg_pos(const SegGrid& g, int n, int d1, int d2)
{
    const int so = mg_slot(d1, d2);
    return g.slot_pos ? g.slot_pos[(size_t)n * 9 + so] : so;
}

// Block of node n's row coupling it to its neighbour at (d1, 0) along
// direction 1; the node's own block (d1 = 0) is at row position 0.
template <typename T>
__device__ __forceinline__ void seg_block(const CgArgs<T>& a, const SegGrid& g, int n, int d1, double (&m)[4])
{
    const int pos = d1 == 0 ? 0 : seg_pos(g, n, d1, 0);
    if (pos < 0) {   // not coupled in the matrix
        m[0] = m[1] = m[2] = m[3] = 0.;
        return;
    }
    const int2 sl = a.slice[n >> 6];
    load_block_d(a, (size_t)sl.x + (n & 63) + (size_t)pos * kWave, m[0], m[1], m[2], m[3]);
}

// Pivot inverses of every segment (one thread per segment column), and
// lambda = 4 for the Chebyshev coefficients.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_seg_setup(CgArgs<T> a, SegGrid g, typename V4<T>::type* minv,
                                                         double* lam)
{
    // along a periodic direction an odd number of segments needs three
    // colours (the last segment touches the first): lambda_max <= 6
    const int nseg1 = (g.n1 + kSeg - 1) / kSeg;
    if (blockIdx.x == 0 && threadIdx.x == 0) mg_set_lam(lam, (g.per1 && (nseg1 & 1)) ? 1.5 * kSegLam : kSegLam);
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    const int nseg = (g.n1 + kSeg - 1) / kSeg;
    if (idx >= nseg * g.n2) return;
    const int s = idx / g.n2, i2 = idx - s * g.n2;
    double p00 = 0., p01 = 0., p11 = 0.;   // Dt_{k-1}^-1 (zero on fixed DOFs)
    uchar2 mp = make_uchar2(0, 0);
    // rows in chunks of kChunk: the loads of a chunk are issued before its
    // (dependent) eliminations and stores, all of them unconditionally (rows
    // past the lattice's end read its last row, and are not used), so that
    // no branch holds a load back
    constexpr int kChunk = 4;
#pragma unroll
    for (int k0 = 0; k0 < kSeg; k0 += kChunk) {
        double d[kChunk][4], l[kChunk][4];
        uchar2 mk[kChunk];
        int nodes[kChunk];
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            const int i1 = s * kSeg + k0 + c;
            const int nc = seg_node(g, min(i1, g.n1 - 1), i2);
            nodes[c] = i1 < g.n1 ? nc : -1;
            mk[c] = a.mask[nc];
            seg_block(a, g, nc, 0, d[c]);
            l[c][0] = l[c][1] = l[c][2] = l[c][3] = 0.;
            if (k0 + c > 0) {
                if (a.vsym) {
                    // the coupling the smoothing steps use: the symmetric layout's
                    T bl[4], bu[4];
                    sym_seg_couplings<T>(a, nc, true, false, bl, bu);
                    for (int j = 0; j < 4; ++j) l[c][j] = bl[j];
                } else {
                    seg_block(a, g, nc, -1, l[c]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            const int n = nodes[c];
            if (n < 0) break;
            const uchar2 m = mk[c];
            double t00 = d[c][0], t01 = 0.5 * (d[c][1] + d[c][2]), t11 = d[c][3];
            // L masked by this node's (rows) and the previous node's (columns) fixed DOFs
            const double l00 = (m.x && mp.x) ? l[c][0] : 0., l01 = (m.x && mp.y) ? l[c][1] : 0.;
            const double l10 = (m.y && mp.x) ? l[c][2] : 0., l11 = (m.y && mp.y) ? l[c][3] : 0.;
            // L P L^T with P = Dt_{k-1}^-1 symmetric
            const double q00 = l00 * p00 + l01 * p01, q01 = l00 * p01 + l01 * p11;
            const double q10 = l10 * p00 + l11 * p01, q11 = l10 * p01 + l11 * p11;
            t00 -= q00 * l00 + q01 * l01;
            t01 -= q00 * l10 + q01 * l11;
            t11 -= q10 * l10 + q11 * l11;
            double i00 = 0., i01 = 0., i11 = 0.;
            if (m.x && m.y) {
                const double det = t00 * t11 - t01 * t01;
                i00 = t11 / det;
                i01 = -t01 / det;
                i11 = t00 / det;
            } else if (m.x) {
                i00 = 1. / t00;
            } else if (m.y) {
                i11 = 1. / t11;
            }
            // the pad carries the row positions of the in-segment couplings
            const int k = k0 + c;
            const int i1 = s * kSeg + k;
            const int ql = k > 0 ? seg_pos(g, n, -1, 0) : -1;
            const int qu = (k < kSeg - 1 && i1 + 1 < g.n1) ? seg_pos(g, n, 1, 0) : -1;
            const unsigned pl = ql >= 0 ? (unsigned)ql : kSegNone, pu = qu >= 0 ? (unsigned)qu : kSegNone;
            minv[n] = typename V4<T>::type{(T)i00, (T)i01, (T)i11, seg_pack<T>(pl | (pu << 8))};
            p00 = i00;
            p01 = i01;
            p11 = i11;
            mp = m;
        }
    }
}

// Smoothing steps of a segment level (same Chebyshev recurrences as
// mg_smooth_node, D^-1 applied by the segment solve):
//   SEG_X0    xo = c0 D^-1 b                      (x1 of the pre-smoothing)
//   SEG_RX0   b = R res (restriction of the finer level's residual), then SEG_X0
//   SEG_FIRST x = xi                              -> do = a1 x + b1 D^-1 (b - A x)
//   MG_CHEB   x = xi + di                         -> xo = x (or x + d if last), do = a di + c D^-1 (b - A x)
//   MG_POST1  x = xi                              -> do = c0 D^-1 (b - A x)
//   SEG_FIRST_X, MG_POST1_X: as SEG_FIRST, MG_POST1, and also xo = x + do
//   MG_CHEB_X x = xi (combined), d = di at the node -> do = a d + c D^-1 (b - A x), xo = x + do
// (SEG_FIRST_X and MG_CHEB_X skip do where dout is null: no later step reads it)
// (the residual steps MG_RESID, MG_RESID_X need no D^-1 and run in k_mg_smooth).
// One workgroup per tile of kSeg x kSegCols lattice nodes: every thread
// forms its node's residual and the per-node factors, then one thread per
// tile column runs the two short recurrences of its segment in LDS.
enum { SEG_X0 = 4, SEG_FIRST = 5, SEG_RX0 = 6 };

template <int W, int MODE, typename T>
__device__ __forceinline__ void mg_smooth_seg_tile(const CgArgs<T>& a, const MgVec<T>& v,
                                                   const typename V2<T>::type* xi, const typename V2<T>::type* di,
                                                   typename V2<T>::type* xo, typename V2<T>::type* dout, int second,
                                                   int last, const SegGrid& g, int tile)
{
    using V = typename V2<T>::type;
    // per node (row k of the tile, column c at k * kSegPitch + c: padded
    // against LDS bank conflicts in the column-wise scans)
    constexpr int kSegCols = seg_cols<W>(), kSegPitch = kSegCols + 1;
    __shared__ T sy[2][kSeg * kSegPitch];    // Dt^-1 rho, then z
    __shared__ T sf[4][kSeg * kSegPitch];    // Dt^-1 L
    __shared__ T sh[4][kSeg * kSegPitch];    // Dt^-1 U
    const int T1 = tile / g.nt2, T2 = tile - T1 * g.nt2;
    const int t1 = threadIdx.x / kSegCols, t2 = threadIdx.x - t1 * kSegCols;
    const int i1 = T1 * kSeg + t1, i2 = T2 * kSegCols + t2;
    const bool in = i1 < g.n1 && i2 < g.n2;
    const int n = in ? seg_node(g, i1, i2) : 0;
    const ChebCoef<T> c = mg_cheb_of<T>(v.lam);
    T rx = 0, ry = 0;
    V xn = V{0, 0}, dn = V{0, 0};
    T f[4] = {0, 0, 0, 0}, h[4] = {0, 0, 0, 0};
    T ax = 0, ay = 0;
    if (in) {
        const uchar2 mk = a.mask[n];
        V bn;
        if constexpr (MODE == SEG_RX0) {
            bn = mg_restrict_value<T>(g.rg, reinterpret_cast<const V*>(g.rres), mk, n);
            const_cast<V*>(v.b)[n] = bn;
        } else {
            bn = v.b[n];
        }
        const typename V4<T>::type mi = a.minv[n];
        const unsigned pp = seg_unpack(mi.w);
        const int pl = pp & 0xff, pu = (pp >> 8) & 0xff;
        constexpr bool lat = W == kLatW, sym = sym_rows<W>();
        const int2 sl = (lat || sym) ? make_int2(0, 9) : a.slice[n >> 6];
        const size_t base = lat ? lat_base(n) : (size_t)sl.x + (n & 63);
        T bl[4] = {0, 0, 0, 0}, bu[4] = {0, 0, 0, 0};   // couplings to the previous / next segment node
        if constexpr (MODE == SEG_X0 || MODE == SEG_RX0) {
            rx = mk.x ? bn.x : (T)0;
            ry = mk.y ? bn.y : (T)0;
            if constexpr (sym) {
                sym_seg_couplings<T>(a, n, pl != (int)kSegNone, pu != (int)kSegNone, bl, bu);
            } else {
                if (pl != (int)kSegNone) load_block(a, base + (size_t)pl * kWave, bl[0], bl[1], bl[2], bl[3]);
                if (pu != (int)kSegNone) load_block(a, base + (size_t)pu * kWave, bu[0], bu[1], bu[2], bu[3]);
            }
        } else {
            // the row product of mg_smooth_node, keeping the two blocks of
            // the segment couplings as they pass
            auto zfun = [&](int m, T& zx, T& zy) {
                if constexpr (MODE == MG_CHEB) {
                    const V xm = xi[m], dm = di[m];
                    zx = xm.x + dm.x;
                    zy = xm.y + dm.y;
                } else {
                    const V xm = xi[m];
                    zx = xm.x;
                    zy = xm.y;
                }
            };
            T yx = 0, yy = 0;
            if constexpr (sym) {
                row_product_sym<T>(a, n, zfun, yx, yy);
                sym_seg_couplings<T>(a, n, pl != (int)kSegNone, pu != (int)kSegNone, bl, bu);
            } else if constexpr (lat) {
                // coarse level: lattice addressing (the coupling slots are 2 and 7)
                T m[9][4];
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    load_block(a, base + (size_t)k * kWave, m[k][0], m[k][1], m[k][2], m[k][3]);
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    T zx, zy;
                    zfun(lat_neighbour(k, i1, i2, g.n1, g.n2, n, g.per1), zx, zy);
                    yx += m[k][0] * zx + m[k][1] * zy;
                    yy += m[k][2] * zx + m[k][3] * zy;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        bl[j] = k == pl ? m[k][j] : bl[j];
                        bu[j] = k == pu ? m[k][j] : bu[j];
                    }
                }
            } else if constexpr (W > 0) {
                int cidx[W];
#pragma unroll
                for (int k = 0; k < W; ++k) cidx[k] = a.col[base + (size_t)(k < sl.y ? k : 0) * kWave];
                T m[W][4];
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    load_block(a, base + (size_t)(k < sl.y ? k : 0) * kWave, m[k][0], m[k][1], m[k][2], m[k][3]);
                    if (k >= sl.y) m[k][0] = m[k][1] = m[k][2] = m[k][3] = 0;
                }
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    T zx, zy;
                    zfun(cidx[k], zx, zy);
                    yx += m[k][0] * zx + m[k][1] * zy;
                    yy += m[k][2] * zx + m[k][3] * zy;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        bl[j] = k == pl ? m[k][j] : bl[j];
                        bu[j] = k == pu ? m[k][j] : bu[j];
                    }
                }
            } else {
                size_t q = base;
                for (int k = 0; k < sl.y; ++k, q += kWave) {
                    T zx, zy, m00, m01, m10, m11;
                    zfun(a.col[q], zx, zy);
                    load_block(a, q, m00, m01, m10, m11);
                    yx += m00 * zx + m01 * zy;
                    yy += m10 * zx + m11 * zy;
                    if (k == pl) bl[0] = m00, bl[1] = m01, bl[2] = m10, bl[3] = m11;
                    if (k == pu) bu[0] = m00, bu[1] = m01, bu[2] = m10, bu[3] = m11;
                }
            }
            rx = mk.x ? bn.x - yx : (T)0;
            ry = mk.y ? bn.y - yy : (T)0;
            xn = xi[n];
            if constexpr (MODE == MG_CHEB || MODE == MG_CHEB_X) dn = di[n];
        }
        apply_minv<T>(mi, rx, ry, ax, ay);
        // Dt^-1 times the couplings to the segment neighbours.  Rows of
        // fixed DOFs are zero in Dt^-1, so y and z vanish on fixed DOFs and
        // the columns of the neighbours' fixed DOFs need no masking.
        f[0] = mi.x * bl[0] + mi.y * bl[2];
        f[1] = mi.x * bl[1] + mi.y * bl[3];
        f[2] = mi.y * bl[0] + mi.z * bl[2];
        f[3] = mi.y * bl[1] + mi.z * bl[3];
        h[0] = mi.x * bu[0] + mi.y * bu[2];
        h[1] = mi.x * bu[1] + mi.y * bu[3];
        h[2] = mi.y * bu[0] + mi.z * bu[2];
        h[3] = mi.y * bu[1] + mi.z * bu[3];
    }
    const int tid = threadIdx.x;
    const int q1 = t1 * kSegPitch + t2;
    sy[0][q1] = ax;
    sy[1][q1] = ay;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        sf[j][q1] = f[j];
        sh[j][q1] = h[j];
    }
    __syncthreads();
    {
        // the two recurrences of each segment as parallel scans over its 16
        // nodes, held by 16 consecutive lanes (column c, row k):
        //   y_k = a_k - F_k y_{k-1}: prefix compositions of y -> v + M y,
        //   z_k = y_k - H_k z_{k+1}: suffix compositions of z -> w + N z
        const int c = tid / kSeg, k = tid - c * kSeg;
        const int q = k * kSegPitch + c;
        T v0 = sy[0][q], v1 = sy[1][q];
        T m00 = -sf[0][q], m01 = -sf[1][q], m10 = -sf[2][q], m11 = -sf[3][q];
        static_assert(kSeg == 16, "a segment is one row of 16 lanes");
        seg_scan_up<1, T>(k, v0, v1, m00, m01, m10, m11);
        seg_scan_up<2, T>(k, v0, v1, m00, m01, m10, m11);
        seg_scan_up<4, T>(k, v0, v1, m00, m01, m10, m11);
        seg_scan_up<8, T>(k, v0, v1, m00, m01, m10, m11);
        // v = y_k
        T n00 = -sh[0][q], n01 = -sh[1][q], n10 = -sh[2][q], n11 = -sh[3][q];
        seg_scan_down<1, T>(k, v0, v1, n00, n01, n10, n11);
        seg_scan_down<2, T>(k, v0, v1, n00, n01, n10, n11);
        seg_scan_down<4, T>(k, v0, v1, n00, n01, n10, n11);
        seg_scan_down<8, T>(k, v0, v1, n00, n01, n10, n11);
        sy[0][q] = v0;
        sy[1][q] = v1;
    }
    __syncthreads();
    if (!in) return;
    const T gx = sy[0][q1], gy = sy[1][q1];
    if constexpr (MODE == SEG_X0 || MODE == SEG_RX0) {
        xo[n] = V{c.c0 * gx, c.c0 * gy};
    } else if constexpr (MODE == SEG_FIRST || MODE == SEG_FIRST_X) {
        const V d = V{c.a1 * xn.x + c.b1 * gx, c.a1 * xn.y + c.b1 * gy};
        if constexpr (MODE == SEG_FIRST_X) {
            xo[n] = V{xn.x + d.x, xn.y + d.y};
            if (dout) dout[n] = d;   // (no later step of degree 2 reads it)
        } else {
            dout[n] = d;
        }
    } else if constexpr (MODE == MG_POST1 || MODE == MG_POST1_X) {
        const V d = V{c.c0 * gx, c.c0 * gy};
        dout[n] = d;
        if constexpr (MODE == MG_POST1_X) xo[n] = V{xn.x + d.x, xn.y + d.y};
    } else if constexpr (MODE == MG_CHEB_X) {
        const T ca = second ? c.a2 : c.a1, cb = second ? c.b2 : c.b1;
        const V d = V{ca * dn.x + cb * gx, ca * dn.y + cb * gy};
        xo[n] = V{xn.x + d.x, xn.y + d.y};
        if (!last && dout) dout[n] = d;
    } else {
        const T ca = second ? c.a2 : c.a1, cb = second ? c.b2 : c.b1;
        const V x = V{xn.x + dn.x, xn.y + dn.y};
        const V d = V{ca * dn.x + cb * gx, ca * dn.y + cb * gy};
        if (last) {
            xo[n] = V{x.x + d.x, x.y + d.y};
        } else {
            xo[n] = x;
            dout[n] = d;
        }
    }
}

template <int W, int MODE, typename T>
__global__ __launch_bounds__(kSeg * seg_cols<W>()) void k_mg_smooth_seg(CgArgs<T> a, MgVec<T> v, const typename V2<T>::type* xi,
                                                          const typename V2<T>::type* di, typename V2<T>::type* xo,
                                                          typename V2<T>::type* dout, int second, int last, SegGrid g)
{
    if (v.sc && v.sc->done) return;
    mg_smooth_seg_tile<W, MODE, T>(a, v, xi, di, xo, dout, second, last, g, xcd_block(g.ntiles));
}

// Coarsest level: x = D^-1 b with the explicit inverse (bs: b in LDS).
template <typename T>
__device__ __forceinline__ void mg_coarse_rows(const double* Dinv, int nd, const double* bs, typename V2<T>::type* x,
                                               int i0, int stride)
{
    using V = typename V2<T>::type;
    for (int i = i0; i < nd / 2; i += stride) {
        double s0 = 0., s1 = 0.;
        const double* r0 = Dinv + (size_t)(2 * i) * nd;
        const double* r1 = r0 + nd;
        for (int j = 0; j < nd; ++j) {
            s0 += r0[j] * bs[j];
            s1 += r1[j] * bs[j];
        }
        x[i] = V{(T)s0, (T)s1};
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_coarse_solve(const double* Dinv, int nd, const typename V2<T>::type* b,
                                                            typename V2<T>::type* x, const CgScal* sc)
{
    using V = typename V2<T>::type;
    if (sc && sc->done) return;
    __shared__ double bs[kMgDenseMax];
    for (int i = threadIdx.x; i < nd / 2; i += kBlock) {
        const V v = b[i];
        bs[2 * i] = v.x;
        bs[2 * i + 1] = v.y;
    }
    __syncthreads();
    mg_coarse_rows<T>(Dinv, nd, bs, x, threadIdx.x, kBlock);
}

// ------------------------------------------------------------ coarse tail

constexpr int kTailThreads = 1024;   // levels of at most this many nodes run in one workgroup
// LDS of the tail: the six vectors of every tail level (the host keeps
// them within kTailVecLds: at most 1536 nodes in FP64) and, where they
// fit in the rest, LDS copies of the levels' rows, pivot inverses and masks
constexpr int kTailVecLds = 144 * 1024;
constexpr int kTailMaxLevels = 16;
// 160 KB per workgroup on gfx950, less the dense solve's right-hand side
// and the levels' Chebyshev coefficients
template <typename T>
__host__ __device__ constexpr int tail_lds_budget()
{
    return 160 * 1024 - kMgDenseMax * 8 - kTailMaxLevels * (int)sizeof(ChebCoef<T>);
}

template <typename T>
struct TailLevel {
    CgArgs<T> a;       // matrix view
    MgVec<T> v;        // vectors in global memory (v.b: right-hand side)
    MgGrid grid;       // to the next coarser level
    int lds;           // byte offset of the level's b, xa, xb, da, db, res in the tail's LDS
    int mat_lds;       // byte offset of the LDS copy of the rows (slot-major, k nn + n), or -1
    int minv_lds;      // byte offset of the LDS copy of the pivot inverses, or -1
    int mask_lds;      // byte offset of the LDS copy of the mask (every tail level has one)
    int where;         // kTailGlobal, kTailMinvLds or kTailAllLds (from minv_lds, mat_lds)
};

// Where a tail level's rows and pivot inverses are read from (its mask and
// vectors are always in LDS): both from global memory, the pivot inverses
// from LDS and the rows from global memory, or both from LDS.  A step picks
// the case once per level, so every load in it has one address space: LDS
// reads through pointers into the tail's LDS arena and global reads through
// address-space-1 pointers (a generic pointer would compile to flat loads,
// which count against the LDS counter as well and turn every LDS wait into
// a wait for all outstanding loads).
constexpr int kTailGlobal = 0, kTailMinvLds = 1, kTailAllLds = 2;

// (hipcc's host pass parses device code as well; the address space is
// given in the device pass only)
#if defined(__HIP_DEVICE_COMPILE__)
#define CP_GLOBAL_AS __attribute__((address_space(1)))
#else
#define CP_GLOBAL_AS
#endif
template <typename P>
__device__ __forceinline__ const CP_GLOBAL_AS P* as_global(const P* p)
{
    return (const CP_GLOBAL_AS P*)p;
}

// The LDS copies of the tail levels' rows, pivot inverses and masks lie in
// one byte range [lo, lo + 16 n16) of the tail's LDS; `img` holds the same
// bytes in global memory (k_mg_tail_image, after each rebuild of the coarse
// levels), so the tail loads them with one coalesced copy.
struct TailImage {
    const float4* img;
    int n16;
    int lo;
};

// q = n / d for 0 <= n < 2^20 and d > 0 (the tail's lattice indices): a
// float reciprocal estimate within one of q, then corrected.
__device__ __forceinline__ int small_div(int n, int d)
{
    int q = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
    const int r = n - q * d;
    q += r < 0 ? -1 : (r >= d ? 1 : 0);
    return q;
}

// LDS bytes of the copies of one tail level's rows, pivot inverses and mask.
template <typename T>
__host__ __device__ constexpr int tail_rows_bytes(int nn) { return (int)((9 * (size_t)nn * 4 * sizeof(T) + 15) / 16 * 16); }
template <typename T>
__host__ __device__ constexpr int tail_minv_bytes(int nn) { return (int)(((size_t)nn * 4 * sizeof(T) + 15) / 16 * 16); }
__host__ __device__ constexpr int tail_mask_bytes(int nn) { return (int)((2 * (size_t)nn + 15) / 16 * 16); }

// Level l of the tail with its vectors pointed into the LDS arena.
template <typename T>
__device__ __forceinline__ TailLevel<T> tail_level(const TailLevel<T>* lv, int l, char* lds)
{
    using V = typename V2<T>::type;
    TailLevel<T> L = lv[l];
    V* base = reinterpret_cast<V*>(lds + L.lds);
    const int nn = L.a.nn;
    L.v.b = base;
    L.v.xa = base + nn;
    L.v.xb = base + 2 * nn;
    L.v.da = base + 3 * nn;
    L.v.db = base + 4 * nn;
    L.v.res = base + 5 * nn;
    return L;
}

// The grid of a tail level to the next one: tail levels are lattices
// numbered point by point, so the grid's node tables are absent.
__device__ __forceinline__ MgGrid tail_grid(const MgGrid& g)
{
    MgGrid r = g;
    r.lg = nullptr;
    r.node_of = nullptr;
    r.q8_par = nullptr;
    r.q8_ch = nullptr;
    return r;
}

// Neighbour k of node n = i1 n2 + i2 in the slot order of lat_neighbour (n
// itself where there is none), with selects only: the direction-1 step
// wraps where the lattice is periodic in direction 1.
__device__ __forceinline__ int tail_neighbour(int k, int i1, int i2, int n1, int n2, int n, int per1)
{
    const int idx = k == 0 ? 4 : (k <= 4 ? k - 1 : k);
    const int d1 = idx / 3 - 1, d2 = idx % 3 - 1;
    const int q1 = i1 + d1, j2 = i2 + d2;
    const bool out1 = q1 < 0 || q1 >= n1;
    const int j1 = q1 < 0 ? q1 + n1 : (q1 >= n1 ? q1 - n1 : q1);
    return ((!out1 || per1) && j2 >= 0 && j2 < n2) ? j1 * n2 + j2 : n;
}

// One smoothing step of node n of a tail level whose rows and pivot
// inverses lie where WHERE says: the arithmetic of mg_smooth_node with
// lattice addressing, the Chebyshev coefficients given.
template <int MODE, int WHERE, typename T>
__device__ __forceinline__ void tail_smooth_node(const TailLevel<T>& L, const char* lds, const ChebCoef<T>& c,
                                                 const typename V2<T>::type* xi, const typename V2<T>::type* di,
                                                 typename V2<T>::type* xo, typename V2<T>::type* dout, int second,
                                                 int last, int n)
{
    using V = typename V2<T>::type;
    using M = typename V4<T>::type;
    const CgArgs<T>& a = L.a;
    const MgVec<T>& v = L.v;
    const uchar2* mask = reinterpret_cast<const uchar2*>(lds + L.mask_lds);
    auto minv = [&](int m) -> M {
        if constexpr (WHERE == kTailGlobal) return as_global(a.minv)[m];
        else return reinterpret_cast<const M*>(lds + L.minv_lds)[m];
    };
    // block k of node n's row: the LDS copy (slot-major; FP64 as the two
    // planes one after the other) or the lattice-addressed global rows
    auto row = [&](int k, T(&b)[4]) {
        if constexpr (WHERE == kTailAllLds) {
            const size_t q = (size_t)k * a.nn + n;
            if constexpr (std::is_same<T, float>::value) {
                const float4 u = reinterpret_cast<const float4*>(lds + L.mat_lds)[q];
                b[0] = u.x; b[1] = u.y; b[2] = u.z; b[3] = u.w;
            } else {
                const double2* pa = reinterpret_cast<const double2*>(lds + L.mat_lds);
                const double2 u = pa[q], w = pa[q + (size_t)9 * a.nn];
                b[0] = u.x; b[1] = u.y; b[2] = w.x; b[3] = w.y;
            }
        } else {
            const size_t q = lat_base(n) + (size_t)k * kWave;
            if constexpr (std::is_same<T, float>::value) {
                const float4 u = as_global(reinterpret_cast<const float4*>(a.va))[q];
                b[0] = u.x; b[1] = u.y; b[2] = u.z; b[3] = u.w;
            } else {
                const double2 u = as_global(reinterpret_cast<const double2*>(a.va))[q];
                const double2 w = as_global(reinterpret_cast<const double2*>(a.vb))[q];
                b[0] = u.x; b[1] = u.y; b[2] = w.x; b[3] = w.y;
            }
        }
    };
    const T ca = second ? c.a2 : c.a1, cb = second ? c.b2 : c.b1;
    const int i1 = small_div(n, a.lat_n2), i2 = n - i1 * a.lat_n2;
    T yx = 0, yy = 0;
    // FP32: all nine blocks and neighbour values loaded before the first
    // use; FP64: one block at a time, three unrolled (nine in flight would
    // not fit the 128 registers of a 1024-thread workgroup); the sums run
    // over k = 0..8 in order either way
    constexpr bool kF64 = std::is_same<T, double>::value;
    constexpr int G = kF64 ? 1 : 9, kUnroll = kF64 ? 3 : 1;
#pragma unroll kUnroll
    for (int k0 = 0; k0 < 9; k0 += G) {
        T m[G][4];
        T zx[G], zy[G];
#pragma unroll
        for (int g = 0; g < G; ++g) row(k0 + g, m[g]);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int nb = tail_neighbour(k0 + g, i1, i2, a.lat_n1, a.lat_n2, n, a.lat_per1);
            if constexpr (MODE == MG_PRE1) {
                const V bm = v.b[nb];
                apply_minv<T>(minv(nb), bm.x, bm.y, zx[g], zy[g]);
                zx[g] *= c.c0;
                zy[g] *= c.c0;
            } else if constexpr (MODE == MG_POST1) {
                const V xm = xi[nb];
                zx[g] = xm.x;
                zy[g] = xm.y;
            } else {
                const V xm = xi[nb], dm = di[nb];
                zx[g] = xm.x + dm.x;
                zy[g] = xm.y + dm.y;
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            yx += m[g][0] * zx[g] + m[g][1] * zy[g];
            yy += m[g][2] * zx[g] + m[g][3] * zy[g];
        }
    }
    const uchar2 mk = mask[n];
    const V bn = v.b[n];
    const T rx = mk.x ? bn.x - yx : (T)0, ry = mk.y ? bn.y - yy : (T)0;
    const M mi = minv(n);
    T gx, gy;
    apply_minv<T>(mi, rx, ry, gx, gy);
    if constexpr (MODE == MG_PRE1) {
        T x1x, x1y;
        apply_minv<T>(mi, bn.x, bn.y, x1x, x1y);
        x1x *= c.c0;
        x1y *= c.c0;
        xo[n] = V{x1x, x1y};
        dout[n] = V{c.a1 * x1x + c.b1 * gx, c.a1 * x1y + c.b1 * gy};
    } else if constexpr (MODE == MG_POST1) {
        dout[n] = V{c.c0 * gx, c.c0 * gy};
    } else if constexpr (MODE == MG_CHEB) {
        const V xn = xi[n], dn = di[n];
        const V x = V{xn.x + dn.x, xn.y + dn.y};
        const V d = V{ca * dn.x + cb * gx, ca * dn.y + cb * gy};
        if (last) {
            xo[n] = V{x.x + d.x, x.y + d.y};
        } else {
            xo[n] = x;
            dout[n] = d;
        }
    } else {
        const V xn = xi[n], dn = di[n];
        xo[n] = V{xn.x + dn.x, xn.y + dn.y};
        v.res[n] = V{rx, ry};
    }
}

// Every level from the first one of at most kTailThreads nodes down to the
// coarsest and back, in one workgroup: the same per-node work as the level
// kernels, with workgroup barriers instead of launches between the steps.
// Everything a step reads is on the chip: the vectors and mask of every
// level in LDS; each level's rows and pivot inverses in LDS where the host
// found room (coarsest first; one coalesced copy of the tail image), the
// others read from L2 (rows kept in registers instead spilled to scratch at
// the 128 registers a 1024-thread workgroup leaves each lane); each level's
// Chebyshev coefficients formed once per call.  lv[0] is the tail's finest
// level: its right-hand side, restricted by the level above, is read from
// lv[0].v.b and its result is written to lv[0].v.xb.  The coarsest level is
// solved by its explicit inverse, eight threads per row (in FP32 tails the
// threads load their part of the inverse at the start, so its latency
// overlaps the smoothing).
template <typename T>
__device__ __forceinline__ void mg_tail_body(const TailLevel<T>* lv, int nl, const double* Dinv, int nd,
                                             const TailImage im, const MgGrid& rg, const typename V2<T>::type* rres)
{
    using V = typename V2<T>::type;
    __shared__ __align__(16) char lds[tail_lds_budget<T>()];
    __shared__ ChebCoef<T> cf[kTailMaxLevels];
    __shared__ double bs[kMgDenseMax];
    const int tid = threadIdx.x;
    {
        constexpr int kBatch = 8;
        float4* dst = reinterpret_cast<float4*>(lds + im.lo);
        for (int i0 = tid; i0 < im.n16; i0 += kBatch * kTailThreads) {
            // every lane loads and stores (past the end: the last entry, with
            // the value it holds), so no branch separates a load from its
            // store and the batch's loads are all in flight at once
            float4 t[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) t[u] = as_global(im.img)[min(i0 + u * kTailThreads, im.n16 - 1)];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) dst[min(i0 + u * kTailThreads, im.n16 - 1)] = t[u];
        }
        const TailLevel<T> L0 = lv[0];
        V* b0 = reinterpret_cast<V*>(lds + L0.lds);
        if (rres) {
            // the restriction of the level above's residual (rres, grid rg),
            // done here instead of in a launch of its own
            for (int n = tid; n < L0.a.nn; n += kTailThreads)
                b0[n] = mg_restrict_value<T>(rg, rres, as_global(L0.a.mask)[n], n);
        } else {
            for (int n = tid; n < L0.a.nn; n += kTailThreads) b0[n] = as_global(L0.v.b)[n];
        }
        if (tid < nl) cf[tid] = mg_cheb_of<T>(lv[tid].v.lam);
        __syncthreads();
    }
    auto step = [&](const TailLevel<T>& L, int l, auto mode, const V* xi, const V* di_, V* xo, V* dout, int second,
                    int last) {
        constexpr int MODE = decltype(mode)::value;
        if (tid >= L.a.nn) return;
        const ChebCoef<T> c = cf[l];
        if (L.where == kTailAllLds)
            tail_smooth_node<MODE, kTailAllLds, T>(L, lds, c, xi, di_, xo, dout, second, last, tid);
        else if (L.where == kTailMinvLds)
            tail_smooth_node<MODE, kTailMinvLds, T>(L, lds, c, xi, di_, xo, dout, second, last, tid);
        else
            tail_smooth_node<MODE, kTailGlobal, T>(L, lds, c, xi, di_, xo, dout, second, last, tid);
    };
    using Pre1 = std::integral_constant<int, MG_PRE1>;
    using Cheb = std::integral_constant<int, MG_CHEB>;
    using Resid = std::integral_constant<int, MG_RESID>;
    using Post1 = std::integral_constant<int, MG_POST1>;
    // down: level l's pre-smoothing, then the restriction of its residual
    auto smooth_down = [&](int l) {
        const TailLevel<T> L = tail_level<T>(lv, l, lds);
        step(L, l, Pre1(), nullptr, nullptr, L.v.xa, L.v.da, 0, 0);
        __syncthreads();
        step(L, l, Cheb(), L.v.xa, L.v.da, L.v.xb, L.v.db, 1, 0);
        __syncthreads();
        step(L, l, Resid(), L.v.xb, L.v.db, L.v.xa, nullptr, 0, 0);
        __syncthreads();
    };
    auto restrict_down = [&](int l) {
        const TailLevel<T> L = tail_level<T>(lv, l, lds);
        const TailLevel<T> Cn = tail_level<T>(lv, l + 1, lds);
        if (tid < Cn.a.nn) {
            const int K1 = small_div(tid, L.grid.n2c);
            const uchar2 mc = reinterpret_cast<const uchar2*>(lds + Cn.mask_lds)[tid];
            const_cast<V*>(Cn.v.b)[tid] = mg_restrict_lat<T>(tail_grid(L.grid), L.v.res, mc, K1, tid - K1 * L.grid.n2c);
        }
        __syncthreads();
    };
    for (int l = 0; l < nl - 2; ++l) {
        smooth_down(l);
        restrict_down(l);
    }
    // row i = tid / 8 of the inverse, its columns j = tid % 8 + 8 m; FP32
    // tails load their part of it before the last restriction, whose work
    // then covers the loads' latency
    const int di = tid >> 3, dj0 = tid & 7;
    constexpr int kDenseCols = kMgDenseMax / 8;
    double dv[kDenseCols];
    if (nl >= 2) smooth_down(nl - 2);
    if constexpr (std::is_same<T, float>::value) {
#pragma unroll
        for (int m = 0; m < kDenseCols; ++m) {
            const int j = dj0 + 8 * m;
            dv[m] = (di < nd && j < nd) ? as_global(Dinv)[(size_t)di * nd + j] : 0.;
        }
    }
    if (nl >= 2) restrict_down(nl - 2);
    {
        const TailLevel<T> Cn = tail_level<T>(lv, nl - 1, lds);
        for (int i = tid; i < nd / 2; i += kTailThreads) {
            const V b = Cn.v.b[i];
            bs[2 * i] = b.x;
            bs[2 * i + 1] = b.y;
        }
        __syncthreads();
        // the eight lanes of a row sum its columns in a fixed order
        double s = 0.;
        if (di < nd) {
            if constexpr (std::is_same<T, float>::value) {
#pragma unroll
                for (int m = 0; m < kDenseCols; ++m) {
                    const int j = dj0 + 8 * m;
                    if (j < nd) s += dv[m] * bs[j];
                }
            } else {
                for (int j = dj0; j < nd; j += 8) s += as_global(Dinv)[(size_t)di * nd + j] * bs[j];
            }
        }
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        if (di < nd && dj0 == 0) {
            T* xc = reinterpret_cast<T*>(Cn.v.xb);
            xc[di] = (T)s;
        }
        __syncthreads();
    }
    for (int l = nl - 2; l >= 0; --l) {
        const TailLevel<T> L = tail_level<T>(lv, l, lds);
        const TailLevel<T> Cn = tail_level<T>(lv, l + 1, lds);
        if (tid < L.a.nn) {
            const int f1 = small_div(tid, L.grid.n2f);
            mg_prolong_lat<T>(tail_grid(L.grid), Cn.v.xb, L.v.xa, L.v.xb,
                              reinterpret_cast<const uchar2*>(lds + L.mask_lds), tid, f1, tid - f1 * L.grid.n2f);
        }
        __syncthreads();
        step(L, l, Post1(), L.v.xb, nullptr, nullptr, L.v.da, 0, 0);
        __syncthreads();
        step(L, l, Cheb(), L.v.xb, L.v.da, L.v.xa, L.v.db, 0, 0);
        __syncthreads();
        step(L, l, Cheb(), L.v.xa, L.v.db, L.v.xb, nullptr, 1, 1);
        __syncthreads();
    }
    {
        // the result of the first tail level, for the prolongation above
        const TailLevel<T> L0 = lv[0];
        const V* x0 = reinterpret_cast<const V*>(lds + L0.lds) + 2 * L0.a.nn;
        for (int n = tid; n < L0.a.nn; n += kTailThreads) L0.v.xb[n] = x0[n];
    }
}

template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_mg_tail(const TailLevel<T>* lv, int nl, const double* Dinv, int nd,
                                                          const CgScal* sc, TailImage im, MgGrid rg,
                                                          const typename V2<T>::type* rres)
{
    if (sc && sc->done) return;
    mg_tail_body<T>(lv, nl, Dinv, nd, im, rg, rres);
}

// The tail image (TailImage): one workgroup per tail level copies the
// level's rows (slot-major, from its lattice rows), pivot inverses and mask
// to where they lie in the tail's LDS, less lo.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_tail_image(const TailLevel<T>* lv, char* img, int lo)
{
    const TailLevel<T> L = lv[blockIdx.x];
    const int nn = L.a.nn;
    if (L.mat_lds >= 0) {
        for (int i = threadIdx.x; i < 9 * nn; i += kBlock) {
            const int k = i / nn, n = i - k * nn;
            const size_t src = lat_base(n) + (size_t)k * kWave;
            if constexpr (std::is_same<T, float>::value) {
                reinterpret_cast<float4*>(img + L.mat_lds - lo)[i] = reinterpret_cast<const float4*>(L.a.va)[src];
            } else {
                double2* pa = reinterpret_cast<double2*>(img + L.mat_lds - lo);
                pa[i] = reinterpret_cast<const double2*>(L.a.va)[src];
                pa[i + 9 * nn] = reinterpret_cast<const double2*>(L.a.vb)[src];
            }
        }
    }
    if (L.minv_lds >= 0)
        for (int n = threadIdx.x; n < nn; n += kBlock)
            reinterpret_cast<typename V4<T>::type*>(img + L.minv_lds - lo)[n] = L.a.minv[n];
    if (L.mask_lds >= 0)
        for (int n = threadIdx.x; n < nn; n += kBlock) reinterpret_cast<uchar2*>(img + L.mask_lds - lo)[n] = L.a.mask[n];
}

// ------------------------------------------------------------ CG with the V-cycle

// w = mask (A u), partials gamma = (r, u), delta = (w, u), rho = (r, r).
template <int W, typename T>
__device__ __forceinline__ void mg_cg_dots_block(const CgArgs<T>& a, const typename V2<T>::type* u, int lb)
{
    using V = typename V2<T>::type;
    const int n = lb * kBlock + threadIdx.x;
    double v[3] = {0., 0., 0.};
    if (n < a.nn) {
        T yx, yy;
        row_apply<W, T>(a, n,
                        [&](int m, T& zx, T& zy) {
                            const V um = u[m];
                            zx = um.x;
                            zy = um.y;
                        },
                        yx, yy);
        const uchar2 mk = a.mask[n];
        const V wn = V{mk.x ? yx : (T)0, mk.y ? yy : (T)0};
        a.w[0][n] = wn;
        const V rn = a.r[0][n], un = u[n];
        v[0] = (double)rn.x * un.x + (double)rn.y * un.y;
        v[1] = (double)wn.x * un.x + (double)wn.y * un.y;
        v[2] = (double)rn.x * rn.x + (double)rn.y * rn.y;
    }
    block_partial<3>(v, a.partials, lb);
}

template <int W, typename T>
__global__ __launch_bounds__(kBlock) void k_mg_cg_dots(CgArgs<T> a, const typename V2<T>::type* u)
{
    if (a.sc->done) return;
    mg_cg_dots_block<W, T>(a, u, xcd_block(a.nblocks));
}

// p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s (own node only).
template <typename T>
// The updated residual's (r, r) goes to per-block partials (partials_alt),
// so that k_mg_cg_check can stop the solve before its V-cycle.
__device__ __forceinline__ void mg_cg_update_block(const CgArgs<T>& a, const typename V2<T>::type* u, int lb)
{
    using V = typename V2<T>::type;
    const int n = lb * kBlock + threadIdx.x;
    double v[1] = {0.};
    if (n < a.nn) {
        const T alpha = (T)a.sc->alpha, beta = (T)a.sc->beta;
        const V un = u[n], wn = a.w[0][n], pp = a.p[n], sp = a.s[0][n], xn = a.x[n], rn = a.r[0][n];
        const V p = V{un.x + beta * pp.x, un.y + beta * pp.y};
        const V s = V{wn.x + beta * sp.x, wn.y + beta * sp.y};
        a.p[n] = p;
        a.s[0][n] = s;
        a.x[n] = V{xn.x + alpha * p.x, xn.y + alpha * p.y};
        const V r = V{rn.x - alpha * s.x, rn.y - alpha * s.y};
        a.r[0][n] = r;
        v[0] = (double)r.x * r.x + (double)r.y * r.y;
    }
    block_partial<1>(v, a.partials_alt, lb);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_cg_update(CgArgs<T> a, const typename V2<T>::type* u)
{
    if (a.sc->done) return;
    mg_cg_update_block<T>(a, u, xcd_block(a.nblocks));
}

// Stop test of the iteration on the updated residual, before the V-cycle:
// a solve that has converged skips its last preconditioner application
// (the V-cycle, dots and step kernels return at once once done is set).
// The (r, r) summed here in block order equals the one the dots kernel
// would sum, so the stop decision is the one the step would take.
template <typename T>
__device__ __forceinline__ void mg_cg_check_body(const CgArgs<T>& a)
{
    double v[1];
    reduce_partials<1>(a.partials_alt, a.nblocks, v);
    if (threadIdx.x == 0) {
        CgScal* sc = a.sc;
        const double rho = v[0];
        if (rho == rho && !(rho * (1. / sc->bnorm2) > sc->tol2)) {
            sc->iters += 1;
            sc->rho = rho;
            sc->done = 1;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_cg_check(CgArgs<T> a)
{
    if (a.sc->done) return;
    mg_cg_check_body<T>(a);
}

// ------------------------------------------------------------ batched launches
//
// The serving loop runs the CG iterations of several multigrid problems in
// lockstep: each launch below runs one step of the iteration (the CG vector
// update or its stop test, one smoothing step of some level, a restriction,
// a prolongation, the tail, the dots, the scalar step) for every problem of
// the batch that has that step.  Tile t of a launch is {op, blk}: block blk
// of what op's own launch of that step runs alone, with that launch's
// arguments args[op], copied into registers first (as a kernel's own
// arguments are).  Each problem therefore computes what it computes alone,
// block for block, and its partial sums are reduced in the same order: its
// results do not depend on the batch.  (The library is compiled with
// -ffp-contract=on, so that a multiply-add is fused by the expression it
// comes from and not by what the optimiser sees around it: the same
// expression rounds the same way in every kernel that inlines it.)

template <typename T>
struct StepOp {   // smoothing step (point or segment smoother)
    CgArgs<T> a;
    MgVec<T> v;
    const typename V2<T>::type* xi;
    const typename V2<T>::type* di;
    typename V2<T>::type* xo;
    typename V2<T>::type* dout;
    int second, last;
    SegGrid g;
};

template <typename T>
struct GridOp {   // restriction (src = finer residual, dst = coarse b) or prolongation
    MgGrid gr;
    const typename V2<T>::type* src;
    const typename V2<T>::type* xin;
    typename V2<T>::type* dst;
    const uchar2* mask;
    int nn;
    const CgScal* sc;
};

template <typename T>
struct TailOp {
    const TailLevel<T>* lv;
    int nl, nd;
    const double* Dinv;
    const CgScal* sc;
    TailImage im;
    // the level above the tail: its grid and residual, restricted into the
    // tail's right-hand side by the tail itself (rres null: lv[0].v.b holds
    // it already)
    MgGrid rg;
    const typename V2<T>::type* rres;
};

template <typename T>
struct CgOp {
    CgArgs<T> a;
    const typename V2<T>::type* u;
};

template <int W, int MODE, typename T>
__global__ __launch_bounds__(kBlock) void k_b_smooth(const StepOp<T>* ops, const int2* tiles, int ntiles)
{
    const int2 t = tiles[xcd_block(ntiles)];
    const StepOp<T> o = ops[t.x];
    if (o.v.sc && o.v.sc->done) return;
    const int n = t.y * kBlock + threadIdx.x;
    if (n >= o.a.nn) return;
    mg_smooth_node<W, MODE, T>(o.a, o.v, o.xi, o.di, o.xo, o.dout, o.second, o.last, n);
}

template <int W, int MODE, typename T>
__global__ __launch_bounds__(kSeg * seg_cols<W>()) void k_b_smooth_seg(const StepOp<T>* ops, const int2* tiles,
                                                                       int ntiles)
{
    const int2 t = tiles[xcd_block(ntiles)];
    const StepOp<T> o = ops[t.x];
    if (o.v.sc && o.v.sc->done) return;
    mg_smooth_seg_tile<W, MODE, T>(o.a, o.v, o.xi, o.di, o.xo, o.dout, o.second, o.last, o.g, t.y);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_restrict(const GridOp<T>* ops, const int2* tiles, int ntiles)
{
    const int2 t = tiles[xcd_block(ntiles)];
    const GridOp<T> o = ops[t.x];
    if (o.sc && o.sc->done) return;
    const int C = t.y * kBlock + threadIdx.x;
    if (C >= o.nn) return;
    mg_restrict_node<T>(o.gr, o.src, o.dst, o.mask, C);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_prolong(const GridOp<T>* ops, const int2* tiles, int ntiles)
{
    const int2 t = tiles[xcd_block(ntiles)];
    const GridOp<T> o = ops[t.x];
    if (o.sc && o.sc->done) return;
    const int n = t.y * kBlock + threadIdx.x;
    if (n >= o.nn) return;
    mg_prolong_node<T>(o.gr, o.src, o.xin, o.dst, o.mask, n);
}

// One workgroup per problem.
template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_b_tail(const TailOp<T>* ops)
{
    const TailOp<T> o = ops[blockIdx.x];
    if (o.sc && o.sc->done) return;
    mg_tail_body<T>(o.lv, o.nl, o.Dinv, o.nd, o.im, o.rg, o.rres);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_update(const CgOp<T>* ops, const int2* tiles, int ntiles)
{
    const int2 t = tiles[xcd_block(ntiles)];
    const CgOp<T> o = ops[t.x];
    if (o.a.sc->done) return;
    mg_cg_update_block<T>(o.a, o.u, t.y);
}

template <int W, typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_dots(const CgOp<T>* ops, const int2* tiles, int ntiles)
{
    const int2 t = tiles[xcd_block(ntiles)];
    const CgOp<T> o = ops[t.x];
    if (o.a.sc->done) return;
    mg_cg_dots_block<W, T>(o.a, o.u, t.y);
}

// One workgroup per problem.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_check(const CgOp<T>* ops)
{
    const CgOp<T> o = ops[blockIdx.x];
    if (o.a.sc->done) return;
    mg_cg_check_body<T>(o.a);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_step(const CgOp<T>* ops)
{
    const CgOp<T> o = ops[blockIdx.x];
    if (o.a.sc->done) return;
    cg_step_scalars(o.a.sc, o.a.partials, o.a.nblocks);
}

// Zero p and s before the first update (beta = 0 multiplies them).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_cg_clear(CgArgs<T> a)
{
    using V = typename V2<T>::type;
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    a.p[n] = V{(T)0, (T)0};
    a.s[0][n] = V{(T)0, (T)0};
}

}  // namespace cpk
