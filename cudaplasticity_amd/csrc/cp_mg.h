// cp_mg.h -- gfx950 device code of the geometric multigrid preconditioner.
//
// Structured quad meshes carry a logical grid: node = node_of[i1 * n2 + i2]
// on an n1 x n2 lattice, every matrix coupling within one lattice step.  Each
// coarser level halves both lattice directions (coarse k sits on fine
// min(2k, n-1); a direction of 2 or fewer points is kept) with bilinear
// interpolation P, and carries the Galerkin operator P^T A P of the level
// above, formed on the device as a 3x3-node stencil of 2x2 blocks in the same
// sliced-ELL layout as the fine matrix (slot 0 the node itself, slots 1..8
// the neighbours in raster order; slots outside the lattice point at the
// node with zero values).  Fixed DOFs are removed with the mask: P is zero
// on them, and a coarse DOF that only interpolates to fixed DOFs gets a unit
// diagonal and mask 0.
//
// Smoothing: Chebyshev iteration of degree 3 on the block-Jacobi scaled
// operator over [lam/30, 1.1 lam], lam = the Gershgorin bound of
// |D^-1 A| (no eigenvalue iteration).  Pre- and post-smoothing are the same
// polynomial, so the V-cycle is a symmetric operator and can precondition CG.
// The coarsest level (at most kMgDenseMax DOFs) is solved with an explicit
// inverse formed by Gauss-Jordan elimination in one workgroup.
//
// All sums run in a fixed order: results do not depend on scheduling.
#pragma once

#include "cp_kernels.h"

namespace cpk {

constexpr int kMgDenseMax = 256;   // DOFs of the coarsest level
constexpr double kMgLowFrac = 1. / 30.;

// 1-D coarsening: coarse point count of n fine points.
__host__ __device__ inline int mg_coarse_n(int n) { return n <= 2 ? n : n / 2 + 1; }

// Coarse parents of fine point f of an n-point direction: count (1 or 2),
// coarse indices and bilinear weights.
__host__ __device__ inline int mg_parents(int n, int f, int (&K)[2], double (&w)[2])
{
    if (n <= 2 || (f & 1) == 0) {
        K[0] = n <= 2 ? f : (f >> 1);
        w[0] = 1.;
        return 1;
    }
    if (f == n - 1) {
        K[0] = n / 2;
        w[0] = 1.;
        return 1;
    }
    K[0] = f >> 1;
    K[1] = (f >> 1) + 1;
    w[0] = w[1] = 0.5;
    return 2;
}

// Weight of coarse point K in the interpolation to fine point f (0 if none).
__device__ __forceinline__ double mg_weight(int n, int f, int K)
{
    int P[2];
    double w[2];
    const int np = mg_parents(n, f, P, w);
    double r = 0.;
    for (int p = 0; p < np; ++p)
        if (P[p] == K) r = w[p];
    return r;
}

// Fine position of coarse point K.
__device__ __forceinline__ int mg_pos(int n, int K) { return n <= 2 ? K : (2 * K < n - 1 ? 2 * K : n - 1); }

// Stencil slot of the coarse neighbour at offset (d1, d2), |d| <= 1.
__host__ __device__ inline int mg_slot(int d1, int d2)
{
    const int idx = (d1 + 1) * 3 + (d2 + 1);
    return idx < 4 ? idx + 1 : (idx == 4 ? 0 : idx);
}

template <typename T> struct ChebCoef { T c0, a1, b1, a2, b2; };

// Degree-3 Chebyshev coefficients on [lam/30, 1.1 lam].
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
    int n1f, n2f;
    int n1c, n2c, nnc, nbc;
    const int2* slice_c;
    double2* va_c;         // FP64 destination planes
    double2* vb_c;
    float4* vf_c;          // FP32 destination
    uchar2* mask_c;
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

// One coarse node per thread: (P^T A P)(C, D) for the 3x3 coarse neighbours
// D, summed over C's fine children f (in lattice order) and f's row blocks
// (in storage order).
template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_mg_galerkin(GalArgs<TS, TD> g)
{
    const int C = xcd_block(g.nbc) * kBlock + threadIdx.x;
    if (C >= g.nnc) return;
    const int K1 = C / g.n2c, K2 = C - K1 * g.n2c;
    double acc[9][4];
#pragma unroll
    for (int s = 0; s < 9; ++s) acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.;
    const int p1 = mg_pos(g.n1f, K1), p2 = mg_pos(g.n2f, K2);
    for (int a1 = -1; a1 <= 1; ++a1) {
        const int f1 = p1 + a1;
        if (f1 < 0 || f1 >= g.n1f) continue;
        const double w1 = mg_weight(g.n1f, f1, K1);
        if (w1 == 0.) continue;
        for (int a2 = -1; a2 <= 1; ++a2) {
            const int f2 = p2 + a2;
            if (f2 < 0 || f2 >= g.n2f) continue;
            const double w2 = mg_weight(g.n2f, f2, K2);
            if (w2 == 0.) continue;
            const int lf = f1 * g.n2f + f2;
            const int fn = g.node_of ? g.node_of[lf] : lf;
            const double wf = w1 * w2;
            const uchar2 mf = g.f.mask[fn];
            const int2 sl = g.f.slice[fn >> 6];
            const size_t base = (size_t)sl.x + (fn & 63);
            for (int k = 0; k < sl.y; ++k) {
                const size_t q = base + (size_t)k * kWave;
                const int gn = g.f.col[q];
                double m00, m01, m10, m11;
                load_block_d(g.f, q, m00, m01, m10, m11);
                const uchar2 mg = g.f.mask[gn];
                m00 = (mf.x && mg.x) ? m00 : 0.;
                m01 = (mf.x && mg.y) ? m01 : 0.;
                m10 = (mf.y && mg.x) ? m10 : 0.;
                m11 = (mf.y && mg.y) ? m11 : 0.;
                const int lgg = g.lg ? g.lg[gn] : gn;
                const int g1 = lgg / g.n2f, g2 = lgg - g1 * g.n2f;
                int Q1[2], Q2[2];
                double V1[2], V2[2];
                const int nq1 = mg_parents(g.n1f, g1, Q1, V1), nq2 = mg_parents(g.n2f, g2, Q2, V2);
                for (int b1 = 0; b1 < nq1; ++b1)
                    for (int b2 = 0; b2 < nq2; ++b2) {
                        const int d1 = Q1[b1] - K1, d2 = Q2[b2] - K2;
                        if (d1 < -1 || d1 > 1 || d2 < -1 || d2 > 1) continue;
                        const int s = mg_slot(d1, d2);
                        const double ww = wf * V1[b1] * V2[b2];
#pragma unroll
                        for (int t = 0; t < 9; ++t)
                            if (t == s) {
                                acc[t][0] += ww * m00;
                                acc[t][1] += ww * m01;
                                acc[t][2] += ww * m10;
                                acc[t][3] += ww * m11;
                            }
                    }
            }
        }
    }
    uchar2 mc;
    mc.x = acc[0][0] > 0. ? 1 : 0;
    mc.y = acc[0][3] > 0. ? 1 : 0;
    if (!mc.x) acc[0][0] = 1.;
    if (!mc.y) acc[0][3] = 1.;
    g.mask_c[C] = mc;
    const int2 sl = g.slice_c[C >> 6];
    const size_t base = (size_t)sl.x + (C & 63);
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const size_t q = base + (size_t)t * kWave;
        if (g.va_c) {
            g.va_c[q] = make_double2(acc[t][0], acc[t][1]);
            g.vb_c[q] = make_double2(acc[t][2], acc[t][3]);
        } else {
            g.vf_c[q] = make_float4((float)acc[t][0], (float)acc[t][1], (float)acc[t][2], (float)acc[t][3]);
        }
    }
}

// Block-Jacobi inverse (as k_block_jacobi) and the per-block maximum of the
// Gershgorin row sums of |D^-1 A|.
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
    if (threadIdx.x == 0) *lam = v > 0. ? v : 1.;
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

// In-place inverse of the SPD matrix D (nd x nd) by Gauss-Jordan
// elimination without pivoting, one workgroup.
// Launch with 1024 threads: thread t owns column t & 255 of rows t >> 8,
// (t >> 8) + 4, ...
static __global__ __launch_bounds__(1024) void k_mg_invert(double* D, int nd)
{
    __shared__ double colk[kMgDenseMax];
    __shared__ double piv;
    const int j = threadIdx.x & 255, i0 = threadIdx.x >> 8;
    for (int k = 0; k < nd; ++k) {
        if (threadIdx.x == 0) {
            const double d = D[(size_t)k * nd + k];
            piv = d != 0. ? 1. / d : 0.;
        }
        for (int i = threadIdx.x; i < nd; i += blockDim.x) colk[i] = D[(size_t)i * nd + k];
        __syncthreads();
        const double p = piv;
        // every row but k eliminated against row k (row k and column k are
        // read here and rewritten after the barrier)
        if (j < nd && j != k) {
            const double rkj = p * D[(size_t)k * nd + j];
            for (int i = i0; i < nd; i += 4)
                if (i != k) D[(size_t)i * nd + j] -= colk[i] * rkj;
        }
        __syncthreads();
        for (int jj = threadIdx.x; jj < nd; jj += blockDim.x) {
            if (jj == k) continue;
            D[(size_t)k * nd + jj] *= p;
        }
        for (int i = threadIdx.x; i < nd; i += blockDim.x) {
            if (i == k) continue;
            D[(size_t)i * nd + k] = -colk[i] * p;
        }
        if (threadIdx.x == 0) D[(size_t)k * nd + k] = p;
        __syncthreads();
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

// One smoothing step of a level.  The iterate at the neighbours is formed
// on the fly from the previous step's buffers, so each step is one launch:
//   MG_PRE1  x1 = c0 D^-1 b                    -> xo = x1, do = a1 x1 + b1 D^-1 (b - A x1)
//   MG_CHEB  x = xi + di                       -> xo = x (or x + d if last), do = a di + c D^-1 (b - A x)
//   MG_RESID x = xi + di                       -> xo = x, res = b - A x
//   MG_POST1 x = xi                            -> do = c0 D^-1 (b - A x)
template <int W, int MODE, typename T>
__global__ __launch_bounds__(kBlock) void k_mg_smooth(CgArgs<T> a, MgVec<T> v, const typename V2<T>::type* xi,
                                                      const typename V2<T>::type* di, typename V2<T>::type* xo,
                                                      typename V2<T>::type* dout, int second, int last)
{
    using V = typename V2<T>::type;
    if (v.sc && v.sc->done) return;
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    const ChebCoef<T> c = mg_cheb<T>(*v.lam);
    const T ca = second ? c.a2 : c.a1, cb = second ? c.b2 : c.b1;
    T yx, yy;
    row_product<W, T>(a, n,
                      [&](int m, T& zx, T& zy) {
                          if constexpr (MODE == MG_PRE1) {
                              const V bm = v.b[m];
                              apply_minv<T>(a.minv[m], bm.x, bm.y, zx, zy);
                              zx *= c.c0;
                              zy *= c.c0;
                          } else if constexpr (MODE == MG_POST1) {
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

struct MgGrid {
    const int* lg;        // fine node -> lattice index (nullptr: identity)
    const int* node_of;   // fine lattice index -> node (nullptr: identity)
    int n1f, n2f, n1c, n2c;
};

// Coarse right-hand side: b_c(C) = mask_c * sum over C's fine children f of
// w_f * res_f (fine residuals are already zero on fixed DOFs).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_restrict(MgGrid gr, const typename V2<T>::type* res,
                                                        typename V2<T>::type* bc, const uchar2* mask_c, int nnc,
                                                        int nbc, const CgScal* sc)
{
    using V = typename V2<T>::type;
    if (sc && sc->done) return;
    const int C = xcd_block(nbc) * kBlock + threadIdx.x;
    if (C >= nnc) return;
    const int K1 = C / gr.n2c, K2 = C - K1 * gr.n2c;
    const int p1 = mg_pos(gr.n1f, K1), p2 = mg_pos(gr.n2f, K2);
    double sx = 0., sy = 0.;
    for (int a1 = -1; a1 <= 1; ++a1) {
        const int f1 = p1 + a1;
        if (f1 < 0 || f1 >= gr.n1f) continue;
        const double w1 = mg_weight(gr.n1f, f1, K1);
        if (w1 == 0.) continue;
        for (int a2 = -1; a2 <= 1; ++a2) {
            const int f2 = p2 + a2;
            if (f2 < 0 || f2 >= gr.n2f) continue;
            const double w2 = mg_weight(gr.n2f, f2, K2);
            if (w2 == 0.) continue;
            const int lf = f1 * gr.n2f + f2;
            const V r = res[gr.node_of ? gr.node_of[lf] : lf];
            sx += w1 * w2 * (double)r.x;
            sy += w1 * w2 * (double)r.y;
        }
    }
    const uchar2 mc = mask_c[C];
    bc[C] = V{mc.x ? (T)sx : (T)0, mc.y ? (T)sy : (T)0};
}

// Coarse-grid correction: x_f = xin_f + mask_f * sum over f's parents of w * xc.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_prolong(MgGrid gr, const typename V2<T>::type* xc,
                                                       const typename V2<T>::type* xin, typename V2<T>::type* xout,
                                                       const uchar2* mask_f, int nnf, int nbf, const CgScal* sc)
{
    using V = typename V2<T>::type;
    if (sc && sc->done) return;
    const int n = xcd_block(nbf) * kBlock + threadIdx.x;
    if (n >= nnf) return;
    const int lf = gr.lg ? gr.lg[n] : n;
    const int f1 = lf / gr.n2f, f2 = lf - f1 * gr.n2f;
    int Q1[2], Q2[2];
    double W1[2], W2[2];
    const int nq1 = mg_parents(gr.n1f, f1, Q1, W1), nq2 = mg_parents(gr.n2f, f2, Q2, W2);
    double sx = 0., sy = 0.;
    for (int b1 = 0; b1 < nq1; ++b1)
        for (int b2 = 0; b2 < nq2; ++b2) {
            const V c = xc[Q1[b1] * gr.n2c + Q2[b2]];
            sx += W1[b1] * W2[b2] * (double)c.x;
            sy += W1[b1] * W2[b2] * (double)c.y;
        }
    const uchar2 mk = mask_f[n];
    const V x = xin[n];
    xout[n] = V{x.x + (mk.x ? (T)sx : (T)0), x.y + (mk.y ? (T)sy : (T)0)};
}

// Coarsest level: x = D^-1 b with the explicit inverse, one workgroup.
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
    for (int i = threadIdx.x; i < nd / 2; i += kBlock) {
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

// ------------------------------------------------------------ CG with the V-cycle

// w = mask (A u), partials gamma = (r, u), delta = (w, u), rho = (r, r).
template <int W, typename T>
__global__ __launch_bounds__(kBlock) void k_mg_cg_dots(CgArgs<T> a, const typename V2<T>::type* u)
{
    using V = typename V2<T>::type;
    if (a.sc->done) return;
    const int lb = xcd_block(a.nblocks);
    const int n = lb * kBlock + threadIdx.x;
    double v[3] = {0., 0., 0.};
    if (n < a.nn) {
        T yx, yy;
        row_product<W, T>(a, n,
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

// p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s (own node only).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_cg_update(CgArgs<T> a, const typename V2<T>::type* u)
{
    using V = typename V2<T>::type;
    if (a.sc->done) return;
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    const T alpha = (T)a.sc->alpha, beta = (T)a.sc->beta;
    const V un = u[n], wn = a.w[0][n], pp = a.p[n], sp = a.s[0][n], xn = a.x[n], rn = a.r[0][n];
    const V p = V{un.x + beta * pp.x, un.y + beta * pp.y};
    const V s = V{wn.x + beta * sp.x, wn.y + beta * sp.y};
    a.p[n] = p;
    a.s[0][n] = s;
    a.x[n] = V{xn.x + alpha * p.x, xn.y + alpha * p.y};
    a.r[0][n] = V{rn.x - alpha * s.x, rn.y - alpha * s.y};
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
