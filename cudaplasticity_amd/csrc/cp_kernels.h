// cp_kernels.h -- gfx950 device code of the plasticity hot path.
//
// Layout in HBM (one problem):
//   xy[nn]      double2  node coordinates
//   conn[ne]    int4     element nodes, reference local order
//   ecnu[ne]    double2  secant modulus E_c and Poisson ratio nu_c per element
//   matrix      2x2-block sliced ELL: slices of 64 consecutive nodes (one
//               wave each); slice s holds width_s blocks per node stored
//               block-major (offset_s + k*64 + lane), so a wave's loads of
//               block k are contiguous.  col[] = neighbour node, values as
//               two double2 planes (a00,a01 | a10,a11) in FP64 or one float4
//               plane in FP32.  Row order inside a node = reference CSR
//               column order (self first, then first appearance).
//   inc[]       int2     node -> (element, packed slot bytes k0..k3): where
//               block K_e(i, j) lands in the node's row; assembly gathers
//               per row in element order, so sums are deterministic.
//   vectors     one double2 (or float2) per node: DOFs (2n, 2n+1).
//
// Reductions are deterministic: wave64 butterflies, a fixed-order LDS step
// and a "last block" that sums the per-block partials in block order.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace cpk {

constexpr int kBlock = 256;  // 4 waves of 64
constexpr int kWave = 64;

// ---------------------------------------------------------------- helpers

// Workgroups are dispatched round-robin over the 8 XCDs; map them so that
// each XCD (and its L2) works on one contiguous range of nodes.
__device__ __forceinline__ int xcd_block(int nblocks)
{
    const int b = blockIdx.x;
    const int q = nblocks >> 3, rem = nblocks & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < rem ? x : rem) + k;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block-wide sum of K values; thread 0 returns the totals.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* sh)
{
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sh[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = ((sh[k] + sh[K + k]) + sh[2 * K + k]) + sh[3 * K + k];
    __syncthreads();
}

// Cross-block sums without in-kernel synchronisation: each block writes its
// partial with a plain store and a one-block kernel launched next sums the
// partials in block order.  (Agent-scope fences inside a kernel would write
// back and invalidate a whole XCD L2 per block on this multi-XCD part; the
// kernel boundary provides the coherence once per launch instead.)
template <int K>
__device__ __forceinline__ void block_partial(double (&v)[K], double* partials, int lb)
{
    __shared__ double sh[4 * K];
    block_sum<K>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) partials[(size_t)lb * K + k] = v[k];
}

// Fixed-order sum of nblocks partials by one block; thread 0 gets the totals.
// Thread t sums partials t, t + kBlock, ... in that order, eight of them
// loaded before they are added (every lane loads: past the end, the last
// partial again, and adds +0 for it, which leaves the sum's bits as they
// are), so their loads are in flight together.
template <int K>
__device__ __forceinline__ void reduce_partials(const double* partials, int nblocks, double (&v)[K])
{
    __shared__ double sh[4 * K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.;
    constexpr int kBatch = 8;
    for (int j0 = threadIdx.x; j0 < nblocks; j0 += kBatch * kBlock) {
        double t[kBatch][K];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int j = min(j0 + u * kBlock, nblocks - 1);
#pragma unroll
            for (int k = 0; k < K; ++k) t[u][k] = partials[(size_t)j * K + k];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const bool in = j0 + u * kBlock < nblocks;
#pragma unroll
            for (int k = 0; k < K; ++k) v[k] += in ? t[u][k] : 0.;
        }
    }
    block_sum<K>(v, sh);
}

template <typename T> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };
template <typename T> struct V4;
template <> struct V4<double> { using type = double4; };
template <> struct V4<float> { using type = float4; };

// ---------------------------------------------------------------- material

struct MatParams {
    double E, nu, h, sigmaT;
    int law, plane;
    double K_T, A, m, n;
};

// Material.cu:3-23 (linear / power); Swift law solved by Newton for sigma.
// Host and device share this one definition (cp_material_f is the host call).
__host__ __device__ inline double mat_f(const MatParams& M, double eps)
{
    const double epsT = M.sigmaT / M.E;
    if (eps < epsT) return eps * M.E;
    if (M.law == 0) return M.sigmaT + (eps - epsT) * M.K_T;
    if (M.law == 1)
        return M.sigmaT + M.A * pow(eps - epsT + pow(M.E / (M.A * M.m), 1. / (M.m - 1.)), M.m) -
               M.A * pow(M.E / (M.A * M.m), M.m / (M.m - 1.));
    const double t = M.E * eps / M.sigmaT;
    double s = 1.0;
    for (int it = 0; it < 60; ++it) {
        const double w = 1.0 + t - s;
        const double pw = pow(w, M.n);
        const double g = s - pw;
        const double dg = 1.0 + M.n * pw / w;
        double sn = s - g / dg;
        sn = sn < 1.0 ? 1.0 : (sn > t ? t : sn);
        const double d = fabs(sn - s);
        s = sn;
        if (d <= 1e-15 * s) break;
    }
    return M.sigmaT * s;
}

// formC (PlasticitySolver.cu:226-242): returns C00 (= C11), C01, C22.
__device__ __forceinline__ void elastic_c(int plane, double E, double nu, double& c00, double& c01, double& c22)
{
    if (plane == 0) {
        const double coef = E / (1. - nu * nu);
        c00 = coef;
        c01 = nu * coef;
        c22 = (1. - nu) * 0.5 * coef;
    } else {
        const double coef = E * (1. - nu) / ((1. + nu) * (1 - 2. * nu));
        c00 = coef;
        c01 = nu / (1. - nu) * coef;
        c22 = (1. - 2. * nu) * 0.5 / (1. - nu) * coef;
    }
}

// calcB4 (PlasticitySolver.cu:25-64) at Gauss point g = 2i + j,
// (r, s) = (p_i, p_j), p = -+1/sqrt(3): B0 = dN/dx, B1 = dN/dy, detJ.
__device__ __forceinline__ void shape_grad(const double (&x)[4], const double (&y)[4], int g, double (&B0)[4],
                                           double (&B1)[4], double& detJ)
{
    const double P = 0.577350269189626;
    const double r = (g & 2) ? P : -P;
    const double s = (g & 1) ? P : -P;
    const double J0 = 0.25 * ((1. + s) * x[0] - (1. + s) * x[1] - (1. - s) * x[2] + (1. - s) * x[3]);
    const double J2 = 0.25 * ((1. + r) * x[0] + (1. - r) * x[1] - (1. - r) * x[2] - (1. + r) * x[3]);
    const double J1 = 0.25 * ((1. + s) * y[0] - (1. + s) * y[1] - (1. - s) * y[2] + (1. - s) * y[3]);
    const double J3 = 0.25 * ((1. + r) * y[0] + (1. - r) * y[1] - (1. - r) * y[2] - (1. + r) * y[3]);
    detJ = J0 * J3 - J1 * J2;
    const double inv = 1. / detJ;
    const double i0 = J3 * inv, i1 = -J1 * inv, i2 = -J2 * inv, i3 = J0 * inv;
    B0[0] = 0.25 * ((1. + s) * i0 + (1. + r) * i1);
    B1[0] = 0.25 * ((1. + s) * i2 + (1. + r) * i3);
    B0[1] = 0.25 * ((-1. - s) * i0 + (1. - r) * i1);
    B1[1] = 0.25 * ((-1. - s) * i2 + (1. - r) * i3);
    B0[2] = 0.25 * ((-1. + s) * i0 - (1. - r) * i1);
    B1[2] = 0.25 * ((-1. + s) * i2 - (1. - r) * i3);
    B0[3] = 0.25 * ((1. - s) * i0 - (1. + r) * i1);
    B1[3] = 0.25 * ((1. - s) * i2 - (1. + r) * i3);
}

__device__ __forceinline__ double pick4(const double (&v)[4], int i)
{
    return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}

// 8-node (serendipity) quad, 3 x 3 Gauss rule with the reference's
// constants (sqrt(0.6), 5/9 and 8/9 truncated to 15 decimals: calcB8
// PlasticitySolver.cu:70, formKe8 :190, strain8 :1555).  Gauss point
// g = 3i + j sits at (r, s) = (p_i, p_j).
__device__ __forceinline__ double q8_point(int i) { return i == 0 ? -0.774596669241483 : (i == 1 ? 0. : 0.774596669241483); }
__device__ __forceinline__ double q8_weight(int i) { return i == 1 ? 0.888888888888888 : 0.555555555555555; }

// calcB8 (PlasticitySolver.cu:66-134) at (r, s): B0 = dN/dx, B1 = dN/dy of
// the nodes in local order (corners (1,1), (-1,1), (-1,-1), (1,-1), then
// the mid-sides of edges 1-2, 2-3, 3-4, 4-1), and detJ.
__device__ __forceinline__ void q8_grad(const double (&x)[8], const double (&y)[8], double r, double s,
                                        double (&B0)[8], double (&B1)[8], double& detJ)
{
    double dr[8], ds[8];
    // corners: N = (1 + r ra)(1 + s sa)(r ra + s sa - 1) / 4
    dr[0] = 0.25 * (1. + s) * (2. * r + s);       ds[0] = 0.25 * (1. + r) * (r + 2. * s);
    dr[1] = -0.25 * (1. + s) * (-2. * r + s);     ds[1] = 0.25 * (1. - r) * (-r + 2. * s);
    dr[2] = -0.25 * (1. - s) * (-2. * r - s);     ds[2] = -0.25 * (1. - r) * (-r - 2. * s);
    dr[3] = 0.25 * (1. - s) * (2. * r - s);       ds[3] = -0.25 * (1. + r) * (r - 2. * s);
    // mid-sides
    dr[4] = -r * (1. + s);                        ds[4] = 0.5 * (1. - r * r);
    dr[5] = -0.5 * (1. - s * s);                  ds[5] = -s * (1. - r);
    dr[6] = -r * (1. - s);                        ds[6] = -0.5 * (1. - r * r);
    dr[7] = 0.5 * (1. - s * s);                   ds[7] = -s * (1. + r);
    double J0 = 0., J1 = 0., J2 = 0., J3 = 0.;    // x_r, y_r, x_s, y_s
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        J0 += dr[a] * x[a];
        J1 += dr[a] * y[a];
        J2 += ds[a] * x[a];
        J3 += ds[a] * y[a];
    }
    detJ = J0 * J3 - J1 * J2;
    const double inv = 1. / detJ;
    const double i0 = J3 * inv, i1 = -J1 * inv, i2 = -J2 * inv, i3 = J0 * inv;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        B0[a] = i0 * dr[a] + i1 * ds[a];
        B1[a] = i2 * dr[a] + i3 * ds[a];
    }
}

// ---------------------------------------------------------------- assembly

struct AsmArgs {
    const int* incPtr;
    const int2* inc;
    const int4* conn;
    const double2* xy;
    const double2* ecnu;
    const int2* slice;  // {offset in blocks, width}
    double2* va;        // FP64 plane (a00, a01)
    double2* vb;        // FP64 plane (a10, a11)
    float4* vf;         // FP32 values
    int nn, nblocks, plane;
    double h;
};

// Blocks K_e(i, j), j = 0..3, of element e's stiffness (formKe4,
// PlasticitySolver.cu:156-186) for its local node i, before the thickness
// factor: the 2 x 2 Gauss rule over B_i^T C B_j detJ.
__device__ __forceinline__ void element_row_blocks(const AsmArgs& a, const int4& c, const double2& en, int i,
                                                   double (&blk)[4][4])
{
    double x[4], y[4];
    {
        const double2 p0 = a.xy[c.x], p1 = a.xy[c.y], p2 = a.xy[c.z], p3 = a.xy[c.w];
        x[0] = p0.x; y[0] = p0.y; x[1] = p1.x; y[1] = p1.y;
        x[2] = p2.x; y[2] = p2.y; x[3] = p3.x; y[3] = p3.y;
    }
    double c00, c01, c22;
    elastic_c(a.plane, en.x, en.y, c00, c01, c22);
#pragma unroll
    for (int j = 0; j < 4; ++j) blk[j][0] = blk[j][1] = blk[j][2] = blk[j][3] = 0.;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        double B0[4], B1[4], dJ;
        shape_grad(x, y, g, B0, B1, dJ);
        const double bi0 = pick4(B0, i), bi1 = pick4(B1, i);
        // BC = B_i^T C detJ (rows 0 and 1 of B^T C for node i)
        const double t00 = bi0 * c00 * dJ, t01 = bi0 * c01 * dJ, t02 = bi1 * c22 * dJ;
        const double t10 = bi1 * c01 * dJ, t11 = bi1 * c00 * dJ, t12 = bi0 * c22 * dJ;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            blk[j][0] += t00 * B0[j] + t02 * B1[j];
            blk[j][1] += t01 * B1[j] + t02 * B0[j];
            blk[j][2] += t10 * B0[j] + t12 * B1[j];
            blk[j][3] += t11 * B1[j] + t12 * B0[j];
        }
    }
}

// Local index (0..3) of the row's own node among an incidence's packed row
// slots: the one at row position 0 (the row lists the node itself first).
__device__ __forceinline__ void incidence_slots(unsigned slots, int (&sl)[4], int& i)
{
    sl[0] = (int)(slots & 255u);
    sl[1] = (int)((slots >> 8) & 255u);
    sl[2] = (int)((slots >> 16) & 255u);
    sl[3] = (int)(slots >> 24);
    i = sl[0] == 0 ? 0 : (sl[1] == 0 ? 1 : (sl[2] == 0 ? 2 : 3));
}

// Element stiffness gathered per node: thread n computes, for every element
// e containing n (local index i), the 2x2 blocks K_e(i, j), j = 0..3, and
// adds them to the slots of its row; slot order and element order are
// fixed, so the sums are deterministic.  W = compile-time bound on the row
// width.
template <int W, bool F64>
__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs a)
{
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    double acc[W][4];
#pragma unroll
    for (int k = 0; k < W; ++k) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.;

    for (int q = a.incPtr[n]; q < a.incPtr[n + 1]; ++q) {
        const int2 ic = a.inc[q];
        int sl[4], i;
        incidence_slots((unsigned)ic.y, sl, i);
        double blk[4][4];
        element_row_blocks(a, a.conn[ic.x], a.ecnu[ic.x], i, blk);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int k = 0; k < W; ++k) {
                if (k == sl[j]) {
                    acc[k][0] += blk[j][0] * a.h;
                    acc[k][1] += blk[j][1] * a.h;
                    acc[k][2] += blk[j][2] * a.h;
                    acc[k][3] += blk[j][3] * a.h;
                }
            }
        }
    }
    const int2 sl = a.slice[n >> 6];
    const size_t base = (size_t)sl.x + (n & 63);
#pragma unroll
    for (int k = 0; k < W; ++k) {
        if (k < sl.y) {
            const size_t q = base + (size_t)k * kWave;
            if (F64) {
                a.va[q] = make_double2(acc[k][0], acc[k][1]);
                a.vb[q] = make_double2(acc[k][2], acc[k][3]);
            } else {
                a.vf[q] = make_float4((float)acc[k][0], (float)acc[k][1], (float)acc[k][2], (float)acc[k][3]);
            }
        }
    }
}

// ---------------------------------------------------------------- CG

struct CgScal {
    double alpha, beta, gamma, delta;
    double bnorm2, tol2;
    double gamma0;
    double rho;       // (r, r): the stop test uses the unpreconditioned residual
    double rho0;      // (r_0, r_0)
    int iters, done, status, maxit;
    unsigned counter, pad;
};

template <typename T>
struct CgArgs {
    using V = typename V2<T>::type;
    const int* col;
    const int2* slice;
    const void* va;  // FP64: double2 (a00,a01); FP32: float4
    const void* vb;  // FP64: double2 (a10,a11)
    V* x;
    V* p;
    V* r[2];
    V* w[2];
    V* s[2];
    const V* b;
    const uchar2* mask;
    const typename V4<T>::type* minv;  // 2x2 block-Jacobi inverse (m00, m01, m11, 0) per node
    CgScal* sc;              // k_cg_iter_fused: two slots sc[0], sc[1]
    double* partials;
    double* partials_alt;    // second partials buffer of k_cg_iter_fused
    int nn, nblocks;
    // coarse multigrid levels: the lattice of their 9-slot stencil rows,
    // whose addresses follow from the node (0: rows through col / slice)
    int lat_n1 = 0, lat_n2 = 0;
    int lat_per1 = 0;        // 1: the lattice is periodic in direction 1 (rings, plates with a hole)
    // multigrid level 0 of a lattice numbered node by node (FP32): the
    // operator in the symmetric lattice layout (five blocks per node; see
    // row_product_sym), nullptr otherwise
    const void* vsym = nullptr;
};

template <typename T>
__device__ __forceinline__ void apply_minv(const typename V4<T>::type& m, T rx, T ry, T& ux, T& uy)
{
    ux = m.x * rx + m.y * ry;
    uy = m.y * rx + m.z * ry;
}

// Matrix values are read with ordinary (temporal) loads.  Non-temporal
// loads, which mark the lines for early eviction, cost more than they save
// here: a block is read again by a neighbour (the symmetric lattice layout's
// backward couplings), and the multigrid levels' rows, read five times per
// V-cycle, can stay in the Infinity Cache (measured: the mixed ramps 8-9 %
// faster without them, FP64 unchanged).
typedef double nd2 __attribute__((ext_vector_type(2)));
typedef float nf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void load_block(const CgArgs<double>& a, size_t q, double& m00, double& m01, double& m10,
                                           double& m11)
{
    const nd2 u = reinterpret_cast<const nd2*>(a.va)[q];
    const nd2 v = reinterpret_cast<const nd2*>(a.vb)[q];
    m00 = u.x; m01 = u.y; m10 = v.x; m11 = v.y;
}

__device__ __forceinline__ void load_block(const CgArgs<float>& a, size_t q, float& m00, float& m01, float& m10,
                                           float& m11)
{
    const nf4 u = reinterpret_cast<const nf4*>(a.va)[q];
    m00 = u.x; m01 = u.y; m10 = u.z; m11 = u.w;
}

// r_0 = mask (b - A x), p_{-1} = 0, s_{-1} = 0  (cgInitSparseMask, ConjGradCuda.cu:15-37).
// zero_x: x = 0 and the SpMV is skipped (inner solve of the refinement loop).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init1(CgArgs<T> a, int zero_x)
{
    using V = typename V2<T>::type;
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    T yx = 0, yy = 0;
    if (zero_x) {
        a.x[n] = V{0, 0};
    } else {
        const int2 sl = a.slice[n >> 6];
        size_t q = (size_t)sl.x + (n & 63);
        for (int k = 0; k < sl.y; ++k, q += kWave) {
            const V xm = a.x[a.col[q]];
            T m00, m01, m10, m11;
            load_block(a, q, m00, m01, m10, m11);
            yx += m00 * xm.x + m01 * xm.y;
            yy += m10 * xm.x + m11 * xm.y;
        }
    }
    const V bn = a.b[n];
    const uchar2 mk = a.mask[n];
    a.r[0][n] = V{mk.x ? (T)(bn.x - yx) : (T)0, mk.y ? (T)(bn.y - yy) : (T)0};
    a.p[n] = V{0, 0};
    a.s[1][n] = V{0, 0};
}

// Row of w = A z for node n, z at a neighbour m given by zfun(m).  W > 0:
// the row is unrolled to W blocks with every load issued up front (indices
// of padding blocks point at the node itself and their values are zeroed),
// so a wave has all of its row's loads in flight at once.  W == 0: wide rows
// (8-node meshes: 13 or 21 blocks) in chunks of 8 blocks, each chunk's loads
// issued together; blocks past the slice width are zero.
template <int W, typename T, typename ZF>
__device__ __forceinline__ void row_product(const CgArgs<T>& a, int n, const ZF& zfun, T& yx, T& yy)
{
    const int2 sl = a.slice[n >> 6];
    const size_t base = (size_t)sl.x + (n & 63);
    yx = 0;
    yy = 0;
    if constexpr (W > 0) {
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
        }
    } else {
        constexpr int C = 8;
        for (int k0 = 0; k0 < sl.y; k0 += C) {
            int cidx[C];
            T m[C][4];
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const size_t q = base + (size_t)(k0 + j < sl.y ? k0 + j : k0) * kWave;
                cidx[j] = a.col[q];
                load_block(a, q, m[j][0], m[j][1], m[j][2], m[j][3]);
                if (k0 + j >= sl.y) m[j][0] = m[j][1] = m[j][2] = m[j][3] = 0;
            }
#pragma unroll
            for (int j = 0; j < C; ++j) {
                T zx, zy;
                zfun(cidx[j], zx, zy);
                yx += m[j][0] * zx + m[j][1] * zy;
                yy += m[j][2] * zx + m[j][3] * zy;
            }
        }
    }
}

// u_0 = M^-1 r_0 (M = I without preconditioning), w_0 = mask (A u_0);
// gamma_0 = (r_0, u_0), delta_0 = (w_0, u_0), rho_0 = (r_0, r_0) -> alpha_0.
template <int W, bool PRE, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init2(CgArgs<T> a)
{
    using V = typename V2<T>::type;
    const int lb = xcd_block(a.nblocks);
    const int n = lb * kBlock + threadIdx.x;
    double v[3] = {0., 0., 0.};
    if (n < a.nn) {
        const V* r0 = a.r[0];
        T yx, yy;
        row_product<W, T>(a, n,
                          [&](int m, T& zx, T& zy) {
                              const V rm = r0[m];
                              if constexpr (PRE) apply_minv<T>(a.minv[m], rm.x, rm.y, zx, zy);
                              else { zx = rm.x; zy = rm.y; }
                          },
                          yx, yy);
        const uchar2 mk = a.mask[n];
        const V wn = V{mk.x ? yx : (T)0, mk.y ? yy : (T)0};
        a.w[0][n] = wn;
        const V rn = r0[n];
        T ux = rn.x, uy = rn.y;
        if constexpr (PRE) apply_minv<T>(a.minv[n], rn.x, rn.y, ux, uy);
        v[0] = (double)rn.x * ux + (double)rn.y * uy;
        v[1] = (double)wn.x * ux + (double)wn.y * uy;
        v[2] = (double)rn.x * rn.x + (double)rn.y * rn.y;
    }
    block_partial<3>(v, a.partials, lb);
}

// Scalars of the first iteration from the init2 partials.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_start(CgArgs<T> a)
{
    double v[3];
    reduce_partials<3>(a.partials, a.nblocks, v);
    if (threadIdx.x == 0) {
        CgScal* sc = a.sc;
        const double g = v[0], d = v[1], rho = v[2];
        if (sc->bnorm2 == 0.) sc->bnorm2 = rho;
        sc->gamma0 = g;
        sc->gamma = g;
        sc->delta = d;
        sc->rho = rho;
        sc->rho0 = rho;
        sc->beta = 0.;
        sc->alpha = g / d;
        sc->iters = 0;
        sc->status = (g != g || d != d) ? 2 : 0;
        sc->done = (rho == 0. || sc->status != 0) ? 1 : 0;
    }
}

// One Chronopoulos-Gear CG iteration in a single launch:
//   p_i = r_i + beta_i p_{i-1};  s_i = w_i + beta_i s_{i-1}   (own node)
//   x_{i+1} = x_i + alpha_i p_i; r_{i+1} = r_i - alpha_i s_i  (own node)
//   w_{i+1} = mask (A r_{i+1}), with r_{i+1} at the neighbours formed on the
//   fly from r_i, w_i, s_{i-1} (ping-pong buffers, parity = iteration count)
//   gamma_{i+1} = (r, r), delta_{i+1} = (w, r) -> the last block forms
//   beta_{i+1} = gamma_{i+1}/gamma_i, alpha_{i+1} = gamma_{i+1} /
//   (delta_{i+1} - beta_{i+1} gamma_{i+1}/alpha_i) and the stop test
//   gamma/||b||^2 <= tol^2 (ConjGradCuda.cu:203-222).
// w and s carry the Dirichlet mask, so r stays zero on fixed DOFs.
template <int W, bool PRE, typename T>
__device__ __forceinline__ void cg_iter_body(const CgArgs<T>& a, int lb, int it, T alpha, T beta, double* partials)
{
    using V = typename V2<T>::type;
    const int par = it & 1;
    const V* rin = a.r[par];
    const V* win = a.w[par];
    const V* sprev = a.s[par ^ 1];
    V* rout = a.r[par ^ 1];
    V* wout = a.w[par ^ 1];
    V* sout = a.s[par];
    const int n = lb * kBlock + threadIdx.x;
    double v[3] = {0., 0., 0.};
    if (n < a.nn) {
        T yx, yy;
        row_product<W, T>(a, n,
                          [&](int m, T& zx, T& zy) {
                              const V rm = rin[m], wm = win[m], sm = sprev[m];
                              const T qx = rm.x - alpha * (wm.x + beta * sm.x);
                              const T qy = rm.y - alpha * (wm.y + beta * sm.y);
                              if constexpr (PRE) apply_minv<T>(a.minv[m], qx, qy, zx, zy);
                              else { zx = qx; zy = qy; }
                          },
                          yx, yy);
        const V ri = rin[n], wi = win[n], sp = sprev[n], pp = a.p[n], xi = a.x[n];
        T uix = ri.x, uiy = ri.y;
        typename V4<T>::type mn{};
        if constexpr (PRE) {
            mn = a.minv[n];
            apply_minv<T>(mn, ri.x, ri.y, uix, uiy);
        }
        const V si = V{wi.x + beta * sp.x, wi.y + beta * sp.y};
        const V pi = V{uix + beta * pp.x, uiy + beta * pp.y};
        const V rn = V{ri.x - alpha * si.x, ri.y - alpha * si.y};
        a.p[n] = pi;
        sout[n] = si;
        a.x[n] = V{xi.x + alpha * pi.x, xi.y + alpha * pi.y};
        rout[n] = rn;
        const uchar2 mk = a.mask[n];
        const V wn = V{mk.x ? yx : (T)0, mk.y ? yy : (T)0};
        wout[n] = wn;
        T unx = rn.x, uny = rn.y;
        if constexpr (PRE) apply_minv<T>(mn, rn.x, rn.y, unx, uny);
        v[0] = (double)rn.x * unx + (double)rn.y * uny;
        v[1] = (double)wn.x * unx + (double)wn.y * uny;
        v[2] = (double)rn.x * rn.x + (double)rn.y * rn.y;
    }
    block_partial<3>(v, partials, lb);
}

template <int W, bool PRE, typename T>
__device__ __forceinline__ void cg_iter_body(const CgArgs<T>& a, int lb)
{
    const CgScal* sc = a.sc;
    cg_iter_body<W, PRE, T>(a, lb, sc->iters, (T)sc->alpha, (T)sc->beta, a.partials);
}

// Scalars of iteration i+1 from those of iteration i and the iteration's
// totals v = {gamma, delta, rho}: beta, alpha of the Chronopoulos-Gear
// recurrence and the stop test on (r, r) / ||b||^2.
__device__ __forceinline__ void cg_next_scalars(CgScal& s, const double (&v)[3])
{
    const double g = v[0], d = v[1], rho = v[2];
    const int iters = s.iters + 1;
    s.iters = iters;
    if (g != g || d != d || rho != rho) {
        s.status = 2;
        s.done = 1;
        return;
    }
    const double bp = g / s.gamma;
    const double ap = g / (d - bp * g / s.alpha);
    s.gamma = g;
    s.delta = d;
    s.rho = rho;
    s.beta = bp;
    s.alpha = ap;
    if (!(rho * (1. / s.bnorm2) > s.tol2)) {
        s.done = 1;
    } else if (iters >= s.maxit) {
        s.done = 1;
        s.status = 1;
    }
}

__device__ __forceinline__ void cg_step_scalars(CgScal* sc, const double* partials, int nblocks)
{
    double v[3];
    reduce_partials<3>(partials, nblocks, v);
    if (threadIdx.x == 0) {
        CgScal s = *sc;
        cg_next_scalars(s, v);
        *sc = s;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_step(CgArgs<T> a)
{
    if (a.sc->done) return;
    cg_step_scalars(a.sc, a.partials, a.nblocks);
}

template <int W, bool PRE, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_iter(CgArgs<T> a)
{
    if (a.sc->done) return;
    cg_iter_body<W, PRE, T>(a, xcd_block(a.nblocks));
}

// Largest grid for which k_cg_iter_fused pays: every block re-reads all
// 3 x nblocks partials of the previous iteration.
constexpr int kFusedMaxBlocks = 2048;

// Iteration k with the scalar step of iteration k-1 folded in, so that one
// launch per iteration suffices.  Scalars live in two slots sc[0], sc[1]
// and per-block partials in two buffers partials, partials_alt: launch k
// reads slot (k-1)&1 and the partials of launch k-1, forms the scalars of
// iteration k in every block (same fixed-order sum everywhere, so every
// block gets the same bits), block 0 stores them in slot k&1, and the
// iteration writes its partials to buffer k&1.  Launch 0 takes slot 0 from
// k_cg_start.  A converged or stopped slot is copied forward unchanged, so
// slot (L-1)&1 is current after L launches.
// At least 4 waves per SIMD (at most 128 VGPRs): the FP64 form otherwise
// takes 130 and a 784-block grid (the 632x316 plate) no longer fits on the
// chip in one round.
template <int W, bool PRE, typename T>
__global__ __launch_bounds__(kBlock, 4) void k_cg_iter_fused(CgArgs<T> a, int k)
{
    __shared__ double s_ab[2];
    __shared__ int s_state[2];
    double* parts[2] = {a.partials, a.partials_alt};
    if (k == 0) {
        if (a.sc[0].done) return;
        cg_iter_body<W, PRE, T>(a, xcd_block(a.nblocks), 0, (T)a.sc[0].alpha, (T)a.sc[0].beta, parts[0]);
        return;
    }
    const CgScal* prev = a.sc + ((k - 1) & 1);
    CgScal* cur = a.sc + (k & 1);
    if (prev->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *cur = *prev;
        return;
    }
    double v[3];
    reduce_partials<3>(parts[(k - 1) & 1], a.nblocks, v);
    if (threadIdx.x == 0) {
        CgScal s = *prev;
        cg_next_scalars(s, v);
        if (blockIdx.x == 0) *cur = s;
        s_ab[0] = s.alpha;
        s_ab[1] = s.beta;
        s_state[0] = s.done;
        s_state[1] = s.iters;
    }
    __syncthreads();
    if (s_state[0]) return;
    cg_iter_body<W, PRE, T>(a, xcd_block(a.nblocks), s_state[1], (T)s_ab[0], (T)s_ab[1], parts[k & 1]);
}

// Batched form for the serving loop: one launch advances one CG iteration of
// every problem in the table; tiles[t] = {problem, local block}.  Each
// problem keeps its own scalars, partials and counter, so its result is the
// same alone or in a batch.
template <int W, bool PRE, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_iter_batched(const CgArgs<T>* probs, const int2* tiles, int ntiles)
{
    const int2 t = tiles[xcd_block(ntiles)];
    const CgArgs<T> a = probs[t.x];
    if (a.sc->done) return;
    cg_iter_body<W, PRE, T>(a, t.y);
}

// Batched scalar step: block b handles problem b.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_step_batched(const CgArgs<T>* probs, int nprob)
{
    const CgArgs<T> a = probs[blockIdx.x];
    if (a.sc->done) return;
    cg_step_scalars(a.sc, a.partials, a.nblocks);
}

// 2x2 block-Jacobi inverse from each node's own (first) block, restricted
// to the free DOFs so that M^-1 r stays zero on fixed ones.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_block_jacobi(CgArgs<T> a, typename V4<T>::type* minv)
{
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    const int2 sl = a.slice[n >> 6];
    T m00, m01, m10, m11;
    load_block(a, (size_t)sl.x + (n & 63), m00, m01, m10, m11);
    const double d00 = m00, d11 = m11, d01 = 0.5 * ((double)m01 + (double)m10);
    const uchar2 mk = a.mask[n];
    double i00 = 0., i01 = 0., i11 = 0.;
    if (mk.x && mk.y) {
        const double det = d00 * d11 - d01 * d01;
        i00 = d11 / det;
        i01 = -d01 / det;
        i11 = d00 / det;
    } else if (mk.x) {
        i00 = 1. / d00;
    } else if (mk.y) {
        i11 = 1. / d11;
    }
    minv[n] = typename V4<T>::type{(T)i00, (T)i01, (T)i11, (T)0};
}

static __global__ void k_cg_setup(CgScal* sc, double bnorm2, double tol2, int maxit)
{
    sc->bnorm2 = bnorm2;
    sc->tol2 = tol2;
    sc->rho = 0.;
    sc->rho0 = 0.;
    sc->maxit = maxit;
    sc->iters = 0;
    sc->done = 0;
    sc->status = 0;
    sc->counter = 0u;
}

// Standalone SpMV y = A x (bandwidth measurement of the stored format).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv(CgArgs<T> a, const typename V2<T>::type* xin,
                                                 typename V2<T>::type* yout)
{
    using V = typename V2<T>::type;
    const int n = xcd_block(a.nblocks) * kBlock + threadIdx.x;
    if (n >= a.nn) return;
    const int2 sl = a.slice[n >> 6];
    size_t q = (size_t)sl.x + (n & 63);
    T yx = 0, yy = 0;
    for (int k = 0; k < sl.y; ++k, q += kWave) {
        const V xm = xin[a.col[q]];
        T m00, m01, m10, m11;
        load_block(a, q, m00, m01, m10, m11);
        yx += m00 * xm.x + m01 * xm.y;
        yy += m10 * xm.x + m11 * xm.y;
    }
    yout[n] = V{yx, yy};
}

// ---------------------------------------------------------------- FP64 residual (matrix free)

struct RefScal {
    double rnorm2;
    // sum over free DOFs of (|K| |u|)^2, |K| |u| the internal force with
    // every product taken in absolute value: the size of the rounding error
    // of the FP64 residual is about eps64 times its norm (0: not computed)
    double mag2;
    unsigned counter, pad;
};

struct ResArgs {
    const int* incPtr;
    const int2* inc;
    const int4* conn;
    const double2* xy;
    const double2* ecnu;
    const double2* u;
    const double2* b;      // NULL: output the internal force itself
    const uchar2* mask;
    double2* r64;          // out
    float2* r32;           // out (optional): FP32 copy for the inner solve
    RefScal* sc;
    double* partials;
    int nn, nblocks, plane;
    double h;
};

// Element internal force f_i = h sum_gp detJ B_i^T C (B u_e) gathered per
// node (the element kernel applied to u), r = mask (b - f) in FP64.
static __global__ __launch_bounds__(kBlock) void k_residual64(ResArgs a)
{
    const int lb = xcd_block(a.nblocks);
    const int n = lb * kBlock + threadIdx.x;
    double v[2] = {0., 0.};
    if (n < a.nn) {
        double fx = 0., fy = 0., mx = 0., my = 0.;
        for (int q = a.incPtr[n]; q < a.incPtr[n + 1]; ++q) {
            const int2 ic = a.inc[q];
            const int4 c = a.conn[ic.x];
            const double2 en = a.ecnu[ic.x];
            const unsigned slots = (unsigned)ic.y;
            const int i = (slots & 255u) == 0 ? 0 : (((slots >> 8) & 255u) == 0 ? 1 : (((slots >> 16) & 255u) == 0 ? 2 : 3));
            const int cc[4] = {c.x, c.y, c.z, c.w};
            double x[4], y[4], ux[4], uy[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double2 p = a.xy[cc[j]];
                const double2 w = a.u[cc[j]];
                x[j] = p.x; y[j] = p.y; ux[j] = w.x; uy[j] = w.y;
            }
            double c00, c01, c22;
            elastic_c(a.plane, en.x, en.y, c00, c01, c22);
            double ex = 0., ey = 0., ax = 0., ay = 0.;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                double B0[4], B1[4], dJ;
                shape_grad(x, y, g, B0, B1, dJ);
                const double exx = B0[0] * ux[0] + B0[1] * ux[1] + B0[2] * ux[2] + B0[3] * ux[3];
                const double eyy = B1[0] * uy[0] + B1[1] * uy[1] + B1[2] * uy[2] + B1[3] * uy[3];
                const double gxy = B1[0] * ux[0] + B1[1] * ux[1] + B1[2] * ux[2] + B1[3] * ux[3] +
                                   B0[0] * uy[0] + B0[1] * uy[1] + B0[2] * uy[2] + B0[3] * uy[3];
                const double sx = c00 * exx + c01 * eyy, sy = c01 * exx + c00 * eyy, tx = c22 * gxy;
                const double bi0 = pick4(B0, i), bi1 = pick4(B1, i);
                ex += dJ * (bi0 * sx + bi1 * tx);
                ey += dJ * (bi1 * sy + bi0 * tx);
                // the same sums with every product in absolute value
                double qx = 0., qy = 0., qg = 0.;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    qx += fabs(B0[j] * ux[j]);
                    qy += fabs(B1[j] * uy[j]);
                    qg += fabs(B1[j] * ux[j]) + fabs(B0[j] * uy[j]);
                }
                const double px = fabs(c00) * qx + fabs(c01) * qy, py = fabs(c01) * qx + fabs(c00) * qy,
                             pt = fabs(c22) * qg;
                ax += fabs(dJ) * (fabs(bi0) * px + fabs(bi1) * pt);
                ay += fabs(dJ) * (fabs(bi1) * py + fabs(bi0) * pt);
            }
            fx += ex * a.h;
            fy += ey * a.h;
            mx += ax * fabs(a.h);
            my += ay * fabs(a.h);
        }
        double rx = fx, ry = fy;
        if (a.b) {
            const double2 bn = a.b[n];
            const uchar2 mk = a.mask[n];
            rx = mk.x ? bn.x - fx : 0.;
            ry = mk.y ? bn.y - fy : 0.;
            v[0] = rx * rx + ry * ry;
            const double gx = mk.x ? mx + fabs(bn.x) : 0., gy = mk.y ? my + fabs(bn.y) : 0.;
            v[1] = gx * gx + gy * gy;
        }
        a.r64[n] = make_double2(rx, ry);
        if (a.r32) a.r32[n] = make_float2((float)rx, (float)ry);
    }
    if (a.sc) block_partial<2>(v, a.partials, lb);
}

// Totals of the residual partials: (r, r) and, with two partials per block
// (the 4-node residual), the squared magnitude of |K| |u| + |b|.
template <int NV>
__global__ __launch_bounds__(kBlock) void k_residual_total(RefScal* sc, const double* partials, int nblocks)
{
    double v[NV];
    reduce_partials<NV>(partials, nblocks, v);
    if (threadIdx.x == 0) {
        sc->rnorm2 = v[0];
        sc->mag2 = NV > 1 ? v[NV - 1] : 0.;
    }
}

// u += (double) d on every node (the refinement update).
static __global__ __launch_bounds__(kBlock) void k_add_correction(double2* u, const float2* d, int nn, int nblocks)
{
    const int n = xcd_block(nblocks) * kBlock + threadIdx.x;
    if (n >= nn) return;
    const double2 a = u[n];
    const float2 c = d[n];
    u[n] = make_double2(a.x + (double)c.x, a.y + (double)c.y);
}

// ---------------------------------------------------------------- return map

struct RmScal {
    double rel_err, omega, omega_prev, rr_prev;
    int plain, last, converged, status;
    unsigned counter, pad;
};

struct RmArgs {
    const int4* conn;
    const double2* xy;
    const double2* u;
    const double2* ecnu;
    const double* psi;
    double* F;            // per-element fields, field f at F[f*ne + e]
    double* G;            // plain secant update of psi
    double* rcur;         // log G - log psi
    const double* rprev;
    RmScal* sc;
    double* partials;
    MatParams M;
    int ne, nblocks;
    int k, aitken, fixed_outer, max_outer, elastic;
    int solved_tight;     // this iteration's linear solve ran at the final tolerance
    double tol;
};

// Mean element strain: strain4 (PlasticitySolver.cu:1517-1549), the mean of
// B u over the 4 Gauss points, or strain8 (:1551-1588), (1/4) sum_g w_g B_g u
// over the 9 points.  conn holds NPE / 4 int4 per element.
template <int NPE>
__device__ __forceinline__ void element_strain(const int4* conn, const double2* xy, const double2* u, int e,
                                               double& exx, double& eyy, double& gxy)
{
    exx = eyy = gxy = 0.;
    if constexpr (NPE == 4) {
        const int4 c = conn[e];
        const int cc[4] = {c.x, c.y, c.z, c.w};
        double x[4], y[4], ux[4], uy[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double2 p = xy[cc[j]];
            const double2 w = u[cc[j]];
            x[j] = p.x; y[j] = p.y; ux[j] = w.x; uy[j] = w.y;
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            double B0[4], B1[4], dJ;
            shape_grad(x, y, g, B0, B1, dJ);
            exx += ux[0] * B0[0] + ux[1] * B0[1] + ux[2] * B0[2] + ux[3] * B0[3];
            eyy += uy[0] * B1[0] + uy[1] * B1[1] + uy[2] * B1[2] + uy[3] * B1[3];
            gxy += ux[0] * B1[0] + ux[1] * B1[1] + ux[2] * B1[2] + ux[3] * B1[3] + uy[0] * B0[0] + uy[1] * B0[1] +
                   uy[2] * B0[2] + uy[3] * B0[3];
        }
    } else {
        const int4 c0 = conn[2 * e], c1 = conn[2 * e + 1];
        const int cc[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        double x[8], y[8], ux[8], uy[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double2 p = xy[cc[j]];
            const double2 w = u[cc[j]];
            x[j] = p.x; y[j] = p.y; ux[j] = w.x; uy[j] = w.y;
        }
        for (int gi = 0; gi < 3; ++gi)
            for (int gj = 0; gj < 3; ++gj) {
                double B0[8], B1[8], dJ;
                q8_grad(x, y, q8_point(gi), q8_point(gj), B0, B1, dJ);
                double el = 0., yl = 0., gl = 0.;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    el += ux[k] * B0[k];
                    yl += uy[k] * B1[k];
                    gl += ux[k] * B1[k] + uy[k] * B0[k];
                }
                const double w = q8_weight(gi) * q8_weight(gj);
                exx += el * w;
                eyy += yl * w;
                gxy += gl * w;
            }
    }
    exx *= 0.25;
    eyy *= 0.25;
    gxy *= 0.25;
}

// strain4 / strain8 + plastStress (:1590-1610) + exitNorm2 (:1664-1691) for
// one element per thread, always in FP64, followed by the decision of the
// safeguarded Irons-Tuck step.
template <int NPE>
__global__ __launch_bounds__(kBlock) void k_return_map(RmArgs a)
{
    const int lb = xcd_block(a.nblocks);
    const int e = lb * kBlock + threadIdx.x;
    double v[5] = {0., 0., 0., 0., 0.};
    if (e < a.ne) {
        double exx, eyy, gxy;
        element_strain<NPE>(a.conn, a.xy, a.u, e, exx, eyy, gxy);
        const double2 en = a.ecnu[e];
        double c00, c01, c22;
        elastic_c(a.M.plane, en.x, en.y, c00, c01, c22);
        const double sxx = c00 * exx + c01 * eyy;
        const double syy = c01 * exx + c00 * eyy;
        const double tau = c22 * gxy;
        const double sigI = sqrt(0.5 * (sxx * sxx + syy * syy + (sxx - syy) * (sxx - syy) + 6. * tau * tau));
        const double epsI = sigI / en.x;
        const double sigTab = mat_f(a.M, epsI);
        const double G = (sigTab != 0.) ? a.M.E * epsI / sigTab : 1.;
        const size_t ne = (size_t)a.ne;
        a.F[0 * ne + e] = exx;
        a.F[1 * ne + e] = eyy;
        a.F[2 * ne + e] = gxy;
        a.F[3 * ne + e] = epsI;
        a.F[4 * ne + e] = sxx;
        a.F[5 * ne + e] = syy;
        a.F[6 * ne + e] = tau;
        a.F[7 * ne + e] = sigI;
        a.F[8 * ne + e] = sigTab;
        a.G[e] = G;
        const double d = sigI - sigTab;
        v[0] = d * d;
        v[1] = sigTab * sigTab;
        if (a.aitken) {
            const double rc = log(G) - log(a.psi[e]);
            const double rp = a.rprev[e];
            const double qd = rc - rp;
            a.rcur[e] = rc;
            v[2] = rc * rc;
            v[3] = rp * qd;
            v[4] = qd * qd;
        }
    }
    block_partial<5>(v, a.partials, lb);
}

// Exit measure relErr = ||sigma_i - sigma_tab|| / ||sigma_tab|| (exitCondition,
// PlasticitySolver.cu:1722-1735) and the Irons-Tuck decision:
//   plain step if last, no Aitken, first iteration or |r_k| > |r_{k-1}|;
//   omega_k = -omega_{k-1} (r_{k-1}.(r_k - r_{k-1})) / |r_k - r_{k-1}|^2 in [0.05, 100].
static __global__ __launch_bounds__(kBlock) void k_rm_decide(RmArgs a)
{
    double v[5];
    reduce_partials<5>(a.partials, a.nblocks, v);
    if (threadIdx.x == 0) {
        RmScal* sc = a.sc;
        // unloaded state: sigma_i = sigma_tab = 0 everywhere, so 0/0; the
        // reference's loop leaves there (NaN > tol is false), here the level
        // is converged with rel_err = 0.  A NaN in either sum stays NaN.
        const double rel = (v[0] == 0. && v[1] == 0.) ? 0. : sqrt(v[0] / v[1]);
        sc->rel_err = rel;
        const int nan = rel != rel;
        const int last = a.elastic || nan ||
                         (a.fixed_outer > 0 ? (a.k + 1 == a.fixed_outer)
                                            : ((rel <= a.tol && a.solved_tight) || a.k + 1 == a.max_outer));
        const int plain = last || !a.aitken || a.k == 0 || v[2] > sc->rr_prev;
        double om = 1.;
        if (!plain) {
            om = v[4] > 0. ? -sc->omega_prev * v[3] / v[4] : sc->omega_prev;
            om = om < 0.05 ? 0.05 : (om > 100.0 ? 100.0 : om);
        }
        sc->omega = om;
        sc->omega_prev = om;
        sc->rr_prev = v[2];
        sc->plain = plain;
        sc->last = last;
        sc->converged = last && !nan && (a.elastic || a.fixed_outer > 0 || rel <= a.tol);
        sc->status = nan ? 2 : 0;
    }
}

// psi_{k+1} = G (plain) or exp(log psi + omega r) (Aitken), then
// E_c = E/psi, nu_c = (1 - (1 - 2 nu)/psi)/2 (PlasticitySolver.cu:1607-1609).
static __global__ __launch_bounds__(kBlock) void k_apply_update(double* psi, double2* ecnu, const double* G, const double* rcur,
                                                         double* rprev, const RmScal* sc, double E, double nu, int ne,
                                                         int nblocks, int aitken)
{
    const int e = xcd_block(nblocks) * kBlock + threadIdx.x;
    if (e >= ne) return;
    double p = G[e];
    if (!sc->plain) {
        p = exp(log(psi[e]) + sc->omega * rcur[e]);
        if (p < 1.) p = 1.;
    }
    psi[e] = p;
    ecnu[e] = make_double2(E / p, 0.5 * (1. - (1. - 2. * nu) / p));
    if (aitken) rprev[e] = rcur[e];
}

// plastDeform (PlasticitySolver.cu:1641-1653).
static __global__ __launch_bounds__(kBlock) void k_plastic_strains(double* F, double* psi, const double2* ecnu, double E, double nu,
                                                            int ne, int nblocks)
{
    const int e = xcd_block(nblocks) * kBlock + threadIdx.x;
    if (e >= ne) return;
    const size_t N = (size_t)ne;
    const double Gm = E * 0.5 / (1. + nu);
    const double Ec = ecnu[e].x, nuc = ecnu[e].y;
    const double ex = (1. / Ec - 1. / E) * F[4 * N + e];
    const double ey = (1. / Ec - 1. / E) * F[5 * N + e];
    const double Gc = Ec * 0.5 / (1. + nuc);
    const double gp = (1. / Gc - 1. / Gm) * F[6 * N + e];
    F[12 * N + e] = ex;
    F[13 * N + e] = ey;
    F[14 * N + e] = gp;
    F[15 * N + e] = sqrt(2.) / 3. * sqrt((ex - ey) * (ex - ey) + 6. * gp * gp);
    psi[e] = E / Ec;
}

// ---------------------------------------------------------------- small element-wise kernels

static __global__ void k_init_state(double* psi, double2* ecnu, double* rprev, double E, double nu, int ne)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    psi[e] = 1.;
    ecnu[e] = make_double2(E, nu);
    rprev[e] = 0.;
}

// b = factor * rp0; u = factor * u0 on fixed DOFs (free DOFs keep u, or 0 when reset).
static __global__ void k_set_load(double2* b, double2* u, const double2* rp0, const double2* u0, const uchar2* mask,
                           double factor, int reset, int nn)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const double2 r = rp0[n], p = u0[n];
    const uchar2 mk = mask[n];
    b[n] = make_double2(factor * r.x, factor * r.y);
    double2 w = reset ? make_double2(0., 0.) : u[n];
    if (!mk.x) w.x = factor * p.x;
    if (!mk.y) w.y = factor * p.y;
    u[n] = w;
}

// Load-step predictor: u <- u + ratio (u - u_prev), log psi likewise (psi >= 1);
// the previous state moves to *_prev.
static __global__ void k_predict_u(double2* u, double2* u_prev, double ratio, int nn)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const double2 a = u[n], b = u_prev[n];
    u_prev[n] = a;
    u[n] = make_double2(a.x + ratio * (a.x - b.x), a.y + ratio * (a.y - b.y));
}

static __global__ void k_predict_psi(double* psi, double* psi_prev, double2* ecnu, double ratio, double E, double nu,
                                     int ne)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const double a = psi[e], b = psi_prev[e];
    psi_prev[e] = a;
    double p = exp(log(a) + ratio * (log(a) - log(b)));
    if (!(p >= 1.)) p = 1.;
    psi[e] = p;
    ecnu[e] = make_double2(E / p, 0.5 * (1. - (1. - 2. * nu) / p));
}

static __global__ void k_rm_reset(RmScal* sc)
{
    sc->omega_prev = 1.;
    sc->rr_prev = 0.;
    sc->counter = 0u;
}

}  // namespace cpk
