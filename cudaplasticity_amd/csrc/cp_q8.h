// cp_q8.h -- gfx950 device code of the 8-node (order-2) element path.
//
// The 8-node quad couples a corner node to 20 others, so the node-centric
// assembly of the 4-node path (every row's blocks in registers) would need
// 84 FP64 accumulators per thread.  Here the work is split in two passes:
//
//   k_q8_ke      one thread per (element e, local node i): the 8 blocks
//                K_e(i, j) of formKe8 (PlasticitySolver.cu:188-224) over the
//                3 x 3 rule, in FP64, to keb[64 e + 8 i + j] (2 KB per element);
//   k_q8_gather  one thread per stored matrix block: the sum of its element
//                blocks in element order (a precomputed list), written once
//                into the sliced-ELL planes (FP64 pair or FP32 float4).
//
// The FP64 residual r = mask (b - f(u)) of the refinement loop follows the
// same split: k_q8_force writes the element forces f_e(i), k_q8_fgather sums
// them per node over the node's incidences in element order.  All sums have
// a fixed order, so results are deterministic.
#pragma once

#include "cp_kernels.h"

namespace cpk {

struct Q8Args {
    const int4* conn;      // 2 int4 per element, reference local order
    const double2* xy;
    const double2* ecnu;
    double4* keb;          // element blocks (a00, a01, a10, a11)
    const int* bptr;       // per stored block: range in bsrc
    const int* bsrc;       // keb indices, element order within a block
    const int* bdst;       // position of the block in the sliced-ELL planes
    double2* va;           // FP64 plane (a00, a01)
    double2* vb;           // FP64 plane (a10, a11)
    float4* vf;            // FP32 values
    int ne, nblk, nbe8, nbb, plane;
    double h;
};

// formKe8 row i of element e: K_e(i, j) = h sum_g w_g detJ_g B_i^T C B_j.
static __global__ __launch_bounds__(kBlock) void k_q8_ke(Q8Args a)
{
    const int t = xcd_block(a.nbe8) * kBlock + threadIdx.x;
    if (t >= 8 * a.ne) return;
    const int e = t >> 3, i = t & 7;
    const int4 c0 = a.conn[2 * e], c1 = a.conn[2 * e + 1];
    const int cc[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    double x[8], y[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const double2 p = a.xy[cc[j]];
        x[j] = p.x;
        y[j] = p.y;
    }
    const double2 en = a.ecnu[e];
    double c00, c01, c22;
    elastic_c(a.plane, en.x, en.y, c00, c01, c22);
    double blk[8][4];
#pragma unroll
    for (int j = 0; j < 8; ++j) blk[j][0] = blk[j][1] = blk[j][2] = blk[j][3] = 0.;
    for (int gi = 0; gi < 3; ++gi)
        for (int gj = 0; gj < 3; ++gj) {
            double B0[8], B1[8], dJ;
            q8_grad(x, y, q8_point(gi), q8_point(gj), B0, B1, dJ);
            const double w = q8_weight(gi) * q8_weight(gj);
            double bi0 = B0[0], bi1 = B1[0];
#pragma unroll
            for (int k = 1; k < 8; ++k)
                if (k == i) {
                    bi0 = B0[k];
                    bi1 = B1[k];
                }
            // rows 2i, 2i+1 of B^T C detJ
            const double t00 = bi0 * c00 * dJ, t01 = bi0 * c01 * dJ, t02 = bi1 * c22 * dJ;
            const double t10 = bi1 * c01 * dJ, t11 = bi1 * c00 * dJ, t12 = bi0 * c22 * dJ;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                blk[j][0] += (t00 * B0[j] + t02 * B1[j]) * w;
                blk[j][1] += (t01 * B1[j] + t02 * B0[j]) * w;
                blk[j][2] += (t10 * B0[j] + t12 * B1[j]) * w;
                blk[j][3] += (t11 * B1[j] + t12 * B0[j]) * w;
            }
        }
    double4* out = a.keb + (size_t)t * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        out[j] = make_double4(blk[j][0] * a.h, blk[j][1] * a.h, blk[j][2] * a.h, blk[j][3] * a.h);
}

// One stored block per thread: sum of its element blocks in element order.
template <bool F64>
__global__ __launch_bounds__(kBlock) void k_q8_gather(Q8Args a)
{
    const int b = xcd_block(a.nbb) * kBlock + threadIdx.x;
    if (b >= a.nblk) return;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    for (int p = a.bptr[b]; p < a.bptr[b + 1]; ++p) {
        const double4 k = a.keb[a.bsrc[p]];
        s0 += k.x;
        s1 += k.y;
        s2 += k.z;
        s3 += k.w;
    }
    const int q = a.bdst[b];
    if (F64) {
        a.va[q] = make_double2(s0, s1);
        a.vb[q] = make_double2(s2, s3);
    } else {
        a.vf[q] = make_float4((float)s0, (float)s1, (float)s2, (float)s3);
    }
}

struct Q8ResArgs {
    const int4* conn;
    const double2* xy;
    const double2* ecnu;
    const double2* u;
    double2* fe;           // element forces, 8 per element
    const int* incPtr;
    const int* inc;        // node -> 8 e + i, element order
    const double2* b;      // NULL: output the internal force itself
    const uchar2* mask;
    double2* r64;
    float2* r32;           // optional FP32 copy
    RefScal* sc;
    double* partials;
    int ne, nn, nbe, nbn, plane;
    double h;
};

// Element internal forces f_e(i) = h sum_g w_g detJ_g B_i^T C (B_g u_e).
static __global__ __launch_bounds__(kBlock) void k_q8_force(Q8ResArgs a)
{
    const int e = xcd_block(a.nbe) * kBlock + threadIdx.x;
    if (e >= a.ne) return;
    const int4 c0 = a.conn[2 * e], c1 = a.conn[2 * e + 1];
    const int cc[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    double x[8], y[8], ux[8], uy[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const double2 p = a.xy[cc[j]];
        const double2 w = a.u[cc[j]];
        x[j] = p.x; y[j] = p.y; ux[j] = w.x; uy[j] = w.y;
    }
    const double2 en = a.ecnu[e];
    double c00, c01, c22;
    elastic_c(a.plane, en.x, en.y, c00, c01, c22);
    double fx[8], fy[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) fx[j] = fy[j] = 0.;
    for (int gi = 0; gi < 3; ++gi)
        for (int gj = 0; gj < 3; ++gj) {
            double B0[8], B1[8], dJ;
            q8_grad(x, y, q8_point(gi), q8_point(gj), B0, B1, dJ);
            double exx = 0., eyy = 0., gxy = 0.;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                exx += B0[k] * ux[k];
                eyy += B1[k] * uy[k];
                gxy += B1[k] * ux[k] + B0[k] * uy[k];
            }
            const double wd = q8_weight(gi) * q8_weight(gj) * dJ;
            const double sx = (c00 * exx + c01 * eyy) * wd, sy = (c01 * exx + c00 * eyy) * wd, tx = c22 * gxy * wd;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                fx[k] += B0[k] * sx + B1[k] * tx;
                fy[k] += B1[k] * sy + B0[k] * tx;
            }
        }
    double2* out = a.fe + (size_t)e * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) out[j] = make_double2(fx[j] * a.h, fy[j] * a.h);
}

// f(n) = sum of f_e(i) over the node's incidences (element order);
// r = mask (b - f) in FP64, its square norm partials for the stop test.
static __global__ __launch_bounds__(kBlock) void k_q8_fgather(Q8ResArgs a)
{
    const int lb = xcd_block(a.nbn);
    const int n = lb * kBlock + threadIdx.x;
    double v[1] = {0.};
    if (n < a.nn) {
        double fx = 0., fy = 0.;
        for (int q = a.incPtr[n]; q < a.incPtr[n + 1]; ++q) {
            const double2 f = a.fe[a.inc[q]];
            fx += f.x;
            fy += f.y;
        }
        double rx = fx, ry = fy;
        if (a.b) {
            const double2 bn = a.b[n];
            const uchar2 mk = a.mask[n];
            rx = mk.x ? bn.x - fx : 0.;
            ry = mk.y ? bn.y - fy : 0.;
            v[0] = rx * rx + ry * ry;
        }
        a.r64[n] = make_double2(rx, ry);
        if (a.r32) a.r32[n] = make_float2((float)rx, (float)ry);
    }
    if (a.sc) block_partial<1>(v, a.partials, lb);
}

}  // namespace cpk
