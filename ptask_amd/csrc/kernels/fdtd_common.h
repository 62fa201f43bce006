// fdtd_common.h -- shared device helpers for the Yee update kernels.
//
// Work decomposition used by every FDTD kernel in this directory.  A field array is a stack of
// d3 planes of d1*d2 cells (i fastest).  The launch is viewed as
//   * "column lanes": every thread of every block along grid x, numbered
//     lane = (bx*blockDim.y + threadIdx.y)*blockDim.x + threadIdx.x; lanes stride over the d1*d2
//     columns of a plane, and
//   * "k slabs": the gridDim.y*gridDim.z blocks along y and z split [0,d3) into equal slabs.
// Every (column, k) pair is therefore visited exactly once for ANY launch geometry, including
// the one-thread-per-cell geometry the reference sets with SetBlockAndGridSize
// (ptaskUnitTest/graphfdtd.cpp:439-446).  A lane walks its column upwards through its slab, so
// the k-neighbour plane is carried in a register instead of being re-read.
//
// Block ids are remapped so that each XCD (workgroups are dealt to the 8 XCDs round-robin)
// receives a contiguous run of the (slab, column-block) order: the j-neighbour rows a block
// reads were fetched by blocks of the same XCD and are served from that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

#include "ptk_fdtd.h"

namespace ptk_fdtd {

constexpr unsigned kNumXcd = 8;

// Column-walk unroll depth: planes whose loads are issued together before their stores.
#ifndef PTK_FDTD_UNROLL
#define PTK_FDTD_UNROLL 1
#endif
// XCD-contiguous block order. Measured on MI355X it costs 3-8% on these streaming stencils
// (each XCD then drives a narrower set of HBM channels), so the default is the hardware order.
#ifndef PTK_FDTD_XCD_REMAP
#define PTK_FDTD_XCD_REMAP 0
#endif

struct Sweep {
    unsigned lane;     // first column of this lane
    unsigned nlanes;   // column stride
    int k0, k1;        // this block's k slab [k0, k1)
};

// Bijective XCD-aware renumbering of the linear block id.
__device__ inline unsigned xcd_remap(unsigned b, unsigned n) {
    if (!PTK_FDTD_XCD_REMAP || n < kNumXcd) return b;
    const unsigned xcd = b % kNumXcd, local = b / kNumXcd;
    const unsigned per = n / kNumXcd, rem = n % kNumXcd;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

__device__ inline Sweep make_sweep(int d3) {
    const unsigned gx = gridDim.x, gyz = gridDim.y * gridDim.z;
    const unsigned nblk = gx * gyz;
    const unsigned b = blockIdx.x + gx * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned r = xcd_remap(b, nblk);
    const unsigned bx = r % gx, slab = r / gx;
    Sweep s;
    const unsigned tpb = blockDim.x * blockDim.y * blockDim.z;
    s.lane = bx * tpb + (threadIdx.z * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
    s.nlanes = gx * tpb;
    s.k0 = (int)((unsigned long long)slab * (unsigned)d3 / gyz);
    s.k1 = (int)((unsigned long long)(slab + 1) * (unsigned)d3 / gyz);
    return s;
}

// Flat offset of (i,j,k) in an array of extents (x1,x2,*).
__device__ inline long long at(int i, int j, int k, int x1, int x2) {
    return (long long)i + (long long)j * x1 + (long long)k * ((long long)x1 * x2);
}

}  // namespace ptk_fdtd
