// oz_tiles.h -- layout shared by the tiled q-space kernels (k_oz.hip, FP64;
// k_oz32.hip, FP32): tile size, the host-built integer table and the ordering
// helpers of the rotation matrices.
#pragma once

#include "mdft_internal.h"

namespace oz {

constexpr int kMaxL = MDFT_MAX_MMAX;
constexpr int kMaxW = 2 * kMaxL + 1;
constexpr int kP = 8;            // (q, -q) pairs per tile: 2 kP = 16 (pair, side) lanes per OZ row
constexpr int kTh = 256;         // threads per workgroup
constexpr int kNKhi = 2 * kMaxL + 1;
constexpr int kCMax = 32;        // OZ columns per khi (18 at mmax = 5)
constexpr int kChunkCols = 4;    // khi (or mup) outputs per rotation task

// compact index of the (2m+1)^2 block of order m: sum_{l<m} (2l+1)^2
__host__ __device__ inline int roff(int m) { return m * (4 * m * m - 1) / 3; }

// Integer table (host-built by oz_tables, staged in LDS once per workgroup), in order:
//   dd   [nR]        entry e of the d^m(theta) blocks: Fourier-table offset << 4 | m << 1 | (a - b odd)
//   tin  [ntin][4]   rotation-in tasks {m, mu2, khi0, nkhi}
//   tout [ntout][4]  rotation-back tasks {m, mu2, mup0, nmup}
//   orow [np]        OZ rows in order of decreasing length (the 4 rows of a wave alike)
//   ia0  [np]        first OZ coefficient of each row (the relayout order, cproj_mrso.f90:381-401)
//   pmup [np]        mup of each projection
//   ncol [kNKhi]     OZ columns of each khi;  ctab [kNKhi][kCMax]: A offset << 1 | (nu < 0)
//   poff [kMaxL + 2] first projection of each m
struct Layout {
    int nR, ntin, ntout, np;
    __host__ __device__ int dd() const { return 0; }
    __host__ __device__ int tin() const { return nR; }
    __host__ __device__ int tout() const { return tin() + 4 * ntin; }
    __host__ __device__ int orow() const { return tout() + 4 * ntout; }
    __host__ __device__ int ia0() const { return orow() + np; }
    __host__ __device__ int pmup() const { return ia0() + np; }
    __host__ __device__ int ncol() const { return pmup() + np; }
    __host__ __device__ int ctab() const { return ncol() + kNKhi; }
    __host__ __device__ int poff() const { return ctab() + kNKhi * kCMax; }
    __host__ __device__ int size() const { return poff() + kMaxL + 2; }
};

}  // namespace oz

// Fourier coefficients of d^m(theta) (FP64) and the integer table of oz::Layout for
// the context's angular order (molecular symmetry order 2, mmax <= MDFT_MAX_MMAX).
int oz_tables(mdft_ctx *c, std::vector<double> &ftab, std::vector<int> &itab, int *ntin, int *ntout);
