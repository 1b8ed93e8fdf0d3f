// Index arithmetic of the resident migrate kernel (k_migrate_ph and its
// counting and multi-problem forms), shared by the kernel and by a host-side
// replay of every index one launch can form (tests/bounds_replay.cpp, built
// with g++): plain integer functions, so both compile the same expressions.
#pragma once

#include <cstddef>

#if defined(__HIPCC__)
#define KMIG_HD __host__ __device__
#else
#define KMIG_HD
#endif

namespace kmig {
namespace idx {

KMIG_HD inline int imin(int a, int b) { return a < b ? a : b; }

// Bound cell jx, jz of a wave whose first cell is (cx0, cz0): cell-major
// index cx * ncz + cz, clamped to the last cell row and column (a wave at the
// image's bottom or right edge can reach past them).
KMIG_HD inline int cell(int cx0, int jx, int cz0, int jz, int ncx, int ncz)
{
    return imin(cx0 + jx, ncx - 1) * ncz + imin(cz0 + jz, ncz - 1);
}

// Entry j of a cell's row in a cell-major bounds array with `stride` entries
// per cell (source bounds per shot, receiver bounds per slot, or the union
// bounds per block of shots).
KMIG_HD inline size_t bnd(int cell, size_t stride, int j) { return (size_t)cell * stride + j; }

// Trace-aperture weight read for receiver g of the chunk starting at receiver
// index r0, shot is (fixed receivers): g past nh - r0 reads the first entries
// of the next shot's row (or the pad after the last shot).
KMIG_HD inline size_t wap(int is, int nh, int r0, int g) { return (size_t)is * nh + r0 + g; }

// Trace (row of tap pairs) read for receiver g of the chunk starting at r0,
// shot is (fixed receivers): the chunk's first row plus tap_row_off, where
// receivers past nh, which pad the last chunk and have zero weight, read the
// shot's last real trace, never the next shot's.
KMIG_HD inline size_t chunk_row(int is, int nh, int r0) { return (size_t)is * nh + r0; }
KMIG_HD inline int tap_row_off(int nh, int r0, int g) { return imin(g, nh - 1 - r0); }
KMIG_HD inline size_t tap_row(int is, int nh, int r0, int g)
{
    return chunk_row(is, nh, r0) + tap_row_off(nh, r0, g);
}

// Receiver-column table entry of launch-local shot jc (batch shot
// scell_off + jc), column r0 + g; rows are padded past the last.
KMIG_HD inline size_t ctab(int scell_off, int jc, int ncols, int r0, int g)
{
    return (size_t)(scell_off + jc) * ncols + r0 + g;
}

// Trace of shot is at receiver index ih (a column's trace in that shot).
KMIG_HD inline size_t col_row(int is, int nh, int ih) { return (size_t)is * nh + ih; }

// Survivor word w (launch-local shots 64 w .. 64 w + 63) of receiver chunk k
// in bound cell `cell`, in a mask of nchunk chunks and nwords words per cell.
KMIG_HD inline size_t mask_word(int cell, int nchunk, int k, int nwords, int w)
{
    return ((size_t)cell * nchunk + k) * nwords + w;
}

}  // namespace idx
}  // namespace kmig
