// Interior box geometry of the fused step and the CPML shell around it as boxes of
// quads, one thread per quad: the box lists (host) and the thread -> quad decode (host
// and device), shared by the fused step's kernels and a host-side coverage check.
#pragma once

#include "common.hpp"

namespace mifdtd {

struct FGeom {
    int nx, ny, nzp, nz;
    int T;       // rows per block
    int nrb;     // row blocks over [yb, ye)
    int cx;      // planes per chunk
    int nchunks; // chunks over [xb, xe)
    int xb, xe;  // planes of the interior box
    int yb, ye;  // rows of the interior box
    int kb, ke;  // z range of the interior box (on 128-byte lines of the field arrays)
    int zb0;     // first z segment start, at or below kb (may be negative: lanes below
                 // kb idle); segments sit on 1 KiB boundaries of the arrays
    int nseg;    // z segments of 256 floats from zb0 covering [kb, ke)
    int nblocks;
    int ntri;    // kCeTriple: CE triples in the table
};

// CPML shell: every cell outside the interior box, as up to six boxes (x low/high: all
// rows; y low/high: the box's planes; z low/high: the box's planes and rows), each
// (planes [i0,i1), rows [j0,j1), quads [q0,q1)) with a thread per quad; start[b] is the
// first thread of box b.  Quads run fastest, then the box's shorter one of planes and
// rows, so that the i+1 / i-1 and j+1 / j-1 neighbour rows a quad reads are read by
// nearby threads too (L2 hits).  A quad range may run past the row's Q quads into the
// next row (q >= Q: row j+1, quad q-Q): a "seam" box holds the z-high piece of row j
// and the z-low piece of row j+1, which are adjacent in memory and share a line when
// the arrays are off the line alignment.
constexpr int kMaxShellBoxes = 16;  // at most 7 shell boxes + 3 low-fringe + 3 high-fringe boxes
struct ShellBoxes {
    int n;
    int Q;  // quads per row
    int i0[kMaxShellBoxes], i1[kMaxShellBoxes], j0[kMaxShellBoxes], j1[kMaxShellBoxes], q0[kMaxShellBoxes],
        q1[kMaxShellBoxes];
    int i_inner[kMaxShellBoxes];  // planes before rows in the thread order
    unsigned start[kMaxShellBoxes + 1];
};

__host__ __device__ __forceinline__ void shell_quad(const ShellBoxes &s, unsigned f, int &i, int &j, int &k0)
{
    int b = 0;
    while (b + 1 < s.n && f >= s.start[b + 1]) ++b;
    unsigned r = f - s.start[b];
    const unsigned nq = (unsigned)(s.q1[b] - s.q0[b]), nj = (unsigned)(s.j1[b] - s.j0[b]);
    const unsigned ni = (unsigned)(s.i1[b] - s.i0[b]);
    const unsigned q = r % nq;
    r /= nq;
    // selects rather than branches on i_inner (which the compiler turns into stores of
    // i and j through a stack address)
    const bool ii = s.i_inner[b] != 0;
    const unsigned inner = ii ? ni : nj;
    const int a = (int)(r % inner), o = (int)(r / inner);
    const int qq = s.q0[b] + (int)q;
    const int wrap = qq >= s.Q ? 1 : 0;  // seam boxes: past the row end, the next row
    i = s.i0[b] + (ii ? a : o);
    j = s.j0[b] + (ii ? o : a) + wrap;
    k0 = 4 * (qq - wrap * s.Q);
}

// Appends the box (planes [i0,i1), rows [j0,j1), quads [q0,q1)) to sb unless it is empty.
inline void add_box(ShellBoxes &sb, unsigned long long &total, int i0, int i1, int j0, int j1, int q0, int q1)
{
    if (i1 <= i0 || j1 <= j0 || q1 <= q0) return;
    const int b = sb.n++;
    sb.i0[b] = i0; sb.i1[b] = i1; sb.j0[b] = j0; sb.j1[b] = j1; sb.q0[b] = q0; sb.q1[b] = q1;
    sb.i_inner[b] = (i1 - i0) < (j1 - j0);
    sb.start[b] = (unsigned)total;
    total += (unsigned long long)(i1 - i0) * (j1 - j0) * (q1 - q0);
}

// Line phase of the field arrays in floats: k = 0 of every row sits `phase` floats past
// a 128-byte line (0 when rows do not all start at the same phase, nz_pitch % 32 != 0).
inline int line_phase(const float *e0, int nzp)
{
    if (nzp % 32) return 0;
    return (int)((reinterpret_cast<uintptr_t>(e0) >> 2) & 31u);
}
// The first k >= lo and the last k <= hi on a line boundary of arrays of phase ph.
inline int line_up(int lo, int ph) { return 32 * ((lo + ph + 31) / 32) - ph; }
inline int line_down(int hi, int ph) { return hi + ph < 0 ? -1 : 32 * ((hi + ph) / 32) - ph; }
// Segment origin for arrays off the line alignment (kb on a line): the last k <= kb on a
// 1 KiB boundary of the arrays, so that every segment load of a wave covers one aligned
// KiB (the lanes before kb stay idle: clamped loads, no stores); kb itself when that k
// would be negative.  Arrays placed with kb on a 1 KiB boundary start the segments at kb.
inline int seg_origin(int kb, const float *e0)
{
    const int z = kb - (int)((kb + (reinterpret_cast<uintptr_t>(e0) >> 2)) & 255u);
    return z >= 0 ? z : kb;
}

// The cells outside g's box: x low/high slabs (all rows), y low/high (the box's
// planes), and the z pieces of the box's planes and rows -- as seams (the z-high piece
// of row j with the z-low piece of row j+1, one contiguous run) plus the z-low piece of
// the first row and the z-high piece of the last when both pieces exist, else as z
// low/high boxes.
inline void add_shell(ShellBoxes &sb, unsigned long long &total, const FGeom &g, int Q)
{
    sb.Q = Q;
    add_box(sb, total, 0, g.xb, 0, g.ny, 0, Q);
    add_box(sb, total, g.xe, g.nx, 0, g.ny, 0, Q);
    add_box(sb, total, g.xb, g.xe, 0, g.yb, 0, Q);
    add_box(sb, total, g.xb, g.xe, g.ye, g.ny, 0, Q);
    const int ql = g.kb / 4, qh = g.ke / 4;
    if (ql > 0 && qh < Q && g.ye - g.yb >= 2) {
        add_box(sb, total, g.xb, g.xe, g.yb, g.yb + 1, 0, ql);
        add_box(sb, total, g.xb, g.xe, g.yb, g.ye - 1, qh, Q + ql);
        add_box(sb, total, g.xb, g.xe, g.ye - 1, g.ye, qh, Q);
    } else {
        add_box(sb, total, g.xb, g.xe, g.yb, g.ye, 0, ql);
        add_box(sb, total, g.xb, g.xe, g.yb, g.ye, qh, Q);
    }
}

// The box's low fringe (plane xb, row yb, quad kb/4: the box cells whose step n+1 H the
// shell's step n+2 E update reads) as three disjoint boxes.
inline void add_low_fringe(ShellBoxes &sb, unsigned long long &total, const FGeom &g)
{
    const int qb = g.kb / 4, qe = g.ke / 4;
    add_box(sb, total, g.xb, g.xb + 1, g.yb, g.ye, qb, qe);
    add_box(sb, total, g.xb + 1, g.xe, g.yb, g.yb + 1, qb, qe);
    add_box(sb, total, g.xb + 1, g.xe, g.yb + 1, g.ye, qb, qb + 1);
}

// The box's high fringe (plane xe-1, row ye-1, quad ke/4-1: the box cells whose step n+1
// E the shell's step n+1 H update reads) less the low fringe, as three disjoint boxes
// (needs at least two planes and rows and two quads).
inline void add_high_fringe(ShellBoxes &sb, unsigned long long &total, const FGeom &g)
{
    const int qb = g.kb / 4, qe = g.ke / 4;
    add_box(sb, total, g.xe - 1, g.xe, g.yb + 1, g.ye, qb + 1, qe);
    add_box(sb, total, g.xb + 1, g.xe - 1, g.ye - 1, g.ye, qb + 1, qe);
    add_box(sb, total, g.xb + 1, g.xe - 1, g.yb + 1, g.ye - 1, qe - 1, qe);
}

}  // namespace mifdtd
