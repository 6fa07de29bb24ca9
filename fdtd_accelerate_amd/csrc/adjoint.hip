// Discrete adjoint of the in-place Yee + CPML step (update_e, CPML-E, sources,
// update_h, CPML-H on the KEMP layout), for gradients of an objective with respect to
// the CE coefficients of a design box.
//
// The forward step n maps (E, H, psiE, psiH) -> (E', H', psiE', psiH'):
//   E half:  Ec = E + CE*(K H) on valid E cells, 0 on masked lanes, E on walls;
//            psiE' = b_E psiE + a_E (D H);  E' = Ec -/+ CE psiE'   (+ sources)
//   H half:  Hc = H - 0.5*(Kt E') on valid H cells, 0 on masked lanes, H on walls;
//            psiH' = b_H psiH + a_H (Dt E');  H' = Hc +/- 0.5 psiH'
// with K the forward-difference curl of update_e and Kt = K^T the backward-difference
// curl of update_h.  Its transpose, run backward in time from the adjoint state
// (lE, lH, lpsiE, lpsiH) of the step's outputs, is four passes:
//   A  (E-like): lE += dJ/dE' - 0.5*K(lH restricted to valid H) + Dt^T(a_H t_H),
//                t_H = lpsiH +/- 0.5*lH on valid H slab cells;  lH = 0 on masked H lanes
//   B:           lpsiH = b_H * t_H
//   C  (H-like): lH += Kt(CE * lE restricted to valid E) + D^T(a_E t_E),
//                t_E = lpsiE -/+ CE*lE on valid E slab cells;  grad_CE += lE * (K H_prev)
//                on the design box;  lE = 0 on masked E lanes
//   D:           lpsiE = b_E * t_E
// A and C gather from neighbours (thread per cell), so every pass reads only values no
// thread of the same pass writes.  B only touches the cell's own lH / lpsiH, so it runs
// inside C (before C changes lH of that cell); D only touches the cell's own lE / lpsiE,
// so it is deferred into A of the next (earlier) step, before A changes lE of that cell
// (on the zero initial state it is a no-op).  Two launches per step.  Sources do not
// depend on the state (no adjoint).

#include "common.hpp"

namespace mifdtd {

constexpr int AB = 256;

struct AGeom {
    int nx, ny, nzp, nz, npml;
    unsigned long long ncell;
    size_t sX, sY, sZ;  // packed psi face sizes
    __device__ __forceinline__ size_t off(int f, int c) const
    {
        const int ax = f >> 1, sd = f & 1;
        if (ax == 0) return (size_t)(2 * sd + c) * sX;
        if (ax == 1) return 4 * sX + (size_t)(2 * sd + c) * sY;
        return 4 * sX + 4 * sY + (size_t)(2 * sd + c) * sZ;
    }
};

// Which cells the forward core writes (valid), zeroes (masked lanes) or leaves (walls).
__device__ __forceinline__ bool e_row(int c, int i, int j, int nx, int ny)
{
    return c == 0 ? j < ny - 1 : c == 1 ? i < nx - 1 : (i < nx - 1 && j < ny - 1);
}
__device__ __forceinline__ bool e_lane(int c, int k, int nz) { return c == 2 ? k < nz : k < nz - 1; }
__device__ __forceinline__ bool h_row(int c, int i, int j)
{
    return c == 0 ? j > 0 : c == 1 ? i > 0 : (i > 0 && j > 0);
}
__device__ __forceinline__ bool h_lane(int c, int k, int nz) { return c == 2 ? k < nz : (k > 0 && k < nz); }

__device__ __forceinline__ bool in_grid(int i, int j, int k, const AGeom &g)
{
    return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nzp;
}
__device__ __forceinline__ size_t cidx(int i, int j, int k, const AGeom &g)
{
    return ((size_t)i * g.ny + j) * g.nzp + k;
}
__device__ __forceinline__ bool valid_e(int c, int i, int j, int k, const AGeom &g)
{
    return in_grid(i, j, k, g) && e_row(c, i, j, g.nx, g.ny) && e_lane(c, k, g.nz);
}
__device__ __forceinline__ bool valid_h(int c, int i, int j, int k, const AGeom &g)
{
    return in_grid(i, j, k, g) && h_row(c, i, j) && h_lane(c, k, g.nz);
}

// CPML slab lookup along one axis: returns the table entry (>0) of coordinate q, or -1,
// and the slab side / depth p.  E planes: front p -> p+1, back n-1-npml+p -> npml+1+p;
// H planes: front p -> p+1, back n-npml+p -> npml+2+p.
__device__ __forceinline__ int tab_e(int q, int n, int npml, int &sd, int &p)
{
    sd = 0;
    p = 0;
    if (q >= 0 && q < npml) { p = q; return p + 1; }
    if (q >= n - 1 - npml && q < n - 1) { sd = 1; p = q - (n - 1 - npml); return npml + 1 + p; }
    return -1;
}
__device__ __forceinline__ int tab_h(int q, int n, int npml, int &sd, int &p)
{
    sd = 0;
    p = 0;
    if (q >= 0 && q < npml) { p = q; return p + 1; }
    if (q >= n - npml && q < n) { sd = 1; p = q - (n - npml); return npml + 2 + p; }
    return -1;
}

// Reach of the CPML adjoint terms along one axis (coordinate q of n): pass A adds
// Dt^T(a_H t_H) at an E cell when q or q+1 is an H slab plane (q < npml or
// q >= n-npml-1), and pass D (inside A) acts on E slab planes, a subset; pass C adds
// D^T(a_E t_E) at an H cell when q-1 or q is an E slab plane (q <= npml or
// q >= n-1-npml), and pass B (inside C) acts on H slab planes, a subset.  Elsewhere
// those terms are zero and are not formed.
__device__ __forceinline__ bool near_a(int q, int n, int npml) { return q < npml || q >= n - npml - 1; }
__device__ __forceinline__ bool near_c(int q, int n, int npml) { return q <= npml || q >= n - 1 - npml; }

// Cell (i, j, k) of flat index n (< 2^32: the launch checks it).
__device__ __forceinline__ void decode(unsigned n, const AGeom &g, int &i, int &j, int &k)
{
    const unsigned r = n / (unsigned)g.nzp;
    k = (int)(n - r * (unsigned)g.nzp);
    i = (int)(r / (unsigned)g.ny);
    j = (int)(r - (unsigned)i * (unsigned)g.ny);
}

// packed psi index of cell (i, j, k) on face (axis, side) at depth p
__device__ __forceinline__ size_t psi_at(const AGeom &g, int axis, int sd, int p, int c, int i, int j, int k)
{
    const int f = 2 * axis + sd;
    size_t r;
    if (axis == 0) r = ((size_t)p * g.ny + j) * g.nzp + k;
    else if (axis == 1) r = ((size_t)i * g.npml + p) * g.nzp + k;
    else r = ((size_t)i * g.ny + j) * g.npml + p;
    return g.off(f, c) + r;
}

// ------------------------------------------------------------------- pass A helpers
// t_H at H cell (i,j,k) of component hc for the face (axis), psi comp pc and sign:
// t = lpsiH + sgn*0.5*lH, or 0 when that H cell is not a valid slab cell of the face.
__device__ __forceinline__ float tH(const AGeom &g, const float *__restrict__ lH, const float *__restrict__ lpH,
                                    int axis, int pc, int hc, float sgn, int i, int j, int k, float &a_out,
                                    const float *__restrict__ aH)
{
    a_out = 0.f;
    if (!valid_h(hc, i, j, k, g)) return 0.f;
    const int q = axis == 0 ? i : axis == 1 ? j : k;
    const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
    int sd, p;
    const int t = tab_h(q, n, g.npml, sd, p);
    if (t < 0) return 0.f;
    a_out = aH[t];
    return lpH[psi_at(g, axis, sd, p, pc, i, j, k)] + sgn * 0.5f * lH[cidx(i, j, k, g)];
}

// Dt^T(a t) at E cell q along the axis: a(q) t(q) - a(q+1) t(q+1)
__device__ __forceinline__ float dtT(const AGeom &g, const float *lH, const float *lpH, const float *aH, int axis,
                                     int pc, int hc, float sgn, int i, int j, int k)
{
    float a0, a1;
    const float t0 = tH(g, lH, lpH, axis, pc, hc, sgn, i, j, k, a0, aH);
    const float t1 = tH(g, lH, lpH, axis, pc, hc, sgn, i + (axis == 0), j + (axis == 1), k + (axis == 2), a1, aH);
    return a0 * t0 - a1 * t1;
}

__device__ __forceinline__ float lhv(const AGeom &g, const float *__restrict__ l, int c, int i, int j, int k)
{
    return valid_h(c, i, j, k, g) ? l[cidx(i, j, k, g)] : 0.f;
}

__device__ __forceinline__ void pass_d_cell(const AGeom &g, int i, int j, int k, size_t n, const float *const LE[3],
                                            const float *const CE[3], float *lpE, const float *bE);
__device__ __forceinline__ void pass_b_cell(const AGeom &g, int i, int j, int k, size_t n, const float *const LH[3],
                                            float *lpH, const float *bH);

__global__ __launch_bounds__(AB) void k_adj_a(AGeom g, float *__restrict__ lex, float *__restrict__ ley,
                                              float *__restrict__ lez, float *__restrict__ lhx,
                                              float *__restrict__ lhy, float *__restrict__ lhz,
                                              const float *__restrict__ lpH, const float *__restrict__ aH,
                                              const float *__restrict__ eplane, int iplane,
                                              const float *__restrict__ cex, const float *__restrict__ cey,
                                              const float *__restrict__ cez, float *__restrict__ lpE,
                                              const float *__restrict__ bE)
{
    const unsigned n = blockIdx.x * AB + threadIdx.x;
    if (n >= g.ncell) return;
    int i, j, k;
    decode(n, g, i, j, k);
    const size_t id = (size_t)n;
    const bool nx_ = g.npml > 0 && near_a(i, g.nx, g.npml), ny_ = g.npml > 0 && near_a(j, g.ny, g.npml),
               nz_ = g.npml > 0 && near_a(k, g.nz, g.npml);
    if (nx_ || ny_ || nz_) {  // pass D of the later step, on this cell (reads lE before A changes it)
        const float *LEd[3] = {lex, ley, lez};
        const float *CEd[3] = {cex, cey, cez};
        pass_d_cell(g, i, j, k, id, LEd, CEd, lpE, bE);
    }
    const float *LH[3] = {lhx, lhy, lhz};
    // -0.5*K(lH valid):  ex: (z[j+1]-z) - (y[k+1]-y);  ey: (x[k+1]-x) - (z[i+1]-z);
    //                     ez: (y[i+1]-y) - (x[j+1]-x)
    float ax = 0.f, ay = 0.f, az = 0.f;
    ax = (lhv(g, LH[2], 2, i, j + 1, k) - lhv(g, LH[2], 2, i, j, k)) - (lhv(g, LH[1], 1, i, j, k + 1) - lhv(g, LH[1], 1, i, j, k));
    ay = (lhv(g, LH[0], 0, i, j, k + 1) - lhv(g, LH[0], 0, i, j, k)) - (lhv(g, LH[2], 2, i + 1, j, k) - lhv(g, LH[2], 2, i, j, k));
    az = (lhv(g, LH[1], 1, i + 1, j, k) - lhv(g, LH[1], 1, i, j, k)) - (lhv(g, LH[0], 0, i, j + 1, k) - lhv(g, LH[0], 0, i, j, k));
    float vx = lex[id] - 0.5f * ax, vy = ley[id] - 0.5f * ay, vz = lez[id] - 0.5f * az;
    if (nx_) {  // x: psi1 -> hy from ez (+), psi2 -> hz from ey (-)
        vz += dtT(g, lhy, lpH, aH, 0, 0, 1, +1.f, i, j, k);
        vy += dtT(g, lhz, lpH, aH, 0, 1, 2, -1.f, i, j, k);
    }
    if (ny_) {  // y: psi1 -> hz from ex (+), psi2 -> hx from ez (-)
        vx += dtT(g, lhz, lpH, aH, 1, 0, 2, +1.f, i, j, k);
        vz += dtT(g, lhx, lpH, aH, 1, 1, 0, -1.f, i, j, k);
    }
    if (nz_) {  // z: psi1 -> hx from ey (+), psi2 -> hy from ex (-)
        vy += dtT(g, lhx, lpH, aH, 2, 0, 0, +1.f, i, j, k);
        vx += dtT(g, lhy, lpH, aH, 2, 1, 1, -1.f, i, j, k);
    }
    if (eplane && i == iplane) {
        const size_t pid = (size_t)j * g.nzp + k;
        const size_t ps = (size_t)g.ny * g.nzp;
        vx += 2.f * eplane[pid];
        vy += 2.f * eplane[ps + pid];
        vz += 2.f * eplane[2 * ps + pid];
    }
    lex[id] = vx;
    ley[id] = vy;
    lez[id] = vz;
    // H masked lanes: H' does not depend on H there
    for (int c = 0; c < 3; ++c)
        if (h_row(c, i, j) && !h_lane(c, k, g.nz)) (c == 0 ? lhx : c == 1 ? lhy : lhz)[id] = 0.f;
}

// Pass B on one cell: lpsiH = b_H * (lpsiH + sgn*0.5*lH) on its valid H slab entries.
__device__ __forceinline__ void pass_b_cell(const AGeom &g, int i, int j, int k, size_t n, const float *const LH[3],
                                            float *lpH, const float *bH)
{
    // (axis, psi comp) -> (H comp, sign)
    const int hcs[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    const float sg[2] = {+1.f, -1.f};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int q = ax == 0 ? i : ax == 1 ? j : k;
        const int nn = ax == 0 ? g.nx : ax == 1 ? g.ny : g.nz;
        int sd, p;
        const int t = tab_h(q, nn, g.npml, sd, p);
        if (t < 0) continue;
#pragma unroll
        for (int pc = 0; pc < 2; ++pc) {
            const int hc = hcs[ax][pc];
            if (!valid_h(hc, i, j, k, g)) continue;
            const size_t pi = psi_at(g, ax, sd, p, pc, i, j, k);
            lpH[pi] = bH[t] * (lpH[pi] + sg[pc] * 0.5f * LH[hc][n]);
        }
    }
}

// ------------------------------------------------------------------- pass C helpers
__device__ __forceinline__ float gev(const AGeom &g, const float *__restrict__ l, const float *__restrict__ ce, int c,
                                     int i, int j, int k)
{
    if (!valid_e(c, i, j, k, g)) return 0.f;
    const size_t q = cidx(i, j, k, g);
    return (ce ? ce[q] : 0.5f) * l[q];
}

// t_E at E cell (i,j,k): lpsiE + sgn*CE*lE on valid E slab cells of the face, else 0.
__device__ __forceinline__ float tE(const AGeom &g, const float *__restrict__ lE, const float *__restrict__ ce,
                                    const float *__restrict__ lpE, int axis, int pc, int ec, float sgn, int i, int j,
                                    int k, float &a_out, const float *__restrict__ aE)
{
    a_out = 0.f;
    if (!valid_e(ec, i, j, k, g)) return 0.f;
    const int q = axis == 0 ? i : axis == 1 ? j : k;
    const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
    int sd, p;
    const int t = tab_e(q, n, g.npml, sd, p);
    if (t < 0) return 0.f;
    a_out = aE[t];
    const size_t c = cidx(i, j, k, g);
    return lpE[psi_at(g, axis, sd, p, pc, i, j, k)] + sgn * (ce ? ce[c] : 0.5f) * lE[c];
}

// D^T(a t) at H cell q along the axis: a(q-1) t(q-1) - a(q) t(q)
__device__ __forceinline__ float dT(const AGeom &g, const float *lE, const float *ce, const float *lpE,
                                    const float *aE, int axis, int pc, int ec, float sgn, int i, int j, int k)
{
    float am, a0;
    const float tm = tE(g, lE, ce, lpE, axis, pc, ec, sgn, i - (axis == 0), j - (axis == 1), k - (axis == 2), am, aE);
    const float t0 = tE(g, lE, ce, lpE, axis, pc, ec, sgn, i, j, k, a0, aE);
    return am * tm - a0 * t0;
}

struct ADesign {
    int i0, j0, k0, di, dj, dk;  // box origin and extent
    const float *curl;           // (3, di, dj, dk) curl of H_prev for this step
    double *grad;                // (3, di, dj, dk) accumulated dJ/dCE
};

__global__ __launch_bounds__(AB) void k_adj_c(AGeom g, float *__restrict__ lex, float *__restrict__ ley,
                                              float *__restrict__ lez, float *__restrict__ lhx,
                                              float *__restrict__ lhy, float *__restrict__ lhz,
                                              const float *__restrict__ cex, const float *__restrict__ cey,
                                              const float *__restrict__ cez, const float *__restrict__ lpE,
                                              const float *__restrict__ aE, ADesign dz,
                                              float *__restrict__ lpH, const float *__restrict__ bH)
{
    const unsigned n = blockIdx.x * AB + threadIdx.x;
    if (n >= g.ncell) return;
    int i, j, k;
    decode(n, g, i, j, k);
    const size_t id = (size_t)n;
    const bool nx_ = g.npml > 0 && near_c(i, g.nx, g.npml), ny_ = g.npml > 0 && near_c(j, g.ny, g.npml),
               nz_ = g.npml > 0 && near_c(k, g.nz, g.npml);
    if (nx_ || ny_ || nz_) {  // pass B on this cell (reads lH before C changes it)
        const float *LHb[3] = {lhx, lhy, lhz};
        pass_b_cell(g, i, j, k, id, LHb, lpH, bH);
    }
    const float *LE[3] = {lex, ley, lez};
    const float *CE[3] = {cex, cey, cez};
    // Kt(CE*lE valid): hx: (z - z[j-1]) - (y - y[k-1]); hy: (x - x[k-1]) - (z - z[i-1]);
    //                  hz: (y - y[i-1]) - (x - x[j-1])
    const float bx = (gev(g, LE[2], CE[2], 2, i, j, k) - gev(g, LE[2], CE[2], 2, i, j - 1, k)) -
                     (gev(g, LE[1], CE[1], 1, i, j, k) - gev(g, LE[1], CE[1], 1, i, j, k - 1));
    const float by = (gev(g, LE[0], CE[0], 0, i, j, k) - gev(g, LE[0], CE[0], 0, i, j, k - 1)) -
                     (gev(g, LE[2], CE[2], 2, i, j, k) - gev(g, LE[2], CE[2], 2, i - 1, j, k));
    const float bz = (gev(g, LE[1], CE[1], 1, i, j, k) - gev(g, LE[1], CE[1], 1, i - 1, j, k)) -
                     (gev(g, LE[0], CE[0], 0, i, j, k) - gev(g, LE[0], CE[0], 0, i, j - 1, k));
    float vx = lhx[id] + bx, vy = lhy[id] + by, vz = lhz[id] + bz;
    if (nx_) {  // x: psi1 -> ey from hz (-CE), psi2 -> ez from hy (+CE)
        vz += dT(g, ley, cey, lpE, aE, 0, 0, 1, -1.f, i, j, k);
        vy += dT(g, lez, cez, lpE, aE, 0, 1, 2, +1.f, i, j, k);
    }
    if (ny_) {  // y: psi1 -> ez from hx (-), psi2 -> ex from hz (+)
        vx += dT(g, lez, cez, lpE, aE, 1, 0, 2, -1.f, i, j, k);
        vz += dT(g, lex, cex, lpE, aE, 1, 1, 0, +1.f, i, j, k);
    }
    if (nz_) {  // z: psi1 -> ex from hy (-), psi2 -> ey from hx (+)
        vy += dT(g, lex, cex, lpE, aE, 2, 0, 0, -1.f, i, j, k);
        vx += dT(g, ley, cey, lpE, aE, 2, 1, 1, +1.f, i, j, k);
    }
    lhx[id] = vx;
    lhy[id] = vy;
    lhz[id] = vz;
    // gradient of the CE of the design box: lE * (K H_prev)
    const int di = i - dz.i0, dj = j - dz.j0, dk = k - dz.k0;
    if (dz.grad && di >= 0 && di < dz.di && dj >= 0 && dj < dz.dj && dk >= 0 && dk < dz.dk) {
        const size_t q = ((size_t)di * dz.dj + dj) * dz.dk + dk;
        const size_t s = (size_t)dz.di * dz.dj * dz.dk;
        for (int c = 0; c < 3; ++c)
            if (valid_e(c, i, j, k, g)) dz.grad[c * s + q] += (double)LE[c][id] * (double)dz.curl[c * s + q];
    }
    // E masked lanes: E' does not depend on E there
    for (int c = 0; c < 3; ++c)
        if (e_row(c, i, j, g.nx, g.ny) && !e_lane(c, k, g.nz)) (c == 0 ? lex : c == 1 ? ley : lez)[id] = 0.f;
}

// Pass D on one cell: lpsiE = b_E * (lpsiE + sgn*CE*lE) on its valid E slab entries.
__device__ __forceinline__ void pass_d_cell(const AGeom &g, int i, int j, int k, size_t n, const float *const LE[3],
                                            const float *const CE[3], float *lpE, const float *bE)
{
    const int ecs[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    const float sg[2] = {-1.f, +1.f};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int q = ax == 0 ? i : ax == 1 ? j : k;
        const int nn = ax == 0 ? g.nx : ax == 1 ? g.ny : g.nz;
        int sd, p;
        const int t = tab_e(q, nn, g.npml, sd, p);
        if (t < 0) continue;
#pragma unroll
        for (int pc = 0; pc < 2; ++pc) {
            const int ec = ecs[ax][pc];
            if (!valid_e(ec, i, j, k, g)) continue;
            const size_t pi = psi_at(g, ax, sd, p, pc, i, j, k);
            lpE[pi] = bE[t] * (lpE[pi] + sg[pc] * (CE[ec] ? CE[ec][n] : 0.5f) * LE[ec][n]);
        }
    }
}

// curl of H at the E points of the design box, as update_e forms it:
//   ex: (hz[j+1]-hz) - (hy[k+1]-hy);  ey: (hx[k+1]-hx) - (hz[i+1]-hz);  ez: (hy[i+1]-hy) - (hx[j+1]-hx)
__global__ __launch_bounds__(AB) void k_design_curl(AGeom g, ADesign dz, const float *__restrict__ hx,
                                                    const float *__restrict__ hy, const float *__restrict__ hz,
                                                    float *__restrict__ out)
{
    const unsigned long long s = (unsigned long long)dz.di * dz.dj * dz.dk;
    const unsigned long long n = (unsigned long long)blockIdx.x * AB + threadIdx.x;
    if (n >= s) return;
    const int dk = (int)(n % (unsigned long long)dz.dk);
    const unsigned long long r = n / (unsigned long long)dz.dk;
    const int dj = (int)(r % (unsigned long long)dz.dj);
    const int di = (int)(r / (unsigned long long)dz.dj);
    const int i = dz.i0 + di, j = dz.j0 + dj, k = dz.k0 + dk;
    const size_t id = cidx(i, j, k, g), sj = (size_t)g.nzp, si = (size_t)g.ny * g.nzp;
    out[n] = valid_e(0, i, j, k, g) ? (hz[id + sj] - hz[id]) - (hy[id + 1] - hy[id]) : 0.f;
    out[s + n] = valid_e(1, i, j, k, g) ? (hx[id + 1] - hx[id]) - (hz[id + si] - hz[id]) : 0.f;
    out[2 * s + n] = valid_e(2, i, j, k, g) ? (hy[id + si] - hy[id]) - (hx[id + sj] - hx[id]) : 0.f;
}

static int adj_geom(AGeom &g, int nx, int ny, int nzp, int pad, int npml)
{
    if (int rc = check_grid(nx, ny, nzp, pad)) return rc;
    g.nx = nx;
    g.ny = ny;
    g.nzp = nzp;
    g.nz = nzp - pad;
    g.npml = npml;
    if (npml < 0 || npml > MIFDTD_MAX_NPML) return fail(MIFDTD_EINVAL, "npml %d", npml);
    if (npml && (nx < 2 * npml + 2 || ny < 2 * npml + 2 || g.nz < 2 * npml + 2))
        return fail(MIFDTD_EINVAL, "grid too small for npml %d", npml);
    g.ncell = (unsigned long long)nx * ny * nzp;
    g.sX = (size_t)npml * ny * nzp;
    g.sY = (size_t)nx * npml * nzp;
    g.sZ = (size_t)nx * ny * npml;
    return 0;
}

static int check_box(const AGeom &g, const mifdtd_design_box *d)
{
    if (!d) return fail(MIFDTD_ENULL, "design box is NULL");
    if (d->i0 < 0 || d->j0 < 0 || d->k0 < 0 || d->di < 1 || d->dj < 1 || d->dk < 1 || d->i0 + d->di > g.nx ||
        d->j0 + d->dj > g.ny || d->k0 + d->dk > g.nz)
        return fail(MIFDTD_EINVAL, "design box outside the grid");
    // The gradient accumulates dE'/dCE = K H only: an E cell of a CPML slab also carries
    // the -/+CE*psi term, whose derivative is not formed, so the box must keep off every
    // E slab plane (front [0, npml), back [n-1-npml, n-1) on each axis).
    if (g.npml > 0) {
        const int np_ = g.npml;
        if (d->i0 < np_ || d->j0 < np_ || d->k0 < np_ || d->i0 + d->di > g.nx - 1 - np_ ||
            d->j0 + d->dj > g.ny - 1 - np_ || d->k0 + d->dk > g.nz - 1 - np_)
            return fail(MIFDTD_EINVAL, "design box (%d,%d,%d)+(%d,%d,%d) overlaps a CPML slab (npml %d)", d->i0, d->j0,
                        d->k0, d->di, d->dj, d->dk, np_);
    }
    return 0;
}

}  // namespace mifdtd

using namespace mifdtd;

extern "C" {

int mifdtd_adjoint_step(int nx, int ny, int nz_pitch, int pad, float *const lam[6], const float *const ce[3],
                        int npml, float *lpsiE, float *lpsiH, const float *bE, const float *aE, const float *bH,
                        const float *aH, const float *eplane, int iplane, const mifdtd_design_box *box,
                        const float *curl, double *grad, mifdtd_stream_t stream)
{
    AGeom g;
    if (int rc = adj_geom(g, nx, ny, nz_pitch, pad, npml)) return rc;
    if (!lam) return fail(MIFDTD_ENULL, "adjoint fields NULL");
    for (int c = 0; c < 6; c++)
        if (!lam[c]) return fail(MIFDTD_ENULL, "adjoint field %d NULL", c);
    const float *c0 = ce ? ce[0] : nullptr, *c1 = ce ? ce[1] : nullptr, *c2 = ce ? ce[2] : nullptr;
    if ((c0 || c1 || c2) && !(c0 && c1 && c2)) return fail(MIFDTD_ENULL, "CE arrays all set or all NULL");
    if (npml && (!lpsiE || !lpsiH || !bE || !aE || !bH || !aH)) return fail(MIFDTD_ENULL, "CPML adjoint buffers NULL");
    if (eplane && (iplane < 0 || iplane >= nx)) return fail(MIFDTD_EINVAL, "monitor plane %d outside the grid", iplane);
    ADesign dz{};
    if (grad) {
        if (int rc = check_box(g, box)) return rc;
        if (!curl) return fail(MIFDTD_ENULL, "curl buffer NULL");
        dz.i0 = box->i0;
        dz.j0 = box->j0;
        dz.k0 = box->k0;
        dz.di = box->di;
        dz.dj = box->dj;
        dz.dk = box->dk;
        dz.curl = curl;
        dz.grad = grad;
    }
    if (g.ncell > 0xffffffffull - AB) return fail(MIFDTD_ERANGE, "grid of %llu cells: at most 2^32 - %d", g.ncell, AB);
    const unsigned grid = (unsigned)((g.ncell + AB - 1) / AB);
    hipStream_t st = to_stream(stream);
    k_adj_a<<<grid, AB, 0, st>>>(g, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5], lpsiH, aH, eplane, iplane, c0, c1,
                                 c2, lpsiE, bE);
    MIFDTD_CHECK_LAUNCH("adjoint pass A");
    k_adj_c<<<grid, AB, 0, st>>>(g, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5], c0, c1, c2, lpsiE, aE, dz, lpsiH,
                                 bH);
    MIFDTD_CHECK_LAUNCH("adjoint pass C");
    return 0;
}

int mifdtd_design_curl(int nx, int ny, int nz_pitch, int pad, const float *hx, const float *hy, const float *hz,
                       const mifdtd_design_box *box, float *out, mifdtd_stream_t stream)
{
    AGeom g;
    if (int rc = adj_geom(g, nx, ny, nz_pitch, pad, 0)) return rc;
    if (int rc = check_box(g, box)) return rc;
    if (!hx || !hy || !hz || !out) return fail(MIFDTD_ENULL, "NULL pointer");
    // neighbours i+1, j+1, k+1 of the box must exist
    if (box->i0 + box->di >= nx || box->j0 + box->dj >= ny || box->k0 + box->dk >= nz_pitch)
        return fail(MIFDTD_EINVAL, "design box touches the upper grid edge");
    ADesign dz{};
    dz.i0 = box->i0;
    dz.j0 = box->j0;
    dz.k0 = box->k0;
    dz.di = box->di;
    dz.dj = box->dj;
    dz.dk = box->dk;
    const unsigned long long s = (unsigned long long)box->di * box->dj * box->dk;
    const unsigned grid = (unsigned)((s + AB - 1) / AB);
    k_design_curl<<<grid, AB, 0, to_stream(stream)>>>(g, dz, hx, hy, hz, out);
    MIFDTD_CHECK_LAUNCH("design curl");
    return 0;
}

}  // extern "C"
