// 4th-order symplectic FDTD: the stage updates of S4/Python-C/dielectric.c (H :5-49,
// E :52-96; the same arithmetic as the CUDA kernels of S4/PyCUDA/100-base.py:10-44) and
// a CPML for it, on the reference's layout: (nx, ny, nz) Fortran order, x contiguous,
// no padding.
//
// A lane per x-quad (4 consecutive x cells of one (j, k) row; nx must be a multiple of
// 4), work-groups marching z (below).  Per-cell arithmetic keeps the reference's
// evaluation order:
//   H: h -= (cl*ch) * (27*(((a0 - a1) - b0) + b1) - (((a2 - a3) - b2) + b3))
//   E: e += (dl*ce) * (27*(((a0 - a1) - b0) + b1) - (((a2 - a3) - b2) + b3))
// so that with -ffp-contract=off the result equals the reference C bit for bit.
// Guards (which cells each stage writes) follow dielectric.c:30-42,77-89 including its
// i < nx-1, j < ny-1, k < nz-1 loop bounds.

#include "common.hpp"

namespace mifdtd {

constexpr int S4B = 256;
constexpr int kS4Rows = 8;     // owned rows per work-group of the z-marching stages (T+3 waves)
constexpr int kS4Planes = 32;  // planes per z chunk

// ------------------------------------------------------------------ z-marching stages
// The stage updates with the neighbour rows and planes read once per work-group:
// a work-group owns one x segment of 64 quads (a lane per quad) of T rows [j0, j0+T) and
// marches the planes of a z chunk [k0, k1).  Wave w holds row j0-1+w (E stage) or
// j0-2+w (H stage), so that the T owned rows have the y halo rows (E: j-1, j+1, j+2;
// H: j-2, j-1, j+1) as waves of the same work-group; per plane each wave publishes the
// two components its neighbours need (E: hz, hx; H: ez, ex) in LDS (double-buffered, one
// barrier per plane).  The z neighbours (E: hy, hx at k-1 .. k+2; H: ey, ex at k-2 ..
// k+1) stay in a four-plane register window per owned lane, and the x neighbours come
// from the adjacent lanes (the segment's edge lanes load theirs).  Arithmetic and
// guards are those of the header above, so the values equal the reference's.
struct S4ZGeom {
    int nx, ny, nz;
    int Qx;   // quads per x row
    int nxs;  // x segments of 64 quads per row
    int T;    // owned rows per work-group
    int nyt;  // row tiles
    int cz;   // planes per z chunk
    int nzc;  // z chunks
};

// x neighbours of this lane's quad q of f (x = i0 .. i0+3): f[i0-2], f[i0-1], f[i0+4],
// f[i0+5], zero outside the row.  Called by every lane of the wave (shuffles).
struct XN {
    float m2, m1, p4, p5;
};

__device__ __forceinline__ XN xneigh(const float *f, size_t id, const Q4 &q, int lane, int qi, int Qx)
{
    XN n;
    n.m2 = __shfl_up(q.v[2], 1);
    n.m1 = __shfl_up(q.v[3], 1);
    n.p4 = __shfl_down(q.v[0], 1);
    n.p5 = __shfl_down(q.v[1], 1);
    if (lane == 0) {
        n.m2 = qi > 0 ? f[id - 2] : 0.f;
        n.m1 = qi > 0 ? f[id - 1] : 0.f;
    }
    if (qi + 1 >= Qx) {
        n.p4 = n.p5 = 0.f;
    } else if (lane == 63) {
        n.p4 = f[id + 4];
        n.p5 = f[id + 5];
    }
    return n;
}

__device__ __forceinline__ Q4 lds4(const float *p) { return Q4{{p[0], p[1], p[2], p[3]}}; }

__device__ __forceinline__ void put4(float *p, const Q4 &q)
{
    *reinterpret_cast<float4 *>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
}

// CE of the E stage: the constant ce0, three float arrays, or three uint8 material-index
// arrays into a 3 x 256 table of the float values (staged in LDS; the same products).
enum S4Ce { kS4CeConst = 0, kS4CeArrays = 1, kS4CeTable = 2 };

template <int CEM>
__device__ __forceinline__ Q4 s4_ce(const float *c, const unsigned char *m, const float *tab, size_t id, float ce0)
{
    if (CEM == kS4CeConst) return splat4(ce0);
    if (CEM == kS4CeArrays) return ld4(c + id);
    const uchar4 t = *reinterpret_cast<const uchar4 *>(m + id);
    return Q4{{tab[t.x], tab[t.y], tab[t.z], tab[t.w]}};
}

template <int CEM>
__global__ __launch_bounds__(1024) void k_s4_e_zm(S4ZGeom g, float dl, float *__restrict__ ex, float *__restrict__ ey,
                                                  float *__restrict__ ez, const float *__restrict__ hx,
                                                  const float *__restrict__ hy, const float *__restrict__ hz,
                                                  const float *__restrict__ cx, const float *__restrict__ cy,
                                                  const float *__restrict__ cz, float ce0,
                                                  const unsigned char *__restrict__ mx,
                                                  const unsigned char *__restrict__ my,
                                                  const unsigned char *__restrict__ mz, const float *__restrict__ lut)
{
    extern __shared__ float zl[];  // [2][T+3][2][256]: hz, hx of plane k, rows j0-1 .. j0+T+1; [768] CE table
    const int T = g.T;
    float *tab = zl + (size_t)2 * (T + 3) * 512;
    if (CEM == kS4CeTable) {
        for (int t = threadIdx.x; t < 768; t += blockDim.x) tab[t] = lut[t];
        __syncthreads();
    }
    unsigned b = blockIdx.x;
    const int xs = (int)(b % (unsigned)g.nxs);
    b /= (unsigned)g.nxs;
    const int yt = (int)(b % (unsigned)g.nyt);
    const int zc = (int)(b / (unsigned)g.nyt);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int j0 = yt * T, k0 = zc * g.cz, k1 = min(nz, k0 + g.cz);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int j = j0 - 1 + w;
    const int qi = xs * 64 + lane;
    const bool qact = qi < g.Qx;
    const int i0 = 4 * (qact ? qi : g.Qx - 1);  // inactive lanes address the row's last quad
    const bool rowok = j >= 0 && j < ny;
    const size_t sz = (size_t)nx * ny;
    const size_t rowoff = (size_t)(rowok ? j : 0) * nx + i0;
    const bool own = w >= 1 && w <= T && j < ny;
    const float c27 = 27.f;
    const Q4 z4 = splat4(0.f);
    Q4 hyw[4], hxw[4];  // planes k-1 .. k+2 of this lane's quad (owned rows)
    if (own) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int kk = k0 - 1 + m;
            const bool in = kk >= 0 && kk < nz;
            hyw[m] = in ? ld4(hy + (size_t)kk * sz + rowoff) : z4;
            hxw[m] = in ? ld4(hx + (size_t)kk * sz + rowoff) : z4;
        }
    }
    for (int k = k0; k < k1; ++k) {
        float *buf = zl + (size_t)((k - k0) & 1) * (T + 3) * 512;
        const size_t idk = (size_t)k * sz + rowoff;
        const Q4 hzk = rowok ? ld4(hz + idk) : z4;
        Q4 hxk, e0, e1, e2, c0, c1, c2;
        if (own) {
            const bool in2 = k + 2 < nz;
            hyw[3] = in2 ? ld4(hy + idk + 2 * sz) : z4;
            hxw[3] = in2 ? ld4(hx + idk + 2 * sz) : z4;
            hxk = hxw[1];
            e0 = ld4(ex + idk);
            e1 = ld4(ey + idk);
            e2 = ld4(ez + idk);
            c0 = s4_ce<CEM>(cx, mx, tab, idk, ce0);
            c1 = s4_ce<CEM>(cy, my, tab + 256, idk, ce0);
            c2 = s4_ce<CEM>(cz, mz, tab + 512, idk, ce0);
        } else {
            hxk = rowok ? ld4(hx + idk) : z4;
        }
        put4(buf + (size_t)(w * 2) * 256 + 4 * lane, hzk);
        put4(buf + (size_t)(w * 2 + 1) * 256 + 4 * lane, hxk);
        XN nz_{}, ny_{};
        if (own) {
            nz_ = xneigh(hz, idk, hzk, lane, qi, g.Qx);
            ny_ = xneigh(hy, idk, hyw[1], lane, qi, g.Qx);
        }
        __syncthreads();
        if (own && qact && j <= ny - 2 && k <= nz - 2) {
            const bool rx = j > 0 && j < ny - 2 && k > 0 && k < nz - 2;
            const bool ry = k > 0 && k < nz - 2;
            const bool rz = j > 0 && j < ny - 2;
            const Q4 hzjm = lds4(buf + (size_t)((w - 1) * 2) * 256 + 4 * lane);
            const Q4 hzjp = lds4(buf + (size_t)((w + 1) * 2) * 256 + 4 * lane);
            const Q4 hzj2 = lds4(buf + (size_t)((w + 2) * 2) * 256 + 4 * lane);
            const Q4 hxjm = lds4(buf + (size_t)((w - 1) * 2 + 1) * 256 + 4 * lane);
            const Q4 hxjp = lds4(buf + (size_t)((w + 1) * 2 + 1) * 256 + 4 * lane);
            const Q4 hxj2 = lds4(buf + (size_t)((w + 2) * 2 + 1) * 256 + 4 * lane);
            // x windows f[i0-2 .. i0+5] of hz and hy at plane k
            const float wz[8] = {nz_.m2, nz_.m1, hzk.v[0], hzk.v[1], hzk.v[2], hzk.v[3], nz_.p4, nz_.p5};
            const float wy[8] = {ny_.m2, ny_.m1, hyw[1].v[0], hyw[1].v[1], hyw[1].v[2], hyw[1].v[3], ny_.p4, ny_.p5};
            const int ib = 4 * qi;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const int i = ib + l;
                if (i > nx - 2) continue;
                if (rx)
                    e0.v[l] += dl * c0.v[l] * (c27 * (hzjp.v[l] - hzk.v[l] - hyw[2].v[l] + hyw[1].v[l])
                                               - (hzj2.v[l] - hzjm.v[l] - hyw[3].v[l] + hyw[0].v[l]));
                if (i > 0 && i < nx - 2 && ry)
                    e1.v[l] += dl * c1.v[l] * (c27 * (hxw[2].v[l] - hxw[1].v[l] - wz[l + 3] + hzk.v[l])
                                               - (hxw[3].v[l] - hxw[0].v[l] - wz[l + 4] + wz[l + 1]));
                if (i > 0 && i < nx - 2 && rz)
                    e2.v[l] += dl * c2.v[l] * (c27 * (wy[l + 3] - hyw[1].v[l] - hxjp.v[l] + hxw[1].v[l])
                                               - (wy[l + 4] - wy[l + 1] - hxj2.v[l] + hxjm.v[l]));
            }
            if (rx) st4(ex + idk, e0);
            if (ry) st4(ey + idk, e1);
            if (rz) st4(ez + idk, e2);
        }
        if (own) {
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                hyw[m] = hyw[m + 1];
                hxw[m] = hxw[m + 1];
            }
        }
    }
}

template <bool HC>
__global__ __launch_bounds__(1024) void k_s4_h_zm(S4ZGeom g, float cl, const float *__restrict__ ex,
                                                  const float *__restrict__ ey, const float *__restrict__ ez,
                                                  float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
                                                  const float *__restrict__ cx, const float *__restrict__ cy,
                                                  const float *__restrict__ cz, float ch0)
{
    extern __shared__ float zl[];  // [2][T+3][2][256]: ez, ex of plane k, rows j0-2 .. j0+T
    const int T = g.T;
    unsigned b = blockIdx.x;
    const int xs = (int)(b % (unsigned)g.nxs);
    b /= (unsigned)g.nxs;
    const int yt = (int)(b % (unsigned)g.nyt);
    const int zc = (int)(b / (unsigned)g.nyt);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int j0 = yt * T, k0 = zc * g.cz, k1 = min(nz, k0 + g.cz);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int j = j0 - 2 + w;
    const int qi = xs * 64 + lane;
    const bool qact = qi < g.Qx;
    const int i0 = 4 * (qact ? qi : g.Qx - 1);
    const bool rowok = j >= 0 && j < ny;
    const size_t sz = (size_t)nx * ny;
    const size_t rowoff = (size_t)(rowok ? j : 0) * nx + i0;
    const bool own = w >= 2 && w <= T + 1 && j < ny;
    const float c27 = 27.f;
    const Q4 z4 = splat4(0.f);
    Q4 eyw[4], exw[4];  // planes k-2 .. k+1 of this lane's quad (owned rows)
    if (own) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int kk = k0 - 2 + m;
            const bool in = kk >= 0 && kk < nz;
            eyw[m] = in ? ld4(ey + (size_t)kk * sz + rowoff) : z4;
            exw[m] = in ? ld4(ex + (size_t)kk * sz + rowoff) : z4;
        }
    }
    for (int k = k0; k < k1; ++k) {
        float *buf = zl + (size_t)((k - k0) & 1) * (T + 3) * 512;
        const size_t idk = (size_t)k * sz + rowoff;
        const Q4 ezk = rowok ? ld4(ez + idk) : z4;
        Q4 exk, h0, h1, h2, c0, c1, c2;
        if (own) {
            const bool in1 = k + 1 < nz;
            eyw[3] = in1 ? ld4(ey + idk + sz) : z4;
            exw[3] = in1 ? ld4(ex + idk + sz) : z4;
            exk = exw[2];
            h0 = ld4(hx + idk);
            h1 = ld4(hy + idk);
            h2 = ld4(hz + idk);
            c0 = HC ? ld4(cx + idk) : splat4(ch0);
            c1 = HC ? ld4(cy + idk) : splat4(ch0);
            c2 = HC ? ld4(cz + idk) : splat4(ch0);
        } else {
            exk = rowok ? ld4(ex + idk) : z4;
        }
        put4(buf + (size_t)(w * 2) * 256 + 4 * lane, ezk);
        put4(buf + (size_t)(w * 2 + 1) * 256 + 4 * lane, exk);
        XN nz_{}, ny_{};
        if (own) {
            nz_ = xneigh(ez, idk, ezk, lane, qi, g.Qx);
            ny_ = xneigh(ey, idk, eyw[2], lane, qi, g.Qx);
        }
        __syncthreads();
        if (own && qact && j <= ny - 2 && k <= nz - 2) {
            const bool rx = j > 1 && k > 1;
            const bool ry = k > 1;
            const bool rz = j > 1;
            const Q4 ezjm = lds4(buf + (size_t)((w - 1) * 2) * 256 + 4 * lane);
            const Q4 ezjp = lds4(buf + (size_t)((w + 1) * 2) * 256 + 4 * lane);
            const Q4 ezj2 = lds4(buf + (size_t)((w - 2) * 2) * 256 + 4 * lane);
            const Q4 exjm = lds4(buf + (size_t)((w - 1) * 2 + 1) * 256 + 4 * lane);
            const Q4 exjp = lds4(buf + (size_t)((w + 1) * 2 + 1) * 256 + 4 * lane);
            const Q4 exj2 = lds4(buf + (size_t)((w - 2) * 2 + 1) * 256 + 4 * lane);
            const float wz[8] = {nz_.m2, nz_.m1, ezk.v[0], ezk.v[1], ezk.v[2], ezk.v[3], nz_.p4, nz_.p5};
            const float wy[8] = {ny_.m2, ny_.m1, eyw[2].v[0], eyw[2].v[1], eyw[2].v[2], eyw[2].v[3], ny_.p4, ny_.p5};
            const int ib = 4 * qi;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const int i = ib + l;
                if (i > nx - 2) continue;
                if (rx && j < ny - 1 && k < nz - 1)
                    h0.v[l] -= cl * c0.v[l] * (c27 * (ezk.v[l] - ezjm.v[l] - eyw[2].v[l] + eyw[1].v[l])
                                               - (ezjp.v[l] - ezj2.v[l] - eyw[3].v[l] + eyw[0].v[l]));
                if (i > 1 && ry)
                    h1.v[l] -= cl * c1.v[l] * (c27 * (exw[2].v[l] - exw[1].v[l] - ezk.v[l] + wz[l + 1])
                                               - (exw[3].v[l] - exw[0].v[l] - wz[l + 3] + wz[l]));
                if (i > 1 && rz)
                    h2.v[l] -= cl * c2.v[l] * (c27 * (eyw[2].v[l] - wy[l + 1] - exw[2].v[l] + exjm.v[l])
                                               - (wy[l + 3] - wy[l] - exjp.v[l] + exj2.v[l]));
            }
            if (rx) st4(hx + idk, h0);
            if (ry) st4(hy + idk, h1);
            if (rz) st4(hz + idk, h2);
        }
        if (own) {
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                eyw[m] = eyw[m + 1];
                exw[m] = exw[m + 1];
            }
        }
    }
}

// ------------------------------------------------------------------ CPML
// Thread per slab cell (both sides of one axis in one grid, side-major).  The face's
// two tangential components and the driving fields are picked by AXIS; see the header
// for the sign table.
__device__ __forceinline__ bool s4_valid_e(int c, int i, int j, int k, int nx, int ny, int nz)
{
    if (i > nx - 2 || j > ny - 2 || k > nz - 2) return false;
    if (c == 0) return j > 0 && j < ny - 2 && k > 0 && k < nz - 2;
    if (c == 1) return i > 0 && i < nx - 2 && k > 0 && k < nz - 2;
    return i > 0 && i < nx - 2 && j > 0 && j < ny - 2;
}
__device__ __forceinline__ bool s4_valid_h(int c, int i, int j, int k, int nx, int ny, int nz)
{
    if (i > nx - 2 || j > ny - 2 || k > nz - 2) return false;
    if (c == 0) return j > 1 && k > 1;
    if (c == 1) return i > 1 && k > 1;
    return i > 1 && j > 1;
}

struct S4Face {
    float *psi1[2], *psi2[2];
};

// Slab cell n of one side -> (i, j, k, psi index), psi layout x fastest:
// x face (npml, ny, nz), y face (nx, npml, nz), z face (nx, ny, npml).
template <int AXIS>
__device__ __forceinline__ void s4_slab_cell(unsigned n, int nx, int ny, int npml, int q, int &i, int &j, int &k)
{
    if (AXIS == 0) {
        const unsigned r = n / (unsigned)npml;  // j + ny*k
        j = (int)(r % (unsigned)ny);
        k = (int)(r / (unsigned)ny);
        i = q;
    } else if (AXIS == 1) {
        i = (int)(n % (unsigned)nx);
        const unsigned r = n / (unsigned)nx;  // p + npml*k
        k = (int)(r / (unsigned)npml);
        j = q;
    } else {
        i = (int)(n % (unsigned)nx);
        j = (int)((n / (unsigned)nx) % (unsigned)ny);
        k = q;
    }
}

template <int AXIS, bool HC>
__global__ __launch_bounds__(S4B) void k_s4_cpml_e(int nx, int ny, int nz, int npml, unsigned per_side, float w,
                                                   float *__restrict__ ex, float *__restrict__ ey, float *__restrict__ ez,
                                                   const float *__restrict__ hx, const float *__restrict__ hy,
                                                   const float *__restrict__ hz, const float *__restrict__ cx,
                                                   const float *__restrict__ cy, const float *__restrict__ cz, float ce0,
                                                   S4Face fs, const float *__restrict__ tb, const float *__restrict__ ta)
{
    unsigned n = blockIdx.x * S4B + threadIdx.x;
    if (n >= 2 * per_side) return;
    const int side = n >= per_side;
    if (side) n -= per_side;
    float *psi1 = fs.psi1[side], *psi2 = fs.psi2[side];
    if (!psi1) return;
    const int n_e = AXIS == 0 ? nx : AXIS == 1 ? ny : nz;
    int p;
    if (AXIS == 0) p = (int)(n % (unsigned)npml);
    else if (AXIS == 1) p = (int)((n / (unsigned)nx) % (unsigned)npml);
    else p = (int)(n / ((unsigned)nx * (unsigned)ny));
    const int q = side ? n_e - 1 - npml + p : p;
    const int t = side ? npml + 1 + p : p + 1;
    int i, j, k;
    s4_slab_cell<AXIS>(n, nx, ny, npml, q, i, j, k);
    const size_t id = ((size_t)k * ny + j) * nx + i;
    const size_t s = AXIS == 0 ? 1 : AXIS == 1 ? (size_t)nx : (size_t)nx * ny;
    // x: psi1 -> ey from hz (-), psi2 -> ez from hy (+); y: psi1 -> ez from hx (-),
    // psi2 -> ex from hz (+); z: psi1 -> ex from hy (-), psi2 -> ey from hx (+)
    float *e1 = AXIS == 0 ? ey : AXIS == 1 ? ez : ex;
    float *e2 = AXIS == 0 ? ez : AXIS == 1 ? ex : ey;
    const float *g1 = AXIS == 0 ? hz : AXIS == 1 ? hx : hy;
    const float *g2 = AXIS == 0 ? hy : AXIS == 1 ? hz : hx;
    const float *c1 = AXIS == 0 ? cy : AXIS == 1 ? cz : cx;
    const float *c2 = AXIS == 0 ? cz : AXIS == 1 ? cx : cy;
    const int comp1 = AXIS == 0 ? 1 : AXIS == 1 ? 2 : 0, comp2 = AXIS == 0 ? 2 : AXIS == 1 ? 0 : 1;
    const float b = tb[t], a = ta[t];
    const float c27 = 27.f;
    if (s4_valid_e(comp1, i, j, k, nx, ny, nz)) {
        const float d = c27 * (g1[id + s] - g1[id]) - (g1[id + 2 * s] - g1[id - s]);
        const float tt = b * psi1[n] + a * d;
        psi1[n] = tt;
        e1[id] -= w * (HC ? c1[id] : ce0) * tt;
    }
    if (s4_valid_e(comp2, i, j, k, nx, ny, nz)) {
        const float d = c27 * (g2[id + s] - g2[id]) - (g2[id + 2 * s] - g2[id - s]);
        const float tt = b * psi2[n] + a * d;
        psi2[n] = tt;
        e2[id] += w * (HC ? c2[id] : ce0) * tt;
    }
}

template <int AXIS, bool HC>
__global__ __launch_bounds__(S4B) void k_s4_cpml_h(int nx, int ny, int nz, int npml, unsigned per_side, float w,
                                                   const float *__restrict__ ex, const float *__restrict__ ey,
                                                   const float *__restrict__ ez, float *__restrict__ hx,
                                                   float *__restrict__ hy, float *__restrict__ hz,
                                                   const float *__restrict__ cx, const float *__restrict__ cy,
                                                   const float *__restrict__ cz, float ch0, S4Face fs,
                                                   const float *__restrict__ tb, const float *__restrict__ ta)
{
    unsigned n = blockIdx.x * S4B + threadIdx.x;
    if (n >= 2 * per_side) return;
    const int side = n >= per_side;
    if (side) n -= per_side;
    float *psi1 = fs.psi1[side], *psi2 = fs.psi2[side];
    if (!psi1) return;
    const int n_e = AXIS == 0 ? nx : AXIS == 1 ? ny : nz;
    int p;
    if (AXIS == 0) p = (int)(n % (unsigned)npml);
    else if (AXIS == 1) p = (int)((n / (unsigned)nx) % (unsigned)npml);
    else p = (int)(n / ((unsigned)nx * (unsigned)ny));
    const int q = side ? n_e - npml + p : p;
    const int t = side ? npml + 2 + p : p + 1;
    int i, j, k;
    s4_slab_cell<AXIS>(n, nx, ny, npml, q, i, j, k);
    const size_t id = ((size_t)k * ny + j) * nx + i;
    const size_t s = AXIS == 0 ? 1 : AXIS == 1 ? (size_t)nx : (size_t)nx * ny;
    // x: psi1 -> hy from ez (+), psi2 -> hz from ey (-); y: psi1 -> hz from ex (+),
    // psi2 -> hx from ez (-); z: psi1 -> hx from ey (+), psi2 -> hy from ex (-)
    float *h1 = AXIS == 0 ? hy : AXIS == 1 ? hz : hx;
    float *h2 = AXIS == 0 ? hz : AXIS == 1 ? hx : hy;
    const float *g1 = AXIS == 0 ? ez : AXIS == 1 ? ex : ey;
    const float *g2 = AXIS == 0 ? ey : AXIS == 1 ? ez : ex;
    const float *c1 = AXIS == 0 ? cy : AXIS == 1 ? cz : cx;
    const float *c2 = AXIS == 0 ? cz : AXIS == 1 ? cx : cy;
    const int comp1 = AXIS == 0 ? 1 : AXIS == 1 ? 2 : 0, comp2 = AXIS == 0 ? 2 : AXIS == 1 ? 0 : 1;
    const float b = tb[t], a = ta[t];
    const float c27 = 27.f;
    if (s4_valid_h(comp1, i, j, k, nx, ny, nz)) {
        const float d = c27 * (g1[id] - g1[id - s]) - (g1[id + s] - g1[id - 2 * s]);
        const float tt = b * psi1[n] + a * d;
        psi1[n] = tt;
        h1[id] += w * (HC ? c1[id] : ch0) * tt;
    }
    if (s4_valid_h(comp2, i, j, k, nx, ny, nz)) {
        const float d = c27 * (g2[id] - g2[id - s]) - (g2[id + s] - g2[id - 2 * s]);
        const float tt = b * psi2[n] + a * d;
        psi2[n] = tt;
        h2[id] -= w * (HC ? c2[id] : ch0) * tt;
    }
}

static int s4_check(int nx, int ny, int nz)
{
    if (nx < 4 || ny < 4 || nz < 4) return fail(MIFDTD_EINVAL, "4th-order grid (%d,%d,%d) smaller than 4", nx, ny, nz);
    if (nx % 4) return fail(MIFDTD_EALIGN, "4th-order nx %d is not a multiple of 4", nx);
    if ((unsigned long long)nx * ny * nz / 4 >= (1ull << 31)) return fail(MIFDTD_ERANGE, "4th-order grid too large");
    return 0;
}

static int s4_zgeom(int nx, int ny, int nz, S4ZGeom &z, unsigned &nblocks)
{
    z.nx = nx;
    z.ny = ny;
    z.nz = nz;
    z.Qx = nx / 4;
    z.nxs = (z.Qx + 63) / 64;
    z.T = kS4Rows;
    z.nyt = (ny + z.T - 1) / z.T;
    z.cz = kS4Planes;
    z.nzc = (nz + z.cz - 1) / z.cz;
    const unsigned long long nb = (unsigned long long)z.nxs * z.nyt * z.nzc;
    if (nb >= (1ull << 31)) return fail(MIFDTD_ERANGE, "4th-order grid too large");
    nblocks = (unsigned)nb;
    return 0;
}

static int s4_core(bool h_half, int nx, int ny, int nz, float w, float *ex, float *ey, float *ez, float *hx, float *hy,
                   float *hz, const float *c0, const float *c1, const float *c2, float cconst, hipStream_t st)
{
    if (int rc = s4_check(nx, ny, nz)) return rc;
    const float *ptrs[6] = {ex, ey, ez, hx, hy, hz};
    for (const float *p : ptrs) {
        if (!p) return fail(MIFDTD_ENULL, "field pointer is NULL");
        if (!aligned16(p)) return fail(MIFDTD_EALIGN, "field pointer not 16-byte aligned");
    }
    const bool hc = c0 || c1 || c2;
    if (hc && !(c0 && c1 && c2)) return fail(MIFDTD_ENULL, "coefficient arrays must be all set or all NULL");
    if (hc && (!aligned16(c0) || !aligned16(c1) || !aligned16(c2)))
        return fail(MIFDTD_EALIGN, "coefficient pointer not 16-byte aligned");
    S4ZGeom z;
    unsigned nb = 0;
    if (int rc = s4_zgeom(nx, ny, nz, z, nb)) return rc;
    const unsigned threads = 64u * (unsigned)(z.T + 3);
    const size_t lds = sizeof(float) * 2 * (size_t)(z.T + 3) * 512;
    if (h_half) {
        if (hc) k_s4_h_zm<true><<<nb, threads, lds, st>>>(z, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst);
        else k_s4_h_zm<false><<<nb, threads, lds, st>>>(z, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst);
        MIFDTD_CHECK_LAUNCH("s4 update_h launch");
    } else {
        if (hc)
            k_s4_e_zm<kS4CeArrays><<<nb, threads, lds, st>>>(z, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst,
                                                                       nullptr, nullptr, nullptr, nullptr);
        else
            k_s4_e_zm<kS4CeConst><<<nb, threads, lds, st>>>(z, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst,
                                                                      nullptr, nullptr, nullptr, nullptr);
        MIFDTD_CHECK_LAUNCH("s4 update_e launch");
    }
    return 0;
}

template <int AXIS>
static int s4_cpml_axis(bool h_half, int nx, int ny, int nz, const mifdtd_cpml_half *cp, float w, float *ex, float *ey,
                        float *ez, float *hx, float *hy, float *hz, const float *c0, const float *c1, const float *c2,
                        float cconst, hipStream_t st)
{
    S4Face fs;
    bool any = false;
    for (int sd = 0; sd < 2; sd++) {
        fs.psi1[sd] = cp->psi1[2 * AXIS + sd];
        fs.psi2[sd] = cp->psi2[2 * AXIS + sd];
        if ((fs.psi1[sd] == nullptr) != (fs.psi2[sd] == nullptr))
            return fail(MIFDTD_ENULL, "psi1/psi2 of face %d must both be set or both NULL", 2 * AXIS + sd);
        any = any || fs.psi1[sd];
    }
    if (!any) return 0;
    const int npml = cp->npml;
    const unsigned per_side = AXIS == 0 ? (unsigned)npml * ny * nz
                            : AXIS == 1 ? (unsigned)nx * npml * nz
                                        : (unsigned)nx * ny * npml;
    const unsigned grid = (2 * per_side + S4B - 1) / S4B;
    const bool hc = c0 != nullptr;
    if (h_half) {
        if (hc) k_s4_cpml_h<AXIS, true><<<grid, S4B, 0, st>>>(nx, ny, nz, npml, per_side, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst, fs, cp->b, cp->a);
        else k_s4_cpml_h<AXIS, false><<<grid, S4B, 0, st>>>(nx, ny, nz, npml, per_side, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst, fs, cp->b, cp->a);
    } else {
        if (hc) k_s4_cpml_e<AXIS, true><<<grid, S4B, 0, st>>>(nx, ny, nz, npml, per_side, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst, fs, cp->b, cp->a);
        else k_s4_cpml_e<AXIS, false><<<grid, S4B, 0, st>>>(nx, ny, nz, npml, per_side, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst, fs, cp->b, cp->a);
    }
    MIFDTD_CHECK_LAUNCH("s4 cpml launch");
    return 0;
}

static int s4_cpml(bool h_half, int nx, int ny, int nz, const mifdtd_cpml_half *cp, float w, float *ex, float *ey,
                   float *ez, float *hx, float *hy, float *hz, const float *c0, const float *c1, const float *c2,
                   float cconst, hipStream_t st)
{
    if (int rc = s4_check(nx, ny, nz)) return rc;
    if (!cp) return fail(MIFDTD_ENULL, "CPML descriptor is NULL");
    const int npml = cp->npml;
    if (npml < 1 || npml > MIFDTD_MAX_NPML) return fail(MIFDTD_EINVAL, "npml %d outside [1,%d]", npml, MIFDTD_MAX_NPML);
    if (nx < 2 * npml + 4 || ny < 2 * npml + 4 || nz < 2 * npml + 4)
        return fail(MIFDTD_EINVAL, "4th-order grid (%d,%d,%d) too small for npml %d", nx, ny, nz, npml);
    if (!cp->b || !cp->a) return fail(MIFDTD_ENULL, "CPML tables are NULL");
    if ((c0 || c1 || c2) && !(c0 && c1 && c2)) return fail(MIFDTD_ENULL, "coefficient arrays must be all set or all NULL");
    if (int rc = s4_cpml_axis<0>(h_half, nx, ny, nz, cp, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst, st)) return rc;
    if (int rc = s4_cpml_axis<1>(h_half, nx, ny, nz, cp, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst, st)) return rc;
    return s4_cpml_axis<2>(h_half, nx, ny, nz, cp, w, ex, ey, ez, hx, hy, hz, c0, c1, c2, cconst, st);
}

}  // namespace mifdtd

using namespace mifdtd;

extern "C" {

void s4_update_h(int nx, int ny, int nz, float cl, float *ex, float *ey, float *ez, float *hx, float *hy, float *hz,
                 float *cx, float *cy, float *cz)
{
    if (!cx || !cy || !cz) {
        fail(MIFDTD_ENULL, "s4_update_h needs the three coefficient arrays");
        return;
    }
    s4_core(true, nx, ny, nz, cl, ex, ey, ez, hx, hy, hz, cx, cy, cz, 0.f, nullptr);
}

void s4_update_e(int nx, int ny, int nz, float dl, float *ex, float *ey, float *ez, float *hx, float *hy, float *hz,
                 float *cx, float *cy, float *cz)
{
    if (!cx || !cy || !cz) {
        fail(MIFDTD_ENULL, "s4_update_e needs the three coefficient arrays");
        return;
    }
    s4_core(false, nx, ny, nz, dl, ex, ey, ez, hx, hy, hz, cx, cy, cz, 0.f, nullptr);
}

int mifdtd_s4_update_h(int nx, int ny, int nz, float cl, const float *ex, const float *ey, const float *ez, float *hx,
                       float *hy, float *hz, const float *chx, const float *chy, const float *chz, float ch0,
                       mifdtd_stream_t stream)
{
    return s4_core(true, nx, ny, nz, cl, const_cast<float *>(ex), const_cast<float *>(ey), const_cast<float *>(ez), hx,
                   hy, hz, chx, chy, chz, ch0, to_stream(stream));
}

int mifdtd_s4_update_e(int nx, int ny, int nz, float dl, float *ex, float *ey, float *ez, const float *hx,
                       const float *hy, const float *hz, const float *cex, const float *cey, const float *cez, float ce0,
                       mifdtd_stream_t stream)
{
    return s4_core(false, nx, ny, nz, dl, ex, ey, ez, const_cast<float *>(hx), const_cast<float *>(hy),
                   const_cast<float *>(hz), cex, cey, cez, ce0, to_stream(stream));
}

int mifdtd_s4_update_e_lut(int nx, int ny, int nz, float dl, float *ex, float *ey, float *ez, const float *hx,
                           const float *hy, const float *hz, const unsigned char *mat_x, const unsigned char *mat_y,
                           const unsigned char *mat_z, const float *lut, mifdtd_stream_t stream)
{
    if (int rc = s4_check(nx, ny, nz)) return rc;
    const float *ptrs[6] = {ex, ey, ez, hx, hy, hz};
    for (const float *p : ptrs) {
        if (!p) return fail(MIFDTD_ENULL, "field pointer is NULL");
        if (!aligned16(p)) return fail(MIFDTD_EALIGN, "field pointer not 16-byte aligned");
    }
    if (!mat_x || !mat_y || !mat_z || !lut) return fail(MIFDTD_ENULL, "material index or table is NULL");
    const unsigned char *mats[3] = {mat_x, mat_y, mat_z};
    for (const unsigned char *m : mats)
        if (reinterpret_cast<uintptr_t>(m) & 3u) return fail(MIFDTD_EALIGN, "material index array not 4-byte aligned");
    S4ZGeom z;
    unsigned nb = 0;
    if (int rc = s4_zgeom(nx, ny, nz, z, nb)) return rc;
    const unsigned threads = 64u * (unsigned)(z.T + 3);
    const size_t lds = sizeof(float) * (2 * (size_t)(z.T + 3) * 512 + 768);
    k_s4_e_zm<kS4CeTable><<<nb, threads, lds, to_stream(stream)>>>(z, dl, ex, ey, ez, hx, hy, hz, nullptr, nullptr,
                                                                   nullptr, 0.f, mat_x, mat_y, mat_z, lut);
    MIFDTD_CHECK_LAUNCH("s4 update_e (CE table) launch");
    return 0;
}

int mifdtd_s4_cpml_e(int nx, int ny, int nz, const mifdtd_cpml_half *cp, float dl, float *ex, float *ey, float *ez,
                     const float *hx, const float *hy, const float *hz, const float *cex, const float *cey,
                     const float *cez, float ce0, mifdtd_stream_t stream)
{
    return s4_cpml(false, nx, ny, nz, cp, dl, ex, ey, ez, const_cast<float *>(hx), const_cast<float *>(hy),
                   const_cast<float *>(hz), cex, cey, cez, ce0, to_stream(stream));
}

int mifdtd_s4_cpml_h(int nx, int ny, int nz, const mifdtd_cpml_half *cp, float cl, const float *ex, const float *ey,
                     const float *ez, float *hx, float *hy, float *hz, const float *chx, const float *chy,
                     const float *chz, float ch0, mifdtd_stream_t stream)
{
    return s4_cpml(true, nx, ny, nz, cp, cl, const_cast<float *>(ex), const_cast<float *>(ey), const_cast<float *>(ez),
                   hx, hy, hz, chx, chy, chz, ch0, to_stream(stream));
}

}  // extern "C"
