// Fused single-pass Yee step: core E + CPML-E + E sources + core H + CPML-H + H
// sources + probes of one time step in ONE sweep over the grid, reading the state of
// step n from one buffer set and writing step n+1 to another (ping-pong), so each
// field, coefficient and psi array is read from HBM once and written once per step:
// 60 B/cell (6 fields + 3 CE read, 6 fields written) + 32 B per CPML slab cell.
//
// Mapping (gfx950, 64-lane waves; no barriers, every wave independent):
//  * a wave owns T consecutive y rows of an x chunk of CX planes and streams them
//    plane by plane (x), z segment by z segment (256 floats = 64 lanes x one 16-byte
//    quad), row by row (y): for plane i, segment m it first recomputes E_new of the
//    row below its block (y halo, not stored), then for each own row j computes
//    E_new(i,j) from H(i,j), H(i,j+1) (carried / loaded) and H(i+1,j), and H_new(i,j)
//    from E_new(i,j), E_new(i,j-1) (carried) and E_new(i-1,j);
//  * E_new(i-1,j) is re-read from the output buffer: this wave wrote it one plane
//    earlier (an L2 hit), or, on a chunk's first plane, a small seed pass launched
//    just before wrote it (the owner chunk writes the same values again);
//  * z neighbours: H at k+1 and E_new at k-1 come from the adjacent lane by
//    cross-lane shuffles; across segments the k-1 value is carried from lane 63 and
//    the k+1 value of lane 63 is a scalar load (an L1/L2 hit);
//  * block ids are remapped so that consecutive row blocks of a chunk run on the same
//    XCD, sharing L2 for the halo rows.
// With ping-pong buffers nothing a wave reads is written by any other wave in the
// same launch.  Per-cell arithmetic order equals the in-place path (core, x, y, z
// CPML, sources) under -ffp-contract=off: values are bit-identical to it; only the
// sign of zeros stored in masked lanes (value*0 in the in-place path) may differ.

#include "common.hpp"

namespace mifdtd {

constexpr int kMaxRows = 7;      // owned rows per block (a block has rows+1 waves)
constexpr int kMaxSrc = 4;
constexpr int kMaxProbe = 8;

// psi of all faces packed in one buffer per half (layout of include/mifdtd.h):
// face f = 2*axis+side, psi1 then psi2, x faces (npml,ny,nzp), y (nx,npml,nzp), z (nx,ny,npml)
struct FCpml {
    int npml;
    const float *bE, *aE, *bH, *aH;
    const float *eIn;   // psi E of step n
    float *eOut;        // psi E of step n+1
    float *h;           // psi H (in place)
    size_t sX, sY, sZ;  // face sizes
    __device__ __forceinline__ size_t off(int f, int c) const  // c = 0 (psi1) or 1 (psi2)
    {
        const int ax = f >> 1, sd = f & 1;
        if (ax == 0) return (size_t)(2 * sd + c) * sX;
        if (ax == 1) return 4 * sX + (size_t)(2 * sd + c) * sY;
        return 4 * sX + 4 * sY + (size_t)(2 * sd + c) * sZ;
    }
};

struct FSrc {
    int n;
    int comp[kMaxSrc], i0[kMaxSrc], i1[kMaxSrc], j0[kMaxSrc], j1[kMaxSrc], k0[kMaxSrc], k1[kMaxSrc], ow[kMaxSrc];
    float tval[kMaxSrc], svs[kMaxSrc];
    const float *sv[kMaxSrc];
};

struct FProbe {
    int n;
    int comp[kMaxProbe], i[kMaxProbe], j[kMaxProbe], k[kMaxProbe];
    float *out;
};

struct FGeom {
    int nx, ny, nzp, nz;
    int T;        // rows per wave
    int nrb;      // row blocks (ceil(ny/T))
    int cx;       // planes per chunk
    int nchunks;
    int nseg;     // z segments of 256 floats
    int nwaves;
    int nblocks;
};

struct Ctx {  // buffers of one launch (restrict-qualified at the kernel boundary)
    const float *ex, *ey, *ez, *hx, *hy, *hz;   // step n
    float *oex, *oey, *oez, *ohx, *ohy, *ohz;   // step n+1
    const float *cex, *cey, *cez;
};

// slab position helpers: table index (>0), side and plane p of coordinate q, or -1
// outside the slabs.  Both ends are checked: along z, q runs into the pad lanes.
__device__ __forceinline__ int slab_e(int q, int n_e, int npml, int &side, int &p)
{
    side = 0;
    p = 0;
    if (q < npml) { p = q; return p + 1; }
    if (q >= n_e - 1 - npml && q < n_e - 1) { side = 1; p = q - (n_e - 1 - npml); return npml + 1 + p; }
    return -1;
}
__device__ __forceinline__ int slab_h(int q, int n_e, int npml, int &side, int &p)
{
    side = 0;
    p = 0;
    if (q < npml) { p = q; return p + 1; }
    if (q >= n_e - npml && q < n_e) { side = 1; p = q - (n_e - npml); return npml + 2 + p; }
    return -1;
}

__device__ __forceinline__ void apply_src(const FSrc &s, int c, int i, int j, int k0, Q4 &v)
{
#pragma unroll
    for (int n = 0; n < kMaxSrc; ++n) {
        if (n < s.n && s.comp[n] == c && i >= s.i0[n] && i <= s.i1[n] && j >= s.j0[n] && j <= s.j1[n]) {
            const int lj = s.j1[n] - s.j0[n] + 1, lk = s.k1[n] - s.k0[n] + 1;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const int k = k0 + l;
                if (k >= s.k0[n] && k <= s.k1[n]) {
                    const float w = s.sv[n] ? s.sv[n][((size_t)(i - s.i0[n]) * lj + (j - s.j0[n])) * lk + (k - s.k0[n])]
                                            : s.svs[n];
                    const float add = s.tval[n] * w;
                    v.v[l] = s.ow[n] ? add : v.v[l] + add;
                }
            }
        }
    }
}

__device__ __forceinline__ Q4 zero4() { return splat4(0.f); }

// Addressing: within a launch everything but the lane is wave-uniform, so each access
// is a uniform (scalar) row base -- array + (i*ny + j)*nzp + 256*m -- plus the lane's
// constant 4*lane offset.
__device__ __forceinline__ size_t row_base(const FGeom &g, int i, int j, int m)
{
    return ((size_t)i * g.ny + j) * g.nzp + (size_t)m * 256;
}

// element 0 of the next lane's quad (k+4); lane 63 reads the next segment's first
// element (same row) from memory when that segment exists
__device__ __forceinline__ float next_el0(const Q4 &q, const float *row, bool last_lane, bool has_next_seg)
{
    float v = __shfl_down(q.v[0], 1);
    if (last_lane) v = has_next_seg ? row[256] : 0.f;
    return v;
}

__device__ __forceinline__ void load_H(const Ctx &c, const FGeom &g, int i, int j, int m, int l4, bool act, Q4 h[3])
{
    if (i < g.nx && j >= 0 && j < g.ny) {
        const size_t rb = row_base(g, i, j, m);
        h[0] = act ? ld4(c.hx + rb + l4) : zero4();
        h[1] = act ? ld4(c.hy + rb + l4) : zero4();
        h[2] = act ? ld4(c.hz + rb + l4) : zero4();
    } else {
        h[0] = h[1] = h[2] = zero4();
    }
}

// E_new of row j (0 <= j < ny) of plane i for this lane's quad (segment m, lane offset
// l4 = 4*lane, k0 = 256*m + l4).  hc = H(i,j), hu = H(i,j+1).  Loads H(i+1,j), E(i,j),
// CE(i,j) and psi-E; stores psi-E of step n+1 when `store`.  wantY = false skips ey.
template <bool HC, bool PML>
__device__ __forceinline__ void compute_E(const FGeom &g, const Ctx &c, const FCpml &cp, const FSrc &src,
                                          int i, int j, int m, int l4, bool act, bool last_lane, bool has_next_seg,
                                          const Q4 hc[3], const Q4 hu[3], bool wantY, bool store,
                                          Q4 &exn, Q4 &eyn, Q4 &ezn)
{
    const int nx = g.nx, ny = g.ny, nzp = g.nzp, nz = g.nz;
    const int k0 = m * 256 + l4;
    const size_t rb = row_base(g, i, j, m);
    const size_t sx = (size_t)ny * nzp;
    const int npml = PML ? cp.npml : 0;
    // k+1 neighbours of H(i,j)
    const float hy1 = next_el0(hc[1], c.hy + rb, last_lane, has_next_seg);
    const float hx1 = next_el0(hc[0], c.hx + rb, last_lane, has_next_seg);
    const bool zslab = PML && (k0 < npml || k0 + 3 >= nz - 1 - npml);
    // ---------------------------------------------------------------- ex
    {
        Q4 e = act ? ld4(c.ex + rb + l4) : zero4();
        if (act && j < ny - 1) {
            const Q4 cc = HC ? ld4(c.cex + rb + l4) : splat4(0.5f);
            const Q4 &h1 = hu[2], &h0 = hc[2], &y0 = hc[1];
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float y1 = l < 3 ? y0.v[l + 1] : hy1;
                const float v = e.v[l] + cc.v[l] * ((h1.v[l] - h0.v[l]) - (y1 - y0.v[l]));
                e.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
            }
            if (PML) {
                int ys, yp;
                const int yt = slab_e(j, ny, npml, ys, yp);
                if (yt > 0) {  // y face psi2: ex += cex*psi, from hz(j+1)-hz(j)
                    const float b = cp.bE[yt], a = cp.aE[yt];
                    const size_t pb = cp.off(2 + ys, 1) + ((size_t)i * npml + yp) * nzp + (size_t)m * 256;
                    const Q4 ps = ld4(cp.eIn + pb + l4);
                    Q4 np = ps;
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        if (k0 + l < nz - 1) {
                            const float tt = b * ps.v[l] + a * (h1.v[l] - h0.v[l]);
                            np.v[l] = tt;
                            e.v[l] = e.v[l] + cc.v[l] * tt;
                        }
                    }
                    if (store) st4(cp.eOut + pb + l4, np);
                }
                if (zslab) {
                    const size_t zb = ((size_t)i * ny + j) * npml;
#pragma unroll
                    for (int l = 0; l < 4; ++l) {  // z face psi1: ex -= cex*psi, from hy(k+1)-hy(k)
                        int zs, zp;
                        const int zt = slab_e(k0 + l, nz, npml, zs, zp);
                        if (zt > 0) {
                            const size_t pid = cp.off(4 + zs, 0) + zb + zp;
                            const float y1 = l < 3 ? y0.v[l + 1] : hy1;
                            const float tt = cp.bE[zt] * cp.eIn[pid] + cp.aE[zt] * (y1 - y0.v[l]);
                            if (store) cp.eOut[pid] = tt;
                            e.v[l] = e.v[l] - cc.v[l] * tt;
                        }
                    }
                }
            }
        }
        if (act) apply_src(src, 0, i, j, k0, e);
        exn = e;
    }
    // ------------------------------------------------ H(i+1,j) for ey, ez
    const bool nxt = act && i + 1 < nx;
    const Q4 hyN = nxt ? ld4(c.hy + rb + sx + l4) : zero4();
    const Q4 hzN = (nxt && wantY) ? ld4(c.hz + rb + sx + l4) : zero4();
    // ---------------------------------------------------------------- ey
    if (wantY) {
        Q4 e = act ? ld4(c.ey + rb + l4) : zero4();
        if (act && i < nx - 1) {
            const Q4 cc = HC ? ld4(c.cey + rb + l4) : splat4(0.5f);
            const Q4 &x0 = hc[0], &z1 = hzN, &z0 = hc[2];
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float x1 = l < 3 ? x0.v[l + 1] : hx1;
                const float v = e.v[l] + cc.v[l] * ((x1 - x0.v[l]) - (z1.v[l] - z0.v[l]));
                e.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
            }
            if (PML) {
                int xs, xp;
                const int xt = slab_e(i, nx, npml, xs, xp);
                if (xt > 0) {  // x face psi1: ey -= cey*psi, from hz(i+1)-hz(i)
                    const float b = cp.bE[xt], a = cp.aE[xt];
                    const size_t pb = cp.off(xs, 0) + ((size_t)xp * ny + j) * nzp + (size_t)m * 256;
                    const Q4 ps = ld4(cp.eIn + pb + l4);
                    Q4 np = ps;
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        if (k0 + l < nz - 1) {
                            const float tt = b * ps.v[l] + a * (z1.v[l] - z0.v[l]);
                            np.v[l] = tt;
                            e.v[l] = e.v[l] - cc.v[l] * tt;
                        }
                    }
                    if (store) st4(cp.eOut + pb + l4, np);
                }
                if (zslab) {
                    const size_t zb = ((size_t)i * ny + j) * npml;
#pragma unroll
                    for (int l = 0; l < 4; ++l) {  // z face psi2: ey += cey*psi, from hx(k+1)-hx(k)
                        int zs, zp;
                        const int zt = slab_e(k0 + l, nz, npml, zs, zp);
                        if (zt > 0) {
                            const size_t pid = cp.off(4 + zs, 1) + zb + zp;
                            const float x1 = l < 3 ? x0.v[l + 1] : hx1;
                            const float tt = cp.bE[zt] * cp.eIn[pid] + cp.aE[zt] * (x1 - x0.v[l]);
                            if (store) cp.eOut[pid] = tt;
                            e.v[l] = e.v[l] + cc.v[l] * tt;
                        }
                    }
                }
            }
        }
        if (act) apply_src(src, 1, i, j, k0, e);
        eyn = e;
    } else {
        eyn = zero4();
    }
    // ---------------------------------------------------------------- ez
    {
        Q4 e = act ? ld4(c.ez + rb + l4) : zero4();
        if (act && i < nx - 1 && j < ny - 1) {
            const Q4 cc = HC ? ld4(c.cez + rb + l4) : splat4(0.5f);
            const Q4 &y1 = hyN, &y0 = hc[1], &x1 = hu[0], &x0 = hc[0];
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float v = e.v[l] + cc.v[l] * ((y1.v[l] - y0.v[l]) - (x1.v[l] - x0.v[l]));
                e.v[l] = (k0 + l < nz) ? v : v * 0.f;
            }
            if (PML) {
                int xs, xp;
                const int xt = slab_e(i, nx, npml, xs, xp);
                if (xt > 0) {  // x face psi2: ez += cez*psi, from hy(i+1)-hy(i)
                    const float b = cp.bE[xt], a = cp.aE[xt];
                    const size_t pb = cp.off(xs, 1) + ((size_t)xp * ny + j) * nzp + (size_t)m * 256;
                    const Q4 ps = ld4(cp.eIn + pb + l4);
                    Q4 np = ps;
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        if (k0 + l < nz) {
                            const float tt = b * ps.v[l] + a * (y1.v[l] - y0.v[l]);
                            np.v[l] = tt;
                            e.v[l] = e.v[l] + cc.v[l] * tt;
                        }
                    }
                    if (store) st4(cp.eOut + pb + l4, np);
                }
                int ys, yp;
                const int yt = slab_e(j, ny, npml, ys, yp);
                if (yt > 0) {  // y face psi1: ez -= cez*psi, from hx(j+1)-hx(j)
                    const float b = cp.bE[yt], a = cp.aE[yt];
                    const size_t pb = cp.off(2 + ys, 0) + ((size_t)i * npml + yp) * nzp + (size_t)m * 256;
                    const Q4 ps = ld4(cp.eIn + pb + l4);
                    Q4 np = ps;
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        if (k0 + l < nz) {
                            const float tt = b * ps.v[l] + a * (x1.v[l] - x0.v[l]);
                            np.v[l] = tt;
                            e.v[l] = e.v[l] - cc.v[l] * tt;
                        }
                    }
                    if (store) st4(cp.eOut + pb + l4, np);
                }
            }
        }
        if (act) apply_src(src, 2, i, j, k0, e);
        ezn = e;
    }
}

// H_new of row j (0 <= j < ny) of plane i for this lane's quad: hc = H(i,j) of step n,
// exn/eyn/ezn = E_new(i,j), exb/ezb = E_new(i,j-1) (ex, ez), eyp/ezp = E_new(i-1,j)
// (ey, ez), exm/eym = E_new(i,j) at k0-1 (ex, ey).  Updates psi-H in place.
template <bool PML>
__device__ __forceinline__ void compute_H(const FGeom &g, const FCpml &cp, const FSrc &src, int i, int j, int m,
                                          int l4, bool act, const Q4 hc[3], const Q4 &exn, const Q4 &eyn,
                                          const Q4 &ezn, const Q4 &exb, const Q4 &ezb, const Q4 &eyp, const Q4 &ezp,
                                          float exm, float eym, Q4 &hxo, Q4 &hyo, Q4 &hzo)
{
    const int nx = g.nx, ny = g.ny, nzp = g.nzp, nz = g.nz;
    const int k0 = m * 256 + l4;
    const int npml = PML ? cp.npml : 0;
    int hxs = 0, hxp = 0;
    const int hxt = PML ? slab_h(i, nx, npml, hxs, hxp) : -1;
    (void)hxs;
    (void)hxp;
        int ys = 0, yp = 0;
        const int yt = PML ? slab_h(j, ny, npml, ys, yp) : -1;
        const bool zslab = PML && (k0 < npml || k0 + 3 >= nz - npml);
        const size_t zb = PML ? ((size_t)i * ny + j) * npml : 0;
        // hx: j>0 ; -= 0.5*((ez - ez[j-1]) - (ey - ey[k-1]))
        hxo = hc[0];
        if (act && j > 0) {
            const Q4 &z0 = ezn, &zm = ezb, &y0 = eyn;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float ymv = l > 0 ? y0.v[l - 1] : eym;
                const float v = hxo.v[l] - 0.5f * ((z0.v[l] - zm.v[l]) - (y0.v[l] - ymv));
                const int k = k0 + l;
                hxo.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
            }
            if (PML) {
                if (yt > 0) {  // y face psi2: hx -= 0.5*psi, from ez(j)-ez(j-1)
                    const float b = cp.bH[yt], a = cp.aH[yt];
                    float *ps = cp.h + cp.off(2 + ys, 1) + ((size_t)i * npml + yp) * nzp + (size_t)m * 256 + l4;
                    Q4 p4 = ld4(ps);
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        const int k = k0 + l;
                        if (k > 0 && k < nz) {
                            const float tt = b * p4.v[l] + a * (z0.v[l] - zm.v[l]);
                            p4.v[l] = tt;
                            hxo.v[l] = hxo.v[l] - 0.5f * tt;
                        }
                    }
                    st4(ps, p4);
                }
                if (zslab) {
#pragma unroll
                    for (int l = 0; l < 4; ++l) {  // z face psi1: hx += 0.5*psi, from ey(k)-ey(k-1)
                        const int k = k0 + l;
                        int zs, zp;
                        const int zt = slab_h(k, nz, npml, zs, zp);
                        if (zt > 0 && k > 0) {
                            float *ps = cp.h + cp.off(4 + zs, 0) + zb + zp;
                            const float ymv = l > 0 ? y0.v[l - 1] : eym;
                            const float tt = cp.bH[zt] * *ps + cp.aH[zt] * (y0.v[l] - ymv);
                            *ps = tt;
                            hxo.v[l] = hxo.v[l] + 0.5f * tt;
                        }
                    }
                }
            }
        }
        if (act) apply_src(src, 3, i, j, k0, hxo);
        // hy: i>0 ; -= 0.5*((ex - ex[k-1]) - (ez - ez[i-1]))
        hyo = hc[1];
        if (act && i > 0) {
            const Q4 &x0 = exn, &z0 = ezn, &zp_ = ezp;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float xmv = l > 0 ? x0.v[l - 1] : exm;
                const float v = hyo.v[l] - 0.5f * ((x0.v[l] - xmv) - (z0.v[l] - zp_.v[l]));
                const int k = k0 + l;
                hyo.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
            }
            if (PML) {
                if (hxt > 0) {  // x face psi1: hy += 0.5*psi, from ez(i)-ez(i-1)
                    const float b = cp.bH[hxt], a = cp.aH[hxt];
                    float *ps = cp.h + cp.off(hxs, 0) + ((size_t)hxp * ny + j) * nzp + (size_t)m * 256 + l4;
                    Q4 p4 = ld4(ps);
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        const int k = k0 + l;
                        if (k > 0 && k < nz) {
                            const float tt = b * p4.v[l] + a * (z0.v[l] - zp_.v[l]);
                            p4.v[l] = tt;
                            hyo.v[l] = hyo.v[l] + 0.5f * tt;
                        }
                    }
                    st4(ps, p4);
                }
                if (zslab) {
#pragma unroll
                    for (int l = 0; l < 4; ++l) {  // z face psi2: hy -= 0.5*psi, from ex(k)-ex(k-1)
                        const int k = k0 + l;
                        int zs, zp;
                        const int zt = slab_h(k, nz, npml, zs, zp);
                        if (zt > 0 && k > 0) {
                            float *ps = cp.h + cp.off(4 + zs, 1) + zb + zp;
                            const float xmv = l > 0 ? x0.v[l - 1] : exm;
                            const float tt = cp.bH[zt] * *ps + cp.aH[zt] * (x0.v[l] - xmv);
                            *ps = tt;
                            hyo.v[l] = hyo.v[l] - 0.5f * tt;
                        }
                    }
                }
            }
        }
        if (act) apply_src(src, 4, i, j, k0, hyo);
        // hz: i>0 && j>0 ; -= 0.5*((ey - ey[i-1]) - (ex - ex[j-1]))
        hzo = hc[2];
        if (act && i > 0 && j > 0) {
            const Q4 &y0 = eyn, &yp_ = eyp, &x0 = exn, &xm = exb;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float v = hzo.v[l] - 0.5f * ((y0.v[l] - yp_.v[l]) - (x0.v[l] - xm.v[l]));
                hzo.v[l] = (k0 + l < nz) ? v : v * 0.f;
            }
            if (PML) {
                if (hxt > 0) {  // x face psi2: hz -= 0.5*psi, from ey(i)-ey(i-1)
                    const float b = cp.bH[hxt], a = cp.aH[hxt];
                    float *ps = cp.h + cp.off(hxs, 1) + ((size_t)hxp * ny + j) * nzp + (size_t)m * 256 + l4;
                    Q4 p4 = ld4(ps);
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        if (k0 + l < nz) {
                            const float tt = b * p4.v[l] + a * (y0.v[l] - yp_.v[l]);
                            p4.v[l] = tt;
                            hzo.v[l] = hzo.v[l] - 0.5f * tt;
                        }
                    }
                    st4(ps, p4);
                }
                if (yt > 0) {  // y face psi1: hz += 0.5*psi, from ex(j)-ex(j-1)
                    const float b = cp.bH[yt], a = cp.aH[yt];
                    float *ps = cp.h + cp.off(2 + ys, 0) + ((size_t)i * npml + yp) * nzp + (size_t)m * 256 + l4;
                    Q4 p4 = ld4(ps);
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        if (k0 + l < nz) {
                            const float tt = b * p4.v[l] + a * (x0.v[l] - xm.v[l]);
                            p4.v[l] = tt;
                            hzo.v[l] = hzo.v[l] + 0.5f * tt;
                        }
                    }
                    st4(ps, p4);
                }
            }
        }
        if (act) apply_src(src, 5, i, j, k0, hzo);
}

// One step as one sweep (single pass, ping-pong buffers).  A block of T+1 waves owns
// rows j0 .. j0+T-1 of an x chunk [i0, i1): wave w takes row j0-1+w (w = 0 is the y
// halo row below the block, whose E_new only feeds the block's first row).  For each
// z segment m (256 floats: 64 lanes x one quad) the block marches the planes of the
// chunk, starting with the seed plane i0-1 (E_new only, kept in registers as
// E_new(i-1)):
//   1. every wave computes E_new(i, j) of its row (core + CPML-E + E sources);
//   2. it publishes ex, ez of its row in LDS (double-buffered: the buffer alternates
//      with every exchanged plane, across segments too) for the row above, and lane 63 leaves ex, ey at k = 256m+255 for lane 0 of segment m+1
//      (per plane and row, in LDS);
//   3. one barrier;
//   4. the owned rows compute H_new(i, j) (core + CPML-H + H sources) from E_new(i, j),
//      E_new(i, j-1) (LDS), E_new(i-1, j) (registers) and the k-1 values (lane shuffles,
//      lane 0 from the carry) and store E_new, H_new and the probes.
// H(i, j), H(i, j+1) and H(i+1, j) are read from the step-n buffers (the rows a block
// shares between waves and planes are L2 hits).  Nothing a wave reads is written in
// the same launch except the LDS exchange, which the barrier and the parity orders.
template <bool HC, bool PML>
__global__ __launch_bounds__(512) void k_fused(
    FGeom g,
    const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    float *__restrict__ oex, float *__restrict__ oey, float *__restrict__ oez,
    float *__restrict__ ohx, float *__restrict__ ohy, float *__restrict__ ohz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    const float *__restrict__ psiE_in, float *__restrict__ psiE_out, float *__restrict__ psiH,
    FCpml cp, FSrc src, FProbe prb)
{
    const Ctx c{ex, ey, ez, hx, hy, hz, oex, oey, oez, ohx, ohy, ohz, cex, cey, cez};
    cp.eIn = psiE_in;
    cp.eOut = psiE_out;
    cp.h = psiH;
    extern __shared__ float lds[];
    const int T = g.T;
    float *xbuf = lds;                              // [2][T+1][2][256]: ex, ez of each row
    float *cbuf = lds + 2 * (T + 1) * 2 * 256;      // [cx+1][T+1][2]: ex, ey at k = 256m+255

    // XCD-aware block order: consecutive logical blocks (neighbouring row blocks of a
    // chunk, which share halo rows) land on the same XCD.
    int bid = blockIdx.x;
    {
        const int nb = g.nblocks, nb8 = nb & ~7;
        if (bid < nb8) bid = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    }
    const int rb = bid % g.nrb;
    const int chunk = bid / g.nrb;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int l4 = 4 * lane;
    const bool last_lane = lane == 63;
    const int j0 = rb * T;
    const int j = j0 - 1 + w;
    const bool live = j >= 0 && j < g.ny;
    const bool own = w >= 1 && live;
    const int i0 = chunk * g.cx;
    const int i1 = min(g.nx, i0 + g.cx);
    const int ib = i0 > 0 ? i0 - 1 : i0;

    int it = 0;  // exchanged planes so far: the LDS parity alternates across segments too
    for (int m = 0; m < g.nseg; ++m) {
        const int k0 = m * 256 + l4;
        const bool act = k0 < g.nzp;
        const bool has_next_seg = (m + 1) * 256 < g.nzp;
        Q4 eyp = zero4(), ezp = zero4();  // E_new(i-1, j)
        for (int i = ib; i < i1; ++i) {
            const bool seed = i < i0;
            Q4 hc[3] = {zero4(), zero4(), zero4()};
            Q4 exn = zero4(), eyn = zero4(), ezn = zero4();
            if (live && (own || !seed)) {
                Q4 hu[3];
                load_H(c, g, i, j, m, l4, act, hc);
                load_H(c, g, i, j + 1, m, l4, act, hu);
                compute_E<HC, PML>(g, c, cp, src, i, j, m, l4, act, last_lane, has_next_seg, hc, hu, own,
                                   own && !seed, exn, eyn, ezn);
            }
            if (seed) {  // uniform over the block: no exchange, no barrier
                eyp = eyn;
                ezp = ezn;
                continue;
            }
            const int par = it++ & 1;
            float *xs = xbuf + (size_t)((par * (T + 1) + w) * 2) * 256;
            if (act) {
                *reinterpret_cast<float4 *>(xs + l4) = make_float4(exn.v[0], exn.v[1], exn.v[2], exn.v[3]);
                *reinterpret_cast<float4 *>(xs + 256 + l4) = make_float4(ezn.v[0], ezn.v[1], ezn.v[2], ezn.v[3]);
            }
            // k-1 neighbours: lanes 1..63 from the lane below; lane 0 from lane 63 of the
            // previous segment at this plane (this wave wrote it there); segment 0: k-1 < 0
            // feeds masked lanes only
            float *cs = cbuf + (size_t)(((i - ib) * (T + 1) + w) * 2);
            float exm = __shfl_up(exn.v[3], 1);
            float eym = __shfl_up(eyn.v[3], 1);
            if (lane == 0) {
                exm = m > 0 ? cs[0] : 0.f;
                eym = m > 0 ? cs[1] : 0.f;
            }
            __syncthreads();
            if (last_lane) {
                cs[0] = exn.v[3];
                cs[1] = eyn.v[3];
            }
            if (own) {
                const float *xb = xbuf + (size_t)((par * (T + 1) + (w - 1)) * 2) * 256;
                const Q4 exb = act ? Q4{{xb[l4], xb[l4 + 1], xb[l4 + 2], xb[l4 + 3]}} : zero4();
                const Q4 ezb = act ? Q4{{xb[256 + l4], xb[256 + l4 + 1], xb[256 + l4 + 2], xb[256 + l4 + 3]}} : zero4();
                Q4 hxo, hyo, hzo;
                compute_H<PML>(g, cp, src, i, j, m, l4, act, hc, exn, eyn, ezn, exb, ezb, eyp, ezp, exm, eym, hxo, hyo,
                               hzo);
                const size_t rbase = row_base(g, i, j, m);
                if (act) {
                    st4(oex + rbase + l4, exn);
                    st4(oey + rbase + l4, eyn);
                    st4(oez + rbase + l4, ezn);
                    st4(ohx + rbase + l4, hxo);
                    st4(ohy + rbase + l4, hyo);
                    st4(ohz + rbase + l4, hzo);
                }
#pragma unroll
                for (int n = 0; n < kMaxProbe; ++n) {
                    if (n < prb.n && act && prb.i[n] == i && prb.j[n] == j) {
                        const int cc = prb.comp[n];
#pragma unroll
                        for (int l = 0; l < 4; ++l) {
                            if (prb.k[n] == k0 + l) {
                                const float v = cc == 0 ? exn.v[l] : cc == 1 ? eyn.v[l] : cc == 2 ? ezn.v[l]
                                              : cc == 3 ? hxo.v[l] : cc == 4 ? hyo.v[l] : hzo.v[l];
                                prb.out[n] = v;
                            }
                        }
                    }
                }
            }
            eyp = eyn;
            ezp = ezn;
        }
    }
}

}  // namespace mifdtd

// =========================================================================== C ABI
using namespace mifdtd;

extern "C" {

int mifdtd_fused_run(const mifdtd_fused_desc *d, int nsteps, const float *tvals, float *probe_series,
                     int *parity, mifdtd_stream_t stream)
{
    if (!d) return fail(MIFDTD_ENULL, "fused descriptor is NULL");
    const int nx = d->nx, ny = d->ny, nzp = d->nz_pitch, pad = d->pad;
    if (int rc = check_grid(nx, ny, nzp, pad)) return rc;
    const int T = d->rows_per_thread;
    if (T < 1 || T > kMaxRows) return fail(MIFDTD_EINVAL, "rows_per_thread (rows per block) must be in [1,%d]", kMaxRows);
    const int CX = d->planes_per_chunk;
    if (CX < 1) return fail(MIFDTD_EINVAL, "planes_per_chunk must be >= 1");
    if (nsteps < 0) return fail(MIFDTD_EINVAL, "nsteps < 0");
    for (int s = 0; s < 2; s++)
        for (int q = 0; q < 6; q++) {
            const void *p = s == 0 ? (const void *)d->e[q] : (const void *)d->e_alt[q];
            if (!p) return fail(MIFDTD_ENULL, "field buffer is NULL");
            if (!aligned16(p)) return fail(MIFDTD_EALIGN, "field buffer not 16-byte aligned");
        }
    const bool hc = d->ce[0] != nullptr;
    if (hc && (!d->ce[1] || !d->ce[2])) return fail(MIFDTD_ENULL, "CE arrays must be all set or all NULL");
    if (hc && (!aligned16(d->ce[0]) || !aligned16(d->ce[1]) || !aligned16(d->ce[2])))
        return fail(MIFDTD_EALIGN, "CE arrays not 16-byte aligned");
    const int nz = nzp - pad;
    if (d->nsrc < 0 || d->nsrc > MIFDTD_FUSED_MAX_SRC) return fail(MIFDTD_EINVAL, "nsrc outside [0,%d]", MIFDTD_FUSED_MAX_SRC);
    if (d->nprobe < 0 || d->nprobe > MIFDTD_FUSED_MAX_PROBE) return fail(MIFDTD_EINVAL, "nprobe outside [0,%d]", MIFDTD_FUSED_MAX_PROBE);
    if (d->nprobe && !probe_series) return fail(MIFDTD_ENULL, "probe_series is NULL");
    if (d->nsrc && !tvals) return fail(MIFDTD_ENULL, "tvals is NULL");

    FGeom g;
    g.nx = nx; g.ny = ny; g.nzp = nzp; g.nz = nz;
    g.T = T;
    g.nrb = (ny + T - 1) / T;
    g.cx = CX;
    g.nchunks = (nx + CX - 1) / CX;
    g.nseg = (nzp + 255) / 256;
    const long long nbl = (long long)g.nrb * g.nchunks;
    if (nbl >= (1ll << 31)) return fail(MIFDTD_ERANGE, "too many blocks");
    g.nblocks = (int)nbl;
    g.nwaves = g.nblocks * (T + 1);
    const size_t lds_bytes = sizeof(float) * ((size_t)2 * (T + 1) * 2 * 256 + (size_t)(CX + 1) * (T + 1) * 2);
    if (lds_bytes > 64 * 1024) return fail(MIFDTD_EINVAL, "rows_per_thread %d / planes_per_chunk %d need %zu B of LDS (> 64 KiB)", T, CX, lds_bytes);

    FCpml cp{};
    const bool pml = d->npml > 0;
    if (pml) {
        const int np_ = d->npml;
        if (np_ > MIFDTD_MAX_NPML) return fail(MIFDTD_EINVAL, "npml %d > %d", np_, MIFDTD_MAX_NPML);
        if (nx < 2 * np_ + 2 || ny < 2 * np_ + 2 || nz < 2 * np_ + 2)
            return fail(MIFDTD_EINVAL, "grid too small for npml %d", np_);
        if (!d->bE || !d->aE || !d->bH || !d->aH) return fail(MIFDTD_ENULL, "CPML tables are NULL");
        if (!d->psiE || !d->psiE_alt || !d->psiH) return fail(MIFDTD_ENULL, "packed psi buffers are NULL");
        if (!aligned16(d->psiE) || !aligned16(d->psiE_alt) || !aligned16(d->psiH))
            return fail(MIFDTD_EALIGN, "packed psi buffers not 16-byte aligned");
        cp.npml = np_;
        cp.bE = d->bE; cp.aE = d->aE; cp.bH = d->bH; cp.aH = d->aH;
        cp.h = d->psiH;
        cp.sX = (size_t)np_ * ny * nzp;
        cp.sY = (size_t)nx * np_ * nzp;
        cp.sZ = (size_t)nx * ny * np_;
    }

    FSrc src{};
    src.n = d->nsrc;
    for (int n = 0; n < d->nsrc; n++) {
        const int *b = d->src_box[n];
        if (d->src_comp[n] < 0 || d->src_comp[n] > 5) return fail(MIFDTD_EINVAL, "source component %d", d->src_comp[n]);
        if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] >= nx || b[3] >= ny || b[5] >= nzp || b[0] > b[1] || b[2] > b[3] || b[4] > b[5])
            return fail(MIFDTD_EINVAL, "source box %d outside the grid", n);
        src.comp[n] = d->src_comp[n];
        src.i0[n] = b[0]; src.i1[n] = b[1]; src.j0[n] = b[2]; src.j1[n] = b[3]; src.k0[n] = b[4]; src.k1[n] = b[5];
        src.sv[n] = d->src_sv[n];
        src.svs[n] = d->src_svs[n];
        src.ow[n] = d->src_overwrite[n];
    }
    FProbe prb{};
    prb.n = d->nprobe;
    for (int n = 0; n < d->nprobe; n++) {
        const int *p = d->probe_ijk[n];
        if (d->probe_comp[n] < 0 || d->probe_comp[n] > 5 || p[0] < 0 || p[0] >= nx || p[1] < 0 || p[1] >= ny || p[2] < 0 || p[2] >= nzp)
            return fail(MIFDTD_EINVAL, "probe %d outside the grid", n);
        prb.comp[n] = d->probe_comp[n];
        prb.i[n] = p[0]; prb.j[n] = p[1]; prb.k[n] = p[2];
    }

    const float *in[6];
    float *out[6];
    for (int q = 0; q < 6; q++) { in[q] = d->e[q]; out[q] = d->e_alt[q]; }
    float *pin = d->psiE, *pout = d->psiE_alt;

    hipStream_t st = to_stream(stream);
    const unsigned nthr = 64u * (unsigned)(T + 1);
    for (int s = 0; s < nsteps; s++) {
        for (int n = 0; n < src.n; n++) src.tval[n] = tvals[(size_t)s * src.n + n];
        prb.out = d->nprobe ? probe_series + (size_t)s * d->nprobe : nullptr;
#define MIFDTD_FUSED_LAUNCH(HCV, PMLV)                                                                          \
    k_fused<HCV, PMLV><<<g.nblocks, nthr, lds_bytes, st>>>(g, in[0], in[1], in[2], in[3], in[4], in[5], out[0], \
                                                          out[1], out[2], out[3], out[4], out[5], d->ce[0],     \
                                                          d->ce[1], d->ce[2], pin, pout, d->psiH, cp, src, prb)
        if (hc && pml) MIFDTD_FUSED_LAUNCH(true, true);
        else if (hc) MIFDTD_FUSED_LAUNCH(true, false);
        else if (pml) MIFDTD_FUSED_LAUNCH(false, true);
        else MIFDTD_FUSED_LAUNCH(false, false);
#undef MIFDTD_FUSED_LAUNCH
        MIFDTD_CHECK_LAUNCH("fused step launch");
        for (int q = 0; q < 6; q++) { const float *x = in[q]; in[q] = out[q]; out[q] = const_cast<float *>(x); }
        float *tp = pin;
        pin = pout;
        pout = tp;
    }
    if (parity) *parity = nsteps & 1;
    return 0;
}

}  // extern "C"
