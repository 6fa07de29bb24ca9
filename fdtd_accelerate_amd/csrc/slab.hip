// Slab-ordered step: per time step, for each slab of consecutive x planes, one E pass
// (core E + CPML-E + E sources + E probes) then one H pass (core H + CPML-H + H sources
// + H probes) over that slab, in place.  A slab is sized so that the E, H (and CE) data
// it touches stays resident in the 256 MB Infinity Cache between its two passes: the
// H pass then re-reads E_new and H_old from the cache and HBM sees each array once per
// step (60 B/cell + CPML), while both kernels keep the simple one-quad-per-thread,
// high-occupancy streaming form.
//
// Ordering is the reference's (E of the slab before H of the slab; slabs in increasing
// x): H of plane i needs E_new of planes i and i-1 (done by this or the previous slab's
// E pass) and E of plane i needs H_old of planes i and i+1 (not yet touched: the slab's
// H pass runs after its E pass, the next slab's later).  Per-cell arithmetic equals the
// in-place update_e/update_h + CPML + source path, bit for bit (masked lanes: value*0).

#include "common.hpp"

namespace mifdtd {

constexpr int SB = 256;
constexpr int kSMaxSrc = 4;
constexpr int kSMaxProbe = 8;

struct SCpml {
    int npml;
    const float *bE, *aE, *bH, *aH;
    float *pe, *ph;          // packed psi E / psi H
    size_t sX, sY, sZ;
    __device__ __forceinline__ size_t off(int f, int c) const
    {
        const int ax = f >> 1, sd = f & 1;
        if (ax == 0) return (size_t)(2 * sd + c) * sX;
        if (ax == 1) return 4 * sX + (size_t)(2 * sd + c) * sY;
        return 4 * sX + 4 * sY + (size_t)(2 * sd + c) * sZ;
    }
};

struct SSrc {
    int n;
    int comp[kSMaxSrc], i0[kSMaxSrc], i1[kSMaxSrc], j0[kSMaxSrc], j1[kSMaxSrc], k0[kSMaxSrc], k1[kSMaxSrc], ow[kSMaxSrc];
    float tval[kSMaxSrc], svs[kSMaxSrc];
    const float *sv[kSMaxSrc];
};

struct SProbe {
    int n;
    int comp[kSMaxProbe], i[kSMaxProbe], j[kSMaxProbe], k[kSMaxProbe], slot[kSMaxProbe];
    float *out;  // row of the probe series for this step; probe n writes out[slot[n]]
};

struct SGeom {
    int nx, ny, nzp, nz;
    unsigned Q;        // quads per row
    int i0;            // first plane of the slab
    unsigned nquads;   // quads in the slab
    int remap;         // deal the slab's blocks to XCDs in contiguous runs
};

// Blocks are dealt round-robin to the 8 XCDs; renumber them so that each XCD runs a
// contiguous run of rows and the j-1 / j+1 rows a block reads sit in its own L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    if (nb & 7u) return b;
    return (b & 7u) * (nb >> 3) + (b >> 3);
}

__device__ __forceinline__ int sslab_e(int q, int n_e, int npml, int &side, int &p)
{
    side = 0;
    p = 0;
    if (q < npml) { p = q; return p + 1; }
    if (q >= n_e - 1 - npml && q < n_e - 1) { side = 1; p = q - (n_e - 1 - npml); return npml + 1 + p; }
    return -1;
}
__device__ __forceinline__ int sslab_h(int q, int n_e, int npml, int &side, int &p)
{
    side = 0;
    p = 0;
    if (q < npml) { p = q; return p + 1; }
    if (q >= n_e - npml && q < n_e) { side = 1; p = q - (n_e - npml); return npml + 2 + p; }
    return -1;
}

__device__ __forceinline__ void sapply_src(const SSrc &s, int c, int i, int j, int k0, Q4 &v)
{
#pragma unroll
    for (int n = 0; n < kSMaxSrc; ++n) {
        if (n < s.n && s.comp[n] == c && i >= s.i0[n] && i <= s.i1[n] && j >= s.j0[n] && j <= s.j1[n] &&
            k0 + 3 >= s.k0[n] && k0 <= s.k1[n]) {
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

__device__ __forceinline__ void sprobe(const SProbe &p, int c, int i, int j, int k0, const Q4 &v)
{
#pragma unroll
    for (int n = 0; n < kSMaxProbe; ++n) {
        if (n < p.n && p.comp[n] == c && p.i[n] == i && p.j[n] == j) {
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (p.k[n] == k0 + l) p.out[p.slot[n]] = v.v[l];
        }
    }
}

// Both passes are written load-all / compute / store-all: every global load of a quad
// (fields, neighbours, CE, psi) is issued before the first store, so a wave pays one
// memory round trip even on CPML cells (psi pointers may alias the fields as far as the
// compiler knows, so a load placed after a field store would wait for it).

// E pass on quad f of the slab described by g.
template <bool HC, bool PML>
__device__ __forceinline__ void e_quad(
    const SGeom &g, unsigned f, float *__restrict__ ex, float *__restrict__ ey, float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    const SCpml &cp, const SSrc &src, const SProbe &prb)
{
    const int nx = g.nx, ny = g.ny, nzp = g.nzp, nz = g.nz;
    const unsigned row = f / g.Q;
    const int k0 = (int)(f - row * g.Q) * 4;
    const int di = (int)(row / (unsigned)ny);
    const int j = (int)row - di * ny;
    const int i = g.i0 + di;
    const size_t id = ((size_t)i * ny + j) * nzp + k0;
    const size_t sx = (size_t)ny * nzp;
    const bool jy = j < ny - 1, ix = i < nx - 1;
    const Q4 z4 = splat4(0.f);

    // ---------------------------------------------------------------- loads
    const Q4 hx0 = ld4(hx + id), hy0 = ld4(hy + id), hz0 = ld4(hz + id);
    Q4 e0 = ld4(ex + id), e1 = ld4(ey + id), e2 = ld4(ez + id);
    const Q4 c0 = HC ? ld4(cex + id) : splat4(0.5f);
    const Q4 c1 = HC ? ld4(cey + id) : splat4(0.5f);
    const Q4 c2 = HC ? ld4(cez + id) : splat4(0.5f);
    const Q4 hzU = jy ? ld4(hz + id + nzp) : z4;   // hz(j+1)
    const Q4 hxU = jy ? ld4(hx + id + nzp) : z4;   // hx(j+1)
    const float hyn = jy ? hy[id + 4] : 0.f;       // hy(k0+4): next quad / next row start
    const Q4 hzN = ix ? ld4(hz + id + sx) : z4;    // hz(i+1)
    const Q4 hyN = ix ? ld4(hy + id + sx) : z4;    // hy(i+1)
    const float hxn = ix ? hx[id + 4] : 0.f;       // hx(k0+4)

    const int npml = PML ? cp.npml : 0;
    int xs = 0, xp = 0, ys = 0, yp = 0, xt = -1, yt = -1;
    Q4 px1 = z4, px2 = z4, py1 = z4, py2 = z4;
    float *ppx1 = nullptr, *ppx2 = nullptr, *ppy1 = nullptr, *ppy2 = nullptr;
    float pz1[4] = {0.f, 0.f, 0.f, 0.f}, pz2[4] = {0.f, 0.f, 0.f, 0.f};
    int zt[4] = {-1, -1, -1, -1};
    bool anyz = false;
    if (PML) {
        xt = sslab_e(i, nx, npml, xs, xp);  // x slab planes are all < nx-1
        if (xt > 0) {
            const size_t o = ((size_t)xp * ny + j) * nzp + k0;
            ppx1 = cp.pe + cp.off(xs, 0) + o;
            ppx2 = cp.pe + cp.off(xs, 1) + o;
            px1 = ld4(ppx1);
            px2 = ld4(ppx2);
        }
        yt = sslab_e(j, ny, npml, ys, yp);  // y slab rows are all < ny-1
        if (yt > 0) {
            const size_t o = ((size_t)i * npml + yp) * nzp + k0;
            ppy1 = cp.pe + cp.off(2 + ys, 0) + o;
            ppy2 = cp.pe + cp.off(2 + ys, 1) + o;
            py1 = ld4(ppy1);
            py2 = ld4(ppy2);
        }
        if (k0 < npml || k0 + 3 >= nz - 1 - npml) {
            const size_t o = ((size_t)i * ny + j) * npml;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                int zs, zp;
                zt[l] = sslab_e(k0 + l, nz, npml, zs, zp);
                if (zt[l] > 0) {
                    const float *q = cp.pe + cp.off(4 + zs, 0) + o + zp;  // psi2 sits sZ after psi1
                    pz1[l] = q[0];
                    pz2[l] = q[cp.sZ];
                    anyz = true;
                }
            }
        }
    }

    // -------------------------------------------------------------- compute
    // ex: += cex*((hz[j+1]-hz) - (hy[k+1]-hy)) ; CPML + y psi2 (hz), - z psi1 (hy)
    if (jy) {
        float y1[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) y1[l] = l < 3 ? hy0.v[l + 1] : hyn;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = e0.v[l] + c0.v[l] * ((hzU.v[l] - hz0.v[l]) - (y1[l] - hy0.v[l]));
            e0.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
        if (PML) {
            if (yt > 0) {
                const float b = cp.bE[yt], a = cp.aE[yt];
#pragma unroll
                for (int l = 0; l < 4; ++l)
                    if (k0 + l < nz - 1) {
                        const float tt = b * py2.v[l] + a * (hzU.v[l] - hz0.v[l]);
                        py2.v[l] = tt;
                        e0.v[l] = e0.v[l] + c0.v[l] * tt;
                    }
            }
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (zt[l] > 0) {
                    const float tt = cp.bE[zt[l]] * pz1[l] + cp.aE[zt[l]] * (y1[l] - hy0.v[l]);
                    pz1[l] = tt;
                    e0.v[l] = e0.v[l] - c0.v[l] * tt;
                }
        }
    }
    // ey: += cey*((hx[k+1]-hx) - (hz[i+1]-hz)) ; CPML - x psi1 (hz), + z psi2 (hx)
    if (ix) {
        float x1[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) x1[l] = l < 3 ? hx0.v[l + 1] : hxn;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = e1.v[l] + c1.v[l] * ((x1[l] - hx0.v[l]) - (hzN.v[l] - hz0.v[l]));
            e1.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
        if (PML) {
            if (xt > 0) {
                const float b = cp.bE[xt], a = cp.aE[xt];
#pragma unroll
                for (int l = 0; l < 4; ++l)
                    if (k0 + l < nz - 1) {
                        const float tt = b * px1.v[l] + a * (hzN.v[l] - hz0.v[l]);
                        px1.v[l] = tt;
                        e1.v[l] = e1.v[l] - c1.v[l] * tt;
                    }
            }
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (zt[l] > 0) {
                    const float tt = cp.bE[zt[l]] * pz2[l] + cp.aE[zt[l]] * (x1[l] - hx0.v[l]);
                    pz2[l] = tt;
                    e1.v[l] = e1.v[l] + c1.v[l] * tt;
                }
        }
    }
    // ez: += cez*((hy[i+1]-hy) - (hx[j+1]-hx)) ; CPML + x psi2 (hy), - y psi1 (hx)
    if (ix && jy) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = e2.v[l] + c2.v[l] * ((hyN.v[l] - hy0.v[l]) - (hxU.v[l] - hx0.v[l]));
            e2.v[l] = (k0 + l < nz) ? v : v * 0.f;
        }
        if (PML) {
            if (xt > 0) {
                const float b = cp.bE[xt], a = cp.aE[xt];
#pragma unroll
                for (int l = 0; l < 4; ++l)
                    if (k0 + l < nz) {
                        const float tt = b * px2.v[l] + a * (hyN.v[l] - hy0.v[l]);
                        px2.v[l] = tt;
                        e2.v[l] = e2.v[l] + c2.v[l] * tt;
                    }
            }
            if (yt > 0) {
                const float b = cp.bE[yt], a = cp.aE[yt];
#pragma unroll
                for (int l = 0; l < 4; ++l)
                    if (k0 + l < nz) {
                        const float tt = b * py1.v[l] + a * (hxU.v[l] - hx0.v[l]);
                        py1.v[l] = tt;
                        e2.v[l] = e2.v[l] - c2.v[l] * tt;
                    }
            }
        }
    }
    sapply_src(src, 0, i, j, k0, e0);
    sapply_src(src, 1, i, j, k0, e1);
    sapply_src(src, 2, i, j, k0, e2);

    // --------------------------------------------------------------- stores
    // (unchanged lanes and wall rows store back the value they loaded)
    st4(ex + id, e0);
    st4(ey + id, e1);
    st4(ez + id, e2);
    if (PML) {
        if (xt > 0) {
            st4(ppx1, px1);
            st4(ppx2, px2);
        }
        if (yt > 0) {
            st4(ppy1, py1);
            st4(ppy2, py2);
        }
        if (anyz) {
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (zt[l] > 0) {
                    int zs, zp;
                    sslab_e(k0 + l, nz, npml, zs, zp);
                    float *q = cp.pe + cp.off(4 + zs, 0) + ((size_t)i * ny + j) * npml + zp;
                    if (jy) q[0] = pz1[l];
                    if (ix) q[cp.sZ] = pz2[l];
                }
        }
    }
    sprobe(prb, 0, i, j, k0, e0);
    sprobe(prb, 1, i, j, k0, e1);
    sprobe(prb, 2, i, j, k0, e2);
}

// H pass on quad f of the slab described by g.
template <bool PML>
__device__ __forceinline__ void h_quad(
    const SGeom &g, unsigned f, const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
    const SCpml &cp, const SSrc &src, const SProbe &prb)
{
    const int nx = g.nx, ny = g.ny, nzp = g.nzp, nz = g.nz;
    (void)nx;
    const unsigned row = f / g.Q;
    const int k0 = (int)(f - row * g.Q) * 4;
    const int di = (int)(row / (unsigned)ny);
    const int j = (int)row - di * ny;
    const int i = g.i0 + di;
    const size_t id = ((size_t)i * ny + j) * nzp + k0;
    const size_t sx = (size_t)ny * nzp;
    const bool jp = j > 0, ip = i > 0;
    const Q4 z4 = splat4(0.f);

    // ---------------------------------------------------------------- loads
    const Q4 ex0 = ld4(ex + id), ey0 = ld4(ey + id), ez0 = ld4(ez + id);
    Q4 h0 = ld4(hx + id), h1 = ld4(hy + id), h2 = ld4(hz + id);
    const Q4 ezD = jp ? ld4(ez + id - nzp) : z4;   // ez(j-1)
    const Q4 exD = jp ? ld4(ex + id - nzp) : z4;   // ex(j-1)
    const float eyp = jp ? ey[id - 1] : 0.f;       // ey(k0-1): previous quad / previous row end
    const Q4 ezB = ip ? ld4(ez + id - sx) : z4;    // ez(i-1)
    const Q4 eyB = ip ? ld4(ey + id - sx) : z4;    // ey(i-1)
    const float exp_ = ip ? ex[id - 1] : 0.f;      // ex(k0-1)

    const int npml = PML ? cp.npml : 0;
    int xs = 0, xp = 0, ys = 0, yp = 0, xt = -1, yt = -1;
    Q4 px1 = z4, px2 = z4, py1 = z4, py2 = z4;
    float *ppx1 = nullptr, *ppx2 = nullptr, *ppy1 = nullptr, *ppy2 = nullptr;
    float pz1[4] = {0.f, 0.f, 0.f, 0.f}, pz2[4] = {0.f, 0.f, 0.f, 0.f};
    int zt[4] = {-1, -1, -1, -1};
    bool anyz = false;
    if (PML) {
        xt = ip ? sslab_h(i, nx, npml, xs, xp) : -1;  // plane 0 is the wall
        if (xt > 0) {
            const size_t o = ((size_t)xp * ny + j) * nzp + k0;
            ppx1 = cp.ph + cp.off(xs, 0) + o;
            ppx2 = cp.ph + cp.off(xs, 1) + o;
            px1 = ld4(ppx1);
            px2 = ld4(ppx2);
        }
        yt = jp ? sslab_h(j, ny, npml, ys, yp) : -1;  // row 0 is the wall
        if (yt > 0) {
            const size_t o = ((size_t)i * npml + yp) * nzp + k0;
            ppy1 = cp.ph + cp.off(2 + ys, 0) + o;
            ppy2 = cp.ph + cp.off(2 + ys, 1) + o;
            py1 = ld4(ppy1);
            py2 = ld4(ppy2);
        }
        if (k0 < npml || k0 + 3 >= nz - npml) {
            const size_t o = ((size_t)i * ny + j) * npml;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                int zs, zp;
                const int k = k0 + l;
                zt[l] = k > 0 ? sslab_h(k, nz, npml, zs, zp) : -1;  // k = 0 is the wall
                if (zt[l] > 0) {
                    const float *q = cp.ph + cp.off(4 + zs, 0) + o + zp;
                    pz1[l] = q[0];
                    pz2[l] = q[cp.sZ];
                    anyz = true;
                }
            }
        }
    }

    // -------------------------------------------------------------- compute
    float ym[4], xm[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        ym[l] = l > 0 ? ey0.v[l - 1] : eyp;
        xm[l] = l > 0 ? ex0.v[l - 1] : exp_;
    }
    // hx: j>0 ; -= 0.5*((ez - ez[j-1]) - (ey - ey[k-1])) ; CPML - y psi2 (ez), + z psi1 (ey)
    if (jp) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = h0.v[l] - 0.5f * ((ez0.v[l] - ezD.v[l]) - (ey0.v[l] - ym[l]));
            const int k = k0 + l;
            h0.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
        }
        if (PML) {
            if (yt > 0) {
                const float b = cp.bH[yt], a = cp.aH[yt];
#pragma unroll
                for (int l = 0; l < 4; ++l) {
                    const int k = k0 + l;
                    if (k > 0 && k < nz) {
                        const float tt = b * py2.v[l] + a * (ez0.v[l] - ezD.v[l]);
                        py2.v[l] = tt;
                        h0.v[l] = h0.v[l] - 0.5f * tt;
                    }
                }
            }
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (zt[l] > 0) {
                    const float tt = cp.bH[zt[l]] * pz1[l] + cp.aH[zt[l]] * (ey0.v[l] - ym[l]);
                    pz1[l] = tt;
                    h0.v[l] = h0.v[l] + 0.5f * tt;
                }
        }
    }
    if (ip) {
        // hy: -= 0.5*((ex - ex[k-1]) - (ez - ez[i-1])) ; CPML + x psi1 (ez), - z psi2 (ex)
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = h1.v[l] - 0.5f * ((ex0.v[l] - xm[l]) - (ez0.v[l] - ezB.v[l]));
            const int k = k0 + l;
            h1.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
        }
        if (PML) {
            if (xt > 0) {
                const float b = cp.bH[xt], a = cp.aH[xt];
#pragma unroll
                for (int l = 0; l < 4; ++l) {
                    const int k = k0 + l;
                    if (k > 0 && k < nz) {
                        const float tt = b * px1.v[l] + a * (ez0.v[l] - ezB.v[l]);
                        px1.v[l] = tt;
                        h1.v[l] = h1.v[l] + 0.5f * tt;
                    }
                }
            }
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (zt[l] > 0) {
                    const float tt = cp.bH[zt[l]] * pz2[l] + cp.aH[zt[l]] * (ex0.v[l] - xm[l]);
                    pz2[l] = tt;
                    h1.v[l] = h1.v[l] - 0.5f * tt;
                }
        }
        // hz: j>0 ; -= 0.5*((ey - ey[i-1]) - (ex - ex[j-1])) ; CPML - x psi2 (ey), + y psi1 (ex)
        if (jp) {
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float v = h2.v[l] - 0.5f * ((ey0.v[l] - eyB.v[l]) - (ex0.v[l] - exD.v[l]));
                h2.v[l] = (k0 + l < nz) ? v : v * 0.f;
            }
            if (PML) {
                if (xt > 0) {
                    const float b = cp.bH[xt], a = cp.aH[xt];
#pragma unroll
                    for (int l = 0; l < 4; ++l)
                        if (k0 + l < nz) {
                            const float tt = b * px2.v[l] + a * (ey0.v[l] - eyB.v[l]);
                            px2.v[l] = tt;
                            h2.v[l] = h2.v[l] - 0.5f * tt;
                        }
                }
                if (yt > 0) {
                    const float b = cp.bH[yt], a = cp.aH[yt];
#pragma unroll
                    for (int l = 0; l < 4; ++l)
                        if (k0 + l < nz) {
                            const float tt = b * py1.v[l] + a * (ex0.v[l] - exD.v[l]);
                            py1.v[l] = tt;
                            h2.v[l] = h2.v[l] + 0.5f * tt;
                        }
                }
            }
        }
    }
    sapply_src(src, 3, i, j, k0, h0);
    sapply_src(src, 4, i, j, k0, h1);
    sapply_src(src, 5, i, j, k0, h2);

    // --------------------------------------------------------------- stores
    st4(hx + id, h0);
    st4(hy + id, h1);
    st4(hz + id, h2);
    if (PML) {
        if (xt > 0) {
            st4(ppx1, px1);
            st4(ppx2, px2);
        }
        if (yt > 0) {
            st4(ppy1, py1);
            st4(ppy2, py2);
        }
        if (anyz) {
            const size_t o = ((size_t)i * ny + j) * npml;
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (zt[l] > 0) {
                    int zs, zp;
                    sslab_h(k0 + l, nz, npml, zs, zp);
                    float *q = cp.ph + cp.off(4 + zs, 0) + o + zp;
                    if (jp) q[0] = pz1[l];
                    if (ip) q[cp.sZ] = pz2[l];
                }
        }
    }
    sprobe(prb, 3, i, j, k0, h0);
    sprobe(prb, 4, i, j, k0, h1);
    sprobe(prb, 5, i, j, k0, h2);
}

struct SFields {
    float *e[6];
    const float *ce[3];
};

template <bool HC, bool PML>
__global__ __launch_bounds__(SB) void k_slab_e(SGeom g, SFields F, SCpml cp, SSrc src, SProbe prb)
{
    const unsigned f = (g.remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x) * SB + threadIdx.x;
    if (f < g.nquads)
        e_quad<HC, PML>(g, f, F.e[0], F.e[1], F.e[2], F.e[3], F.e[4], F.e[5], F.ce[0], F.ce[1], F.ce[2], cp, src, prb);
}

template <bool PML>
__global__ __launch_bounds__(SB) void k_slab_h(SGeom g, SFields F, SCpml cp, SSrc src, SProbe prb)
{
    const unsigned f = (g.remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x) * SB + threadIdx.x;
    if (f < g.nquads) h_quad<PML>(g, f, F.e[0], F.e[1], F.e[2], F.e[3], F.e[4], F.e[5], cp, src, prb);
}

}  // namespace mifdtd

// =========================================================================== C ABI
using namespace mifdtd;

extern "C" {

int mifdtd_slab_run(const mifdtd_fused_desc *d, int slab_planes, int nsteps, const float *tvals,
                    float *probe_series, mifdtd_stream_t stream)
{
    if (!d) return fail(MIFDTD_ENULL, "step descriptor is NULL");
    const int nx = d->nx, ny = d->ny, nzp = d->nz_pitch, pad = d->pad;
    if (int rc = check_grid(nx, ny, nzp, pad)) return rc;
    if (slab_planes < 1) return fail(MIFDTD_EINVAL, "slab_planes must be >= 1 : %d", slab_planes);
    if (nsteps < 0) return fail(MIFDTD_EINVAL, "nsteps < 0");
    for (int c = 0; c < 6; c++) {
        if (!d->e[c]) return fail(MIFDTD_ENULL, "field buffer is NULL");
        if (!aligned16(d->e[c])) return fail(MIFDTD_EALIGN, "field buffer not 16-byte aligned");
    }
    const bool hc = d->ce[0] != nullptr;
    if (hc && (!d->ce[1] || !d->ce[2])) return fail(MIFDTD_ENULL, "CE arrays must be all set or all NULL");
    if (hc && (!aligned16(d->ce[0]) || !aligned16(d->ce[1]) || !aligned16(d->ce[2])))
        return fail(MIFDTD_EALIGN, "CE arrays not 16-byte aligned");
    const int nz = nzp - pad;
    if (d->nsrc < 0 || d->nsrc > MIFDTD_FUSED_MAX_SRC) return fail(MIFDTD_EINVAL, "nsrc outside [0,%d]", MIFDTD_FUSED_MAX_SRC);
    if (d->nprobe < 0 || d->nprobe > MIFDTD_FUSED_MAX_PROBE)
        return fail(MIFDTD_EINVAL, "nprobe outside [0,%d]", MIFDTD_FUSED_MAX_PROBE);
    if (d->nprobe && !probe_series) return fail(MIFDTD_ENULL, "probe_series is NULL");
    if (d->nsrc && !tvals) return fail(MIFDTD_ENULL, "tvals is NULL");

    SCpml cp{};
    const bool pml = d->npml > 0;
    if (pml) {
        const int np_ = d->npml;
        if (np_ > MIFDTD_MAX_NPML) return fail(MIFDTD_EINVAL, "npml %d > %d", np_, MIFDTD_MAX_NPML);
        if (nx < 2 * np_ + 2 || ny < 2 * np_ + 2 || nz < 2 * np_ + 2)
            return fail(MIFDTD_EINVAL, "grid (%d,%d,%d) too small for npml %d", nx, ny, nz, np_);
        if (!d->bE || !d->aE || !d->bH || !d->aH) return fail(MIFDTD_ENULL, "CPML tables are NULL");
        if (!d->psiE || !d->psiH) return fail(MIFDTD_ENULL, "packed psi buffers are NULL");
        if (!aligned16(d->psiE) || !aligned16(d->psiH))
            return fail(MIFDTD_EALIGN, "packed psi buffers not 16-byte aligned");
        cp.npml = np_;
        cp.bE = d->bE;
        cp.aE = d->aE;
        cp.bH = d->bH;
        cp.aH = d->aH;
        cp.pe = d->psiE;
        cp.ph = d->psiH;
        cp.sX = (size_t)np_ * ny * nzp;
        cp.sY = (size_t)nx * np_ * nzp;
        cp.sZ = (size_t)nx * ny * np_;
    }
    for (int n = 0; n < d->nsrc; n++) {
        const int *b = d->src_box[n];
        if (d->src_comp[n] < 0 || d->src_comp[n] > 5) return fail(MIFDTD_EINVAL, "source %d component %d", n, d->src_comp[n]);
        if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] >= nx || b[3] >= ny || b[5] >= nzp || b[0] > b[1] || b[2] > b[3] ||
            b[4] > b[5])
            return fail(MIFDTD_EINVAL, "source box %d outside the grid", n);
    }
    // probes split by half step (E components 0..2 are final after the E pass)
    SProbe pe{}, ph{};
    for (int n = 0; n < d->nprobe; n++) {
        const int *p = d->probe_ijk[n];
        const int c = d->probe_comp[n];
        if (c < 0 || c > 5 || p[0] < 0 || p[0] >= nx || p[1] < 0 || p[1] >= ny || p[2] < 0 || p[2] >= nzp)
            return fail(MIFDTD_EINVAL, "probe %d outside the grid", n);
        SProbe &P = c < 3 ? pe : ph;
        const int m = P.n++;
        P.comp[m] = c;
        P.i[m] = p[0];
        P.j[m] = p[1];
        P.k[m] = p[2];
        P.slot[m] = n;
    }

    const unsigned Q = (unsigned)(nzp / 4);
    const unsigned long long plane_q = (unsigned long long)ny * Q;
    const int S = slab_planes < nx ? slab_planes : nx;
    if (plane_q * (unsigned long long)S + SB >= (1ull << 31)) return fail(MIFDTD_ERANGE, "slab of %d planes too large", S);
    hipStream_t st = to_stream(stream);
    SFields F;
    for (int c = 0; c < 6; c++) F.e[c] = d->e[c];
    for (int c = 0; c < 3; c++) F.ce[c] = d->ce[c];
    const int ns = (nx + S - 1) / S;
    // Small slabs: give each XCD a contiguous run of rows (j+-1 reuse in its L2).
    const int remap = S <= 8;
    auto geom = [&](int sl) {
        SGeom g;
        g.nx = nx;
        g.ny = ny;
        g.nzp = nzp;
        g.nz = nz;
        g.Q = Q;
        g.i0 = sl * S;
        g.nquads = (unsigned)(plane_q * (unsigned long long)(nx - g.i0 < S ? nx - g.i0 : S));
        g.remap = remap;
        return g;
    };
    auto sources = [&](int step, bool e_half) {
        SSrc X{};
        for (int n = 0; n < d->nsrc; n++) {
            if ((d->src_comp[n] < 3) != e_half) continue;
            const int *b = d->src_box[n];
            const int m = X.n++;
            X.comp[m] = d->src_comp[n];
            X.i0[m] = b[0];
            X.i1[m] = b[1];
            X.j0[m] = b[2];
            X.j1[m] = b[3];
            X.k0[m] = b[4];
            X.k1[m] = b[5];
            X.sv[m] = d->src_sv[n];
            X.svs[m] = d->src_svs[n];
            X.ow[m] = d->src_overwrite[n];
            X.tval[m] = tvals[(size_t)step * d->nsrc + n];
        }
        return X;
    };
    auto probes = [&](const SProbe &P, int step) {
        SProbe X = P;
        X.out = d->nprobe ? probe_series + (size_t)step * d->nprobe : nullptr;
        return X;
    };
    for (int n = 0; n < nsteps; n++) {
        const SSrc se = sources(n, true), sh = sources(n, false);
        const SProbe pen = probes(pe, n), phn = probes(ph, n);
        for (int sl = 0; sl < ns; sl++) {
            const SGeom g = geom(sl);
            const unsigned grid = (g.nquads + SB - 1) / SB;
            if (hc && pml) k_slab_e<true, true><<<grid, SB, 0, st>>>(g, F, cp, se, pen);
            else if (hc) k_slab_e<true, false><<<grid, SB, 0, st>>>(g, F, cp, se, pen);
            else if (pml) k_slab_e<false, true><<<grid, SB, 0, st>>>(g, F, cp, se, pen);
            else k_slab_e<false, false><<<grid, SB, 0, st>>>(g, F, cp, se, pen);
            MIFDTD_CHECK_LAUNCH("slab E launch");
            if (pml) k_slab_h<true><<<grid, SB, 0, st>>>(g, F, cp, sh, phn);
            else k_slab_h<false><<<grid, SB, 0, st>>>(g, F, cp, sh, phn);
            MIFDTD_CHECK_LAUNCH("slab H launch");
        }
    }
    return 0;
}

}  // extern "C"
