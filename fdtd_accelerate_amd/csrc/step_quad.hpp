// Per-quad whole-step updates shared by the slab engine and the batched engine: the E
// pass of one quad (core E + CPML-E + E sources + E probes) and the H pass (core H +
// CPML-H + H sources + H probes), with the same per-cell arithmetic as the in-place
// update_e/update_h + CPML + source kernels (bit for bit, masked lanes value*0).
#pragma once

#include "common.hpp"

namespace mifdtd {

constexpr int SB = 256;
constexpr int kSMaxSrc = 4;
constexpr int kSMaxProbe = 8;

struct SCpml {
    int npml;
    int axes;                // bit a: CPML on both faces of axis a (x = 1, y = 2, z = 4)
    const float *bE, *aE, *bH, *aH;
    float *pe, *ph;          // packed psi E / psi H
    float *pe_out;           // where the E pass stores psi E (pe for in-place updates)
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

// A Drude box (auxiliary current; kuFDTD matter.py:104-221, drude_ade.c): on cells of the
// inclusive box [i0,i1] x [j0,j1] x [k0,k1], right after the core E update and before
// CPML-E and the sources, each component takes e' = e - CE J ; J' = CJA J + CJB e'.
// a/b: CJA/CJB, C-order box arrays (lj x lk per plane); J is read from jin and J' written
// to jout (the step's two current sets).
struct SDrude {
    int i0, i1, j0, j1, k0, k1;
    int lj, lk;
    const float *a[3], *b[3], *jin[3];
    float *jout[3];
    __device__ __forceinline__ bool row_in(int i, int j) const { return i >= i0 && i <= i1 && j >= j0 && j <= j1; }
};

// Drude update of component c on the quad (i, j, k0..k0+3) of a row inside the box (row_in):
// e: the core-updated values, ce: the quad's CE; returns J' in jn (stored by the caller
// for the cells it owns).
__device__ __forceinline__ void drude_quad(const SDrude &d, int c, int i, int j, int k0, const Q4 &ce, Q4 &e, Q4 &jn)
{
    const size_t rowb = ((size_t)(i - d.i0) * d.lj + (j - d.j0)) * d.lk;
    bool in[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) in[l] = k0 + l >= d.k0 && k0 + l <= d.k1;
    Q4 f, a, b;
    if (in[0] && in[3] && ((rowb + (size_t)(k0 - d.k0)) & 3u) == 0) {  // whole quad, 16-byte aligned
        const size_t n0 = rowb + (size_t)(k0 - d.k0);
        f = ld4(d.jin[c] + n0);
        a = ld4(d.a[c] + n0);
        b = ld4(d.b[c] + n0);
    } else {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const size_t n = in[l] ? rowb + (size_t)(k0 + l - d.k0) : rowb;
            f.v[l] = in[l] ? d.jin[c][n] : 0.f;
            a.v[l] = in[l] ? d.a[c][n] : 0.f;
            b.v[l] = in[l] ? d.b[c][n] : 0.f;
        }
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        if (in[l]) {
            const float et = e.v[l] - ce.v[l] * f.v[l];
            e.v[l] = et;
            jn.v[l] = a.v[l] * f.v[l] + b.v[l] * et;
        }
    }
}

// Stores the J' of drude_quad for the quad's cells inside the box.
__device__ __forceinline__ void drude_store(const SDrude &d, int c, int i, int j, int k0, const Q4 &jn)
{
    const size_t rowb = ((size_t)(i - d.i0) * d.lj + (j - d.j0)) * d.lk;
    const size_t n0 = rowb + (size_t)(k0 - d.k0);
    if (k0 >= d.k0 && k0 + 3 <= d.k1 && (n0 & 3u) == 0) {
        st4(d.jout[c] + n0, jn);
        return;
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int k = k0 + l;
        if (k >= d.k0 && k <= d.k1) d.jout[c][rowb + (size_t)(k - d.k0)] = jn.v[l];
    }
}

// TFSF face patches of one half step (kuFDTD tfsf.py:138-226 tables, normal incidence):
// patch p adds 0.5 * (coef * line[base + along]) to component comp on its inclusive box,
// along being the offset along the propagation axis from the box start; patches apply in
// table order after the sources.  line: the 1-D incident line of the other half (H for
// the E patches, E for the H patches).
constexpr int kTfMaxPatch = 12;
struct STfsf {
    int n, axis, line_n;
    int comp[kTfMaxPatch], box[kTfMaxPatch][6], base[kTfMaxPatch];
    float coef[kTfMaxPatch];
    const float *line;
};

__device__ __forceinline__ void tfsf_add(const STfsf &t, int p, int i, int j, int k0, Q4 &f)
{
    const int *b = t.box[p];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int k = k0 + l;
        if (k >= b[4] && k <= b[5]) {
            const int along = t.axis == 0 ? i - b[0] : t.axis == 1 ? j - b[2] : k - b[4];
            const float v = 0.5f * (t.coef[p] * t.line[t.base[p] + along]);
            f.v[l] = f.v[l] + v;
        }
    }
}

// The patches of one half step on the quad (i, j, k0..k0+3); c0: 0 for E, 3 for H; mask:
// the patches to consider (bit p; patches known not to touch the quad may be left out),
// applied in table order.
__device__ __forceinline__ void tfsf_quad(const STfsf &t, int c0, int i, int j, int k0, Q4 &f0, Q4 &f1, Q4 &f2,
                                          unsigned mask = ~0u)
{
    for (int p = 0; p < t.n; ++p) {
        if (!((mask >> p) & 1u)) continue;
        const int *b = t.box[p];
        if (i < b[0] || i > b[1] || j < b[2] || j > b[3]) continue;
        if (k0 + 3 < b[4] || k0 > b[5]) continue;  // the quad misses the patch (z-face patches: all lanes but one)
        const int c = t.comp[p] - c0;
        if (c == 0) tfsf_add(t, p, i, j, k0, f0);
        else if (c == 1) tfsf_add(t, p, i, j, k0, f1);
        else tfsf_add(t, p, i, j, k0, f2);
    }
}

// Bit p set for each patch of t whose box meets planes [ia, ib], row j and cells [ka, kb].
__device__ __forceinline__ unsigned tfsf_mask(const STfsf &t, int ia, int ib, int j, int ka, int kb)
{
    unsigned m = 0;
    for (int p = 0; p < t.n; ++p) {
        const int *b = t.box[p];
        if (b[0] <= ib && ia <= b[1] && j >= b[2] && j <= b[3] && b[4] <= kb && ka <= b[5]) m |= 1u << p;
    }
    return m;
}

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

// How CE reaches the E update: the constant 0.5, three float arrays, or three uint8
// material-index arrays into a 3 x 256 float table (same float values, so the products
// are identical; 3 B per cell instead of 12).  The per-quad passes read the 3 KiB table
// from global memory, where it stays in L1/L2; the sweep kernels stage it in LDS.
enum CeMode { kCeConst = 0, kCeArrays = 1, kCeTable = 2, kCeTriple = 3 };  // kCeTriple: the fused sweep only

// CE of a quad as uint8 material indices m[0..3] into a 256-entry float table.
__device__ __forceinline__ Q4 ce_from_table(const unsigned char *__restrict__ m, const float *tab)
{
    const uchar4 t = *reinterpret_cast<const uchar4 *>(m);
    Q4 q;
    q.v[0] = tab[t.x];
    q.v[1] = tab[t.y];
    q.v[2] = tab[t.z];
    q.v[3] = tab[t.w];
    return q;
}

// E pass on the quad (i, j, k0..k0+3): reads E from exi.. and writes it to exo.. (the
// same arrays for an in-place update, the next-step buffers for a ping-pong one); psi E
// is read from cp.pe and written to cp.pe_out.  CE: CM = 0 the constant 0.5, 1 the float
// arrays cex.., 2 the uint8 material indices mx.. into tab ([3][256] floats, the same
// values as the arrays, so the products are identical).
constexpr int kFeatDrude = 1, kFeatTfsf = 2;  // optional E-pass features (template bits)

template <int CM, bool PML, int FEAT = 0>
__device__ __forceinline__ void e_quad_at(
    const SGeom &g, int i, int j, int k0, const float *exi, const float *eyi, const float *ezi, float *exo,
    float *eyo, float *ezo, const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    const SCpml &cp, const SSrc &src, const SProbe &prb, const unsigned char *__restrict__ mx = nullptr,
    const unsigned char *__restrict__ my = nullptr, const unsigned char *__restrict__ mz = nullptr,
    const float *tab = nullptr, const SDrude *dr = nullptr, const STfsf *tf = nullptr)
{
    constexpr bool DR = (FEAT & kFeatDrude) != 0;
    constexpr bool TF = (FEAT & kFeatTfsf) != 0;
    const int nx = g.nx, ny = g.ny, nzp = g.nzp, nz = g.nz;
    const size_t id = ((size_t)i * ny + j) * nzp + k0;
    const size_t sx = (size_t)ny * nzp;
    const bool jy = j < ny - 1, ix = i < nx - 1;
    const Q4 z4 = splat4(0.f);

    // ---------------------------------------------------------------- loads
    const Q4 hx0 = ld4(hx + id), hy0 = ld4(hy + id), hz0 = ld4(hz + id);
    Q4 e0 = ld4(exi + id), e1 = ld4(eyi + id), e2 = ld4(ezi + id);
    const Q4 c0 = CM == 2 ? ce_from_table(mx + id, tab) : CM ? ld4(cex + id) : splat4(0.5f);
    const Q4 c1 = CM == 2 ? ce_from_table(my + id, tab + 256) : CM ? ld4(cey + id) : splat4(0.5f);
    const Q4 c2 = CM == 2 ? ce_from_table(mz + id, tab + 512) : CM ? ld4(cez + id) : splat4(0.5f);
    const Q4 hzU = jy ? ld4(hz + id + nzp) : z4;   // hz(j+1)
    const Q4 hxU = jy ? ld4(hx + id + nzp) : z4;   // hx(j+1)
    const float hyn = jy ? hy[id + 4] : 0.f;       // hy(k0+4): next quad / next row start
    const Q4 hzN = ix ? ld4(hz + id + sx) : z4;    // hz(i+1)
    const Q4 hyN = ix ? ld4(hy + id + sx) : z4;    // hy(i+1)
    const float hxn = ix ? hx[id + 4] : 0.f;       // hx(k0+4)

    const int npml = PML ? cp.npml : 0;
    int xs = 0, xp = 0, ys = 0, yp = 0, xt = -1, yt = -1;
    Q4 px1 = z4, px2 = z4, py1 = z4, py2 = z4;
    size_t ox1 = 0, ox2 = 0, oy1 = 0, oy2 = 0;  // psi offsets of the x / y face quads
    float pz1[4] = {0.f, 0.f, 0.f, 0.f}, pz2[4] = {0.f, 0.f, 0.f, 0.f};
    int zt[4] = {-1, -1, -1, -1};
    bool anyz = false;
    if (PML) {
        xt = (cp.axes & 1) ? sslab_e(i, nx, npml, xs, xp) : -1;  // x slab planes are all < nx-1
        if (xt > 0) {
            const size_t o = ((size_t)xp * ny + j) * nzp + k0;
            ox1 = cp.off(xs, 0) + o;
            ox2 = cp.off(xs, 1) + o;
            px1 = ld4(cp.pe + ox1);
            px2 = ld4(cp.pe + ox2);
        }
        yt = (cp.axes & 2) ? sslab_e(j, ny, npml, ys, yp) : -1;  // y slab rows are all < ny-1
        if (yt > 0) {
            const size_t o = ((size_t)i * npml + yp) * nzp + k0;
            oy1 = cp.off(2 + ys, 0) + o;
            oy2 = cp.off(2 + ys, 1) + o;
            py1 = ld4(cp.pe + oy1);
            py2 = ld4(cp.pe + oy2);
        }
        if ((cp.axes & 4) && (k0 < npml || k0 + 3 >= nz - 1 - npml)) {
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
    // Drude box: J' of the three components (rows inside the box)
    const bool drow = DR && dr->row_in(i, j);
    Q4 jn0 = z4, jn1 = z4, jn2 = z4;
    float y1[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) y1[l] = l < 3 ? hy0.v[l + 1] : hyn;
    if (jy) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = e0.v[l] + c0.v[l] * ((hzU.v[l] - hz0.v[l]) - (y1[l] - hy0.v[l]));
            e0.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
    }
    if (drow) drude_quad(*dr, 0, i, j, k0, c0, e0, jn0);
    if (jy) {
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
    float x1[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) x1[l] = l < 3 ? hx0.v[l + 1] : hxn;
    if (ix) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = e1.v[l] + c1.v[l] * ((x1[l] - hx0.v[l]) - (hzN.v[l] - hz0.v[l]));
            e1.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
    }
    if (drow) drude_quad(*dr, 1, i, j, k0, c1, e1, jn1);
    if (ix) {
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
    }
    if (drow) drude_quad(*dr, 2, i, j, k0, c2, e2, jn2);
    if (ix && jy) {
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
    if (TF) tfsf_quad(*tf, 0, i, j, k0, e0, e1, e2);

    // --------------------------------------------------------------- stores
    // (unchanged lanes and wall rows store back the value they loaded)
    st4(exo + id, e0);
    st4(eyo + id, e1);
    st4(ezo + id, e2);
    if (drow) {
        drude_store(*dr, 0, i, j, k0, jn0);
        drude_store(*dr, 1, i, j, k0, jn1);
        drude_store(*dr, 2, i, j, k0, jn2);
    }
    if (PML) {
        if (xt > 0) {
            st4(cp.pe_out + ox1, px1);
            st4(cp.pe_out + ox2, px2);
        }
        if (yt > 0) {
            st4(cp.pe_out + oy1, py1);
            st4(cp.pe_out + oy2, py2);
        }
        if (anyz) {
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (zt[l] > 0) {
                    int zs, zp;
                    sslab_e(k0 + l, nz, npml, zs, zp);
                    float *q = cp.pe_out + cp.off(4 + zs, 0) + ((size_t)i * ny + j) * npml + zp;
                    if (jy) q[0] = pz1[l];
                    if (ix) q[cp.sZ] = pz2[l];
                }
        }
    }
    sprobe(prb, 0, i, j, k0, e0);
    sprobe(prb, 1, i, j, k0, e1);
    sprobe(prb, 2, i, j, k0, e2);
}

// E pass on quad f of the slab described by g (in place); CE as in e_quad_at (CM = 2:
// the material indices mx.. into tab).
template <int CM, bool PML>
__device__ __forceinline__ void e_quad(
    const SGeom &g, unsigned f, float *ex, float *ey, float *ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    const SCpml &cp, const SSrc &src, const SProbe &prb, const unsigned char *__restrict__ mx = nullptr,
    const unsigned char *__restrict__ my = nullptr, const unsigned char *__restrict__ mz = nullptr,
    const float *tab = nullptr)
{
    const unsigned row = f / g.Q;
    const int k0 = (int)(f - row * g.Q) * 4;
    const int di = (int)(row / (unsigned)g.ny);
    const int j = (int)row - di * g.ny;
    e_quad_at<CM, PML>(g, g.i0 + di, j, k0, ex, ey, ez, ex, ey, ez, hx, hy, hz, cex, cey, cez, cp, src, prb, mx, my,
                       mz, tab);
}

// H pass on the quad (i, j, k0..k0+3): reads H from hxi.. and writes it to hxo.. (the
// same arrays for an in-place update); psi H is updated in place (cp.ph).
template <bool PML>
__device__ __forceinline__ void h_quad_at(
    const SGeom &g, int i, int j, int k0, const float *__restrict__ ex, const float *__restrict__ ey,
    const float *__restrict__ ez, const float *hxi, const float *hyi, const float *hzi, float *hxo, float *hyo,
    float *hzo, const SCpml &cp, const SSrc &src, const SProbe &prb, const STfsf *tf = nullptr)
{
    const int nx = g.nx, ny = g.ny, nzp = g.nzp, nz = g.nz;
    (void)nx;
    const size_t id = ((size_t)i * ny + j) * nzp + k0;
    const size_t sx = (size_t)ny * nzp;
    const bool jp = j > 0, ip = i > 0;
    const Q4 z4 = splat4(0.f);

    // ---------------------------------------------------------------- loads
    const Q4 ex0 = ld4(ex + id), ey0 = ld4(ey + id), ez0 = ld4(ez + id);
    Q4 h0 = ld4(hxi + id), h1 = ld4(hyi + id), h2 = ld4(hzi + id);
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
        xt = ip && (cp.axes & 1) ? sslab_h(i, nx, npml, xs, xp) : -1;  // plane 0 is the wall
        if (xt > 0) {
            const size_t o = ((size_t)xp * ny + j) * nzp + k0;
            ppx1 = cp.ph + cp.off(xs, 0) + o;
            ppx2 = cp.ph + cp.off(xs, 1) + o;
            px1 = ld4(ppx1);
            px2 = ld4(ppx2);
        }
        yt = jp && (cp.axes & 2) ? sslab_h(j, ny, npml, ys, yp) : -1;  // row 0 is the wall
        if (yt > 0) {
            const size_t o = ((size_t)i * npml + yp) * nzp + k0;
            ppy1 = cp.ph + cp.off(2 + ys, 0) + o;
            ppy2 = cp.ph + cp.off(2 + ys, 1) + o;
            py1 = ld4(ppy1);
            py2 = ld4(ppy2);
        }
        if ((cp.axes & 4) && (k0 < npml || k0 + 3 >= nz - npml)) {
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
    if (tf) tfsf_quad(*tf, 3, i, j, k0, h0, h1, h2);

    // --------------------------------------------------------------- stores
    st4(hxo + id, h0);
    st4(hyo + id, h1);
    st4(hzo + id, h2);
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

// H pass on quad f of the slab described by g (in place).
template <bool PML>
__device__ __forceinline__ void h_quad(
    const SGeom &g, unsigned f, const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    float *hx, float *hy, float *hz, const SCpml &cp, const SSrc &src, const SProbe &prb)
{
    const unsigned row = f / g.Q;
    const int k0 = (int)(f - row * g.Q) * 4;
    const int di = (int)(row / (unsigned)g.ny);
    const int j = (int)row - di * g.ny;
    h_quad_at<PML>(g, g.i0 + di, j, k0, ex, ey, ez, hx, hy, hz, hx, hy, hz, cp, src, prb);
}

struct SFields {
    float *e[6];
    const float *ce[3];
    const unsigned char *cm[3];  // CE as material indices (kCeTable), else NULL
    const float *lut;            // their 3 x 256 table
};

}  // namespace mifdtd
