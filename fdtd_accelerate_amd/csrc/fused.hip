// Fused Yee step with ping-pong buffers: the state of step n is read from one buffer
// set and step n+1 written to the other.  One step is three launches:
//   1. shell E pass: E_new (core + CPML-E + E sources + probes) on every cell outside an
//      interior box, one thread per 4-cell quad (the per-quad arithmetic of the in-place
//      path, e_quad_at);
//   2. interior sweep: core E + E sources + core H + H sources + probes of the interior
//      box in ONE pass (E_new is consumed by the H update on chip), so each field and CE
//      value of the box is read from HBM once and written once: 60 B/cell (6 fields +
//      3 CE read, 6 fields written; 51 B with CE as uint8 material indices);
//   3. shell H pass: H_new (core + CPML-H + H sources + probes) outside the box
//      (h_quad_at).
// The interior box is the set of cells whose E and H updates have no CPML term:
// planes, rows and z cells [npml+1, n-1-npml) per axis (z rounded inward to 128-byte
// lines, 32 cells, so that the shell's z pieces and the sweep share no line).  Its
// E halo (plane xb-1, row yb-1) holds no CPML term either, so the sweep recomputes it;
// the E_new at z = kb-1 it needs comes from the shell E pass, and the shell H pass reads
// the box's E_new.  Without CPML the box is the whole grid and only the sweep runs.
//
// Sweep mapping (gfx950, 64-lane waves), detailed at the kernel:
//  * a work-group of T+1 waves owns T rows of an x chunk; wave w streams row j0-1+w
//    (w = 0 recomputes E_new of the row below the block: the y halo) plane by plane
//    (x) within one z segment of 256 floats (64 lanes x one 16-byte quad) at a time;
//  * H(i+1, j), loaded for E_new(i, j), stays in registers as H(i, j) of the next
//    plane; H(i, j+1) comes from the wave above through LDS; E_new(i-1, j) stays in
//    registers; E_new(i, j-1) comes from the wave below through LDS (one barrier per
//    plane, double-buffered exchange);
//  * z neighbours come from the adjacent lane (shuffles); across segments the k-1 value
//    is carried in LDS and the segment's last lane loads the k+1 element;
//  * block ids are remapped so that consecutive row blocks of a chunk run on the same
//    XCD, sharing L2 for the halo rows.
// Nothing a launch reads from memory is written by another thread of the same launch.
// Per-cell arithmetic order equals the in-place path (core, x, y, z CPML, sources)
// under -ffp-contract=off: values are bit-identical to it; only the sign of zeros
// stored in masked lanes (value*0 in the in-place path) may differ.

#include "common.hpp"
#include "step_quad.hpp"

#include <vector>

namespace mifdtd {

constexpr int kMaxRows = 15;  // owned rows per block (a block has rows+1 waves, at most 1024 threads)
constexpr size_t kMaxLds = 160 * 1024;  // LDS of one work-group on gfx950
constexpr int kMaxSrc = 4;
constexpr int kMaxProbe = 8;

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
    int T;       // rows per block
    int nrb;     // row blocks over [yb, ye)
    int cx;      // planes per chunk
    int nchunks; // chunks over [xb, xe)
    int xb, xe;  // planes of the interior box
    int yb, ye;  // rows of the interior box
    int kb, ke;  // z range of the interior box (multiples of 32)
    int zb0;     // first z segment start: kb rounded down to 256 (segments sit on 1 KiB)
    int nseg;    // z segments of 256 floats from zb0 covering [kb, ke)
    int nblocks;
};

struct Ctx {  // buffers of one launch (restrict-qualified at the kernel boundary)
    const float *ex, *ey, *ez, *hx, *hy, *hz;   // step n
    float *oex, *oey, *oez, *ohx, *ohy, *ohz;   // step n+1
    const float *cex, *cey, *cez;               // CE arrays (kCeArrays)
    const unsigned char *mx, *my, *mz;          // CE material indices (kCeTable)
};

// How CE reaches the E update: the constant 0.5, three float arrays, or three uint8
// material-index arrays into a 3 x 256 float table held in LDS (same float values, so
// the products are identical; 3 B per cell instead of 12).
enum CeMode { kCeConst = 0, kCeArrays = 1, kCeTable = 2 };

// CE of the quad at element rb + lc (rb uniform, lc the lane's element offset)
template <int CEM>
__device__ __forceinline__ Q4 load_ce(const float *cf, const unsigned char *mt, const float *tab, size_t rb, unsigned lc)
{
    if (CEM == kCeConst) return splat4(0.5f);
    if (CEM == kCeArrays) return ld4(cf + rb + lc);
    const uchar4 t = *reinterpret_cast<const uchar4 *>(mt + rb + lc);
    Q4 q;
    q.v[0] = tab[t.x];
    q.v[1] = tab[t.y];
    q.v[2] = tab[t.z];
    q.v[3] = tab[t.w];
    return q;
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
// is a uniform (scalar) row base -- array + (i*ny + j)*nzp + zb0 + 256*m -- plus the
// lane's element offset.
__device__ __forceinline__ size_t row_base(const FGeom &g, int i, int j, int m)
{
    return ((size_t)i * g.ny + j) * g.nzp + g.zb0 + (size_t)m * 256;
}

// E_new of row j of plane i for this lane's quad (segment m, lane offset l4 = 4*lane,
// k0 = zb0 + 256*m + l4; lc = l4 clamped into the box's part of the segment for lanes
// outside it, whose results are never stored).  hc = H(i,j), hux/huz = hx/hz of H(i,j+1), hN = H(i+1,j)
// (all of step n, zero outside the grid), hx_nx/hy_nx = hx/hy of the element after the
// segment for its last lane.  All memory loads of the row are issued first.  src: the
// block's source table (LDS) or null.
template <int CEM>
__device__ __forceinline__ void compute_E(const FGeom &g, const Ctx &c, const float *tab, const FSrc *src, int i, int j,
                                          int m, unsigned l4, unsigned lc, bool end_lane, const Q4 hc[3],
                                          const Q4 &hux, const Q4 &huz, const Q4 hN[3], float hx_nx, float hy_nx,
                                          Q4 &exn, Q4 &eyn, Q4 &ezn)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int k0 = g.zb0 + m * 256 + (int)l4;
    const size_t rb = row_base(g, i, j, m);
    const Q4 e0 = ld4(c.ex + rb + lc);
    const Q4 e1 = ld4(c.ey + rb + lc);
    const Q4 e2 = ld4(c.ez + rb + lc);
    const Q4 c0 = load_ce<CEM>(c.cex, c.mx, tab, rb, lc);
    const Q4 c1 = load_ce<CEM>(c.cey, c.my, tab + 256, rb, lc);
    const Q4 c2 = load_ce<CEM>(c.cez, c.mz, tab + 512, rb, lc);
    // k+1 neighbours of H(i,j)
    float hy1 = __shfl_down(hc[1].v[0], 1);
    float hx1 = __shfl_down(hc[0].v[0], 1);
    if (end_lane) {
        hy1 = hy_nx;
        hx1 = hx_nx;
    }
    // ex: rows j < ny-1 ; += cex*((hz[j+1]-hz) - (hy[k+1]-hy)) on lanes k < nz-1
    Q4 e = e0;
    if (j < ny - 1) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float y1 = l < 3 ? hc[1].v[l + 1] : hy1;
            const float v = e.v[l] + c0.v[l] * ((huz.v[l] - hc[2].v[l]) - (y1 - hc[1].v[l]));
            e.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
    }
    exn = e;
    // ey: planes i < nx-1 ; += cey*((hx[k+1]-hx) - (hz[i+1]-hz)) on lanes k < nz-1
    e = e1;
    if (i < nx - 1) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float x1 = l < 3 ? hc[0].v[l + 1] : hx1;
            const float v = e.v[l] + c1.v[l] * ((x1 - hc[0].v[l]) - (hN[2].v[l] - hc[2].v[l]));
            e.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
    }
    eyn = e;
    // ez: i < nx-1, j < ny-1 ; += cez*((hy[i+1]-hy) - (hx[j+1]-hx)) on lanes k < nz
    e = e2;
    if (i < nx - 1 && j < ny - 1) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = e.v[l] + c2.v[l] * ((hN[1].v[l] - hc[1].v[l]) - (hux.v[l] - hc[0].v[l]));
            e.v[l] = (k0 + l < nz) ? v : v * 0.f;
        }
    }
    ezn = e;
    if (src) {
        apply_src(*src, 0, i, j, k0, exn);
        apply_src(*src, 1, i, j, k0, eyn);
        apply_src(*src, 2, i, j, k0, ezn);
    }
}

// H_new of row j of plane i for this lane's quad: hc = H(i,j) of step n, exn/eyn/ezn =
// E_new(i,j), exb/ezb = E_new(i,j-1) (ex, ez), eyp/ezp = E_new(i-1,j) (ey, ez),
// exm/eym = E_new(i,j) at k0-1 (ex, ey).
__device__ __forceinline__ void compute_H(const FGeom &g, const FSrc *src, int i, int j, int m, unsigned l4,
                                          const Q4 hc[3], const Q4 &exn, const Q4 &eyn, const Q4 &ezn, const Q4 &exb,
                                          const Q4 &ezb, const Q4 &eyp, const Q4 &ezp, float exm, float eym, Q4 &hxo,
                                          Q4 &hyo, Q4 &hzo)
{
    const int nz = g.nz;
    const int k0 = g.zb0 + m * 256 + (int)l4;
    // hx: rows j > 0 ; -= 0.5*((ez - ez[j-1]) - (ey - ey[k-1])) on lanes 0 < k < nz
    hxo = hc[0];
    if (j > 0) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float ymv = l > 0 ? eyn.v[l - 1] : eym;
            const float v = hxo.v[l] - 0.5f * ((ezn.v[l] - ezb.v[l]) - (eyn.v[l] - ymv));
            const int k = k0 + l;
            hxo.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
        }
    }
    // hy: planes i > 0 ; -= 0.5*((ex - ex[k-1]) - (ez - ez[i-1])) on lanes 0 < k < nz
    hyo = hc[1];
    if (i > 0) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float xmv = l > 0 ? exn.v[l - 1] : exm;
            const float v = hyo.v[l] - 0.5f * ((exn.v[l] - xmv) - (ezn.v[l] - ezp.v[l]));
            const int k = k0 + l;
            hyo.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
        }
    }
    // hz: i > 0, j > 0 ; -= 0.5*((ey - ey[i-1]) - (ex - ex[j-1])) on lanes k < nz
    hzo = hc[2];
    if (i > 0 && j > 0) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float v = hzo.v[l] - 0.5f * ((eyn.v[l] - eyp.v[l]) - (exn.v[l] - exb.v[l]));
            hzo.v[l] = (k0 + l < nz) ? v : v * 0.f;
        }
    }
    if (src) {
        apply_src(*src, 3, i, j, k0, hxo);
        apply_src(*src, 4, i, j, k0, hyo);
        apply_src(*src, 5, i, j, k0, hzo);
    }
}

// Interior sweep (single pass, ping-pong buffers).  A block of T+1 waves owns rows
// j0 .. j0+nown-1 (nown = T, fewer in the last row block) of an x chunk [i0, i1): wave w
// takes row j = j0-1+w (w = 0 is the y halo row below the block, whose E_new only feeds
// the block's first row).  For each z segment m (256 floats: 64 lanes x one quad) the
// block marches the planes of the chunk, starting with the seed plane i0-1 (its E_new
// becomes E_new(i-1) of plane i0):
//   1. every wave loads H(i+1, j), which is also its H(i, j) of the next plane, takes
//      H(i, j+1) (hx, hz) from the row above -- published in LDS one plane earlier, or
//      loaded by the top own row -- and computes E_new(i, j) (core + E sources);
//   2. it publishes ex, ez of E_new(i, j) for the row above and hx, hz of H(i+1, j) for
//      the row below in LDS (double-buffered: the buffer alternates with every plane,
//      across segments too); lane 63 leaves ex, ey at k = zb0+256m+255 for lane 0 of
//      segment m+1 (per plane and row, in LDS);
//   3. one barrier;
//   4. the owned rows compute H_new(i, j) (core + H sources) from E_new(i, j),
//      E_new(i, j-1) (LDS), E_new(i-1, j) (registers) and the k-1 values (lane shuffles,
//      lane 0 from the carry, or on the first segment of a box starting inside the row
//      from the output, where the shell E pass stored it) and store E_new, H_new and
//      the probes.
// Each H value of step n is thus read from memory once per block (plus the top row's
// neighbour row and the k+1 element after each segment); E and CE once.  The first plane
// of a segment loads H(i, j) and H(i, j+1) directly.  Loads are unconditional (lanes
// past the segment end read a clamped in-segment address and store nothing) so that a
// plane's loads issue together.  The LDS exchange is ordered by the barrier and the
// parity.  Source and probe tables are copied to LDS once per block and consulted only
// by blocks whose rows they touch.
template <int CEM>
__global__ __launch_bounds__(1024, 4) void k_fused(
    FGeom g,
    const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    float *__restrict__ oex, float *__restrict__ oey, float *__restrict__ oez,
    float *__restrict__ ohx, float *__restrict__ ohy, float *__restrict__ ohz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    const unsigned char *__restrict__ mx, const unsigned char *__restrict__ my,
    const unsigned char *__restrict__ mz, const float *__restrict__ lut, FSrc src, FProbe prb)
{
    const Ctx c{ex, ey, ez, hx, hy, hz, oex, oey, oez, ohx, ohy, ohz, cex, cey, cez, mx, my, mz};
    extern __shared__ float lds[];
    __shared__ FSrc s_src;
    __shared__ FProbe s_prb;
    const int T = g.T;
    float *tab = lds;                                 // [3][256] CE table (kCeTable)
    float *ebuf = lds + (CEM == kCeTable ? 768 : 0);  // [2][T][2][256]: ex, ez of rows w = 0..T-1
    float *hbuf = ebuf + 2 * T * 512;                 // [2][T][2][256]: hx, hz of H(i+1) of rows w = 1..T
    float *cbuf = hbuf + 2 * T * 512;                 // [cx+1][T+1][2]: ex, ey at the segment's last cell

    // XCD-aware block order: consecutive logical blocks (neighbouring row blocks of a
    // chunk, which share halo rows) land on the same XCD.
    int bid = blockIdx.x;
    {
        const int nb = g.nblocks, nb8 = nb & ~7;
        if (bid < nb8) bid = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    }
    const int rb = bid % g.nrb;
    const int chunk = bid / g.nrb;
    const int j0 = g.yb + rb * T;
    const int nown = min(T, g.ye - j0);  // owned rows of this block
    const int i0 = g.xb + chunk * g.cx;
    const int i1 = min(g.xe, i0 + g.cx);
    const int ib = i0 > 0 ? i0 - 1 : i0;
    // does a source / probe touch this block's rows and planes?
    bool src_hit = false, prb_hit = false;
    for (int n = 0; n < kMaxSrc; ++n)
        src_hit |= n < src.n && src.i1[n] >= ib && src.i0[n] < i1 && src.j1[n] >= j0 - 1 && src.j0[n] <= j0 + nown - 1;
    for (int n = 0; n < kMaxProbe; ++n)
        prb_hit |= n < prb.n && prb.i[n] >= i0 && prb.i[n] < i1 && prb.j[n] >= j0 && prb.j[n] <= j0 + nown - 1;
    if (threadIdx.x == 0) {
        if (src_hit) s_src = src;
        if (prb_hit) s_prb = prb;
    }
    if (CEM == kCeTable)
        for (int t = threadIdx.x; t < 768; t += blockDim.x) tab[t] = lut[t];
    __syncthreads();
    const FSrc *sp = src_hit ? &s_src : nullptr;

    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const unsigned l4 = 4u * (unsigned)lane;
    const int j = j0 - 1 + w;
    const bool live = w <= nown && j >= 0;  // rows above the block's own rows idle
    const bool own = w >= 1 && live;
    const bool up_live = j + 1 < g.ny;      // row j+1 exists (j >= -1)
    const bool up_lds = w < nown;           // H(i, j+1) comes from the wave above

    int it = 0;  // planes exchanged so far: the LDS parity alternates across segments too
    for (int m = 0; m < g.nseg; ++m) {
        // the box's part [lo, hi) of segment m (multiples of 4; lanes outside it idle)
        const int base = g.zb0 + m * 256;
        const int lo = max(g.kb - base, 0), hi = min(g.ke - base, 256);
        const bool act = (int)l4 >= lo && (int)l4 < hi;
        const unsigned lc = (int)l4 < lo ? (unsigned)lo : (int)l4 >= hi ? (unsigned)(hi - 4) : l4;
        const bool start_lane = (int)l4 == lo;    // holds the first quad: k-1 from outside
        const bool end_lane = (int)l4 + 4 == hi;  // holds the last quad: k+1 from outside
        // the element after the last cell (its k+1), when inside the row
        const int next_off = base + hi < g.nzp ? hi : 0;
        // first cell of a box starting inside the row: E_new(kb-1) from the output
        // (stored by the shell E pass before this launch)
        const bool carry_mem = m == 0 && g.kb > 0 && own && start_lane;
        Q4 eyp = zero4(), ezp = zero4();  // E_new(i-1, j)
        Q4 hc[3], hux = zero4(), huz = zero4();
        float hx_nx = 0.f, hy_nx = 0.f;  // hx, hy of the element after the segment
        hc[0] = hc[1] = hc[2] = zero4();
        if (live) {
            const size_t r = row_base(g, ib, j, m);
            hc[0] = ld4(hx + r + lc);
            hc[1] = ld4(hy + r + lc);
            hc[2] = ld4(hz + r + lc);
            if (next_off) {
                hx_nx = hx[r + next_off];
                hy_nx = hy[r + next_off];
            }
            if (up_live) {
                const size_t ru = row_base(g, ib, j + 1, m);
                hux = ld4(hx + ru + lc);
                huz = ld4(hz + ru + lc);
            }
        }
        for (int i = ib; i < i1; ++i) {
            const bool seed = i < i0;
            const int par = it & 1;
            ++it;
            Q4 hN[3];
            hN[0] = hN[1] = hN[2] = zero4();
            float hxN_nx = 0.f, hyN_nx = 0.f;
            if (live && i + 1 < g.nx) {
                const size_t r = row_base(g, i + 1, j, m);
                hN[0] = ld4(hx + r + lc);
                hN[1] = ld4(hy + r + lc);
                hN[2] = ld4(hz + r + lc);
                if (next_off) {
                    hxN_nx = hx[r + next_off];
                    hyN_nx = hy[r + next_off];
                }
            }
            float exm_mem = 0.f, eym_mem = 0.f;
            if (carry_mem && !seed) {
                const size_t r = row_base(g, i, j, 0) + lo - 1;
                exm_mem = oex[r];
                eym_mem = oey[r];
            }
            if (i > ib) {  // H(i, j+1): from the row above (LDS) or, for the top row, memory
                if (up_lds) {
                    const float *hb = hbuf + (size_t)(((par ^ 1) * T + w) * 2) * 256;
                    hux = Q4{{hb[l4], hb[l4 + 1], hb[l4 + 2], hb[l4 + 3]}};
                    huz = Q4{{hb[256 + l4], hb[256 + l4 + 1], hb[256 + l4 + 2], hb[256 + l4 + 3]}};
                } else if (live && up_live) {
                    const size_t ru = row_base(g, i, j + 1, m);
                    hux = ld4(hx + ru + lc);
                    huz = ld4(hz + ru + lc);
                } else {
                    hux = huz = zero4();
                }
            }
            Q4 exn = zero4(), eyn = zero4(), ezn = zero4();
            if (live)
                compute_E<CEM>(g, c, tab, sp, i, j, m, l4, lc, end_lane, hc, hux, huz, hN, hx_nx, hy_nx, exn, eyn, ezn);
            if (w < T) {
                float *es = ebuf + (size_t)((par * T + w) * 2) * 256;
                *reinterpret_cast<float4 *>(es + l4) = make_float4(exn.v[0], exn.v[1], exn.v[2], exn.v[3]);
                *reinterpret_cast<float4 *>(es + 256 + l4) = make_float4(ezn.v[0], ezn.v[1], ezn.v[2], ezn.v[3]);
            }
            if (w > 0) {
                float *hs = hbuf + (size_t)((par * T + (w - 1)) * 2) * 256;
                *reinterpret_cast<float4 *>(hs + l4) = make_float4(hN[0].v[0], hN[0].v[1], hN[0].v[2], hN[0].v[3]);
                *reinterpret_cast<float4 *>(hs + 256 + l4) =
                    make_float4(hN[2].v[0], hN[2].v[1], hN[2].v[2], hN[2].v[3]);
            }
            // k-1 neighbours: from the lane below; the first quad of a segment from lane 63
            // of the previous segment at this plane (this wave wrote it there), from
            // memory on a box starting inside the row, else k-1 < 0 feeds masked lanes
            float *cs = cbuf + (size_t)(((i - ib) * (T + 1) + w) * 2);
            float exm = __shfl_up(exn.v[3], 1);
            float eym = __shfl_up(eyn.v[3], 1);
            if (start_lane) {
                exm = m > 0 ? cs[0] : exm_mem;
                eym = m > 0 ? cs[1] : eym_mem;
            }
            __syncthreads();
            if (lane == 63) {
                cs[0] = exn.v[3];
                cs[1] = eyn.v[3];
            }
            if (own && !seed) {
                const float *xb = ebuf + (size_t)((par * T + (w - 1)) * 2) * 256;
                const Q4 exb = Q4{{xb[l4], xb[l4 + 1], xb[l4 + 2], xb[l4 + 3]}};
                const Q4 ezb = Q4{{xb[256 + l4], xb[256 + l4 + 1], xb[256 + l4 + 2], xb[256 + l4 + 3]}};
                Q4 hxo, hyo, hzo;
                compute_H(g, sp, i, j, m, l4, hc, exn, eyn, ezn, exb, ezb, eyp, ezp, exm, eym, hxo, hyo, hzo);
                const size_t rbase = row_base(g, i, j, m);
                if (act) {
                    st4(oex + rbase + l4, exn);
                    st4(oey + rbase + l4, eyn);
                    st4(oez + rbase + l4, ezn);
                    st4(ohx + rbase + l4, hxo);
                    st4(ohy + rbase + l4, hyo);
                    st4(ohz + rbase + l4, hzo);
                }
                if (prb_hit) {
                    const int k0 = g.zb0 + m * 256 + (int)l4;
                    for (int n = 0; n < s_prb.n; ++n) {
                        if (act && s_prb.i[n] == i && s_prb.j[n] == j) {
                            const int cc = s_prb.comp[n];
#pragma unroll
                            for (int l = 0; l < 4; ++l) {
                                if (s_prb.k[n] == k0 + l) {
                                    const float v = cc == 0 ? exn.v[l] : cc == 1 ? eyn.v[l] : cc == 2 ? ezn.v[l]
                                                  : cc == 3 ? hxo.v[l] : cc == 4 ? hyo.v[l] : hzo.v[l];
                                    s_prb.out[n] = v;
                                }
                            }
                        }
                    }
                }
            }
            eyp = eyn;
            ezp = ezn;
            hc[0] = hN[0];
            hc[1] = hN[1];
            hc[2] = hN[2];
            hx_nx = hxN_nx;
            hy_nx = hyN_nx;
        }
    }
}

// CPML shell: every cell outside the interior box, as up to six boxes (x low/high: all
// rows; y low/high: the box's planes; z low/high: the box's planes and rows), each
// (planes [i0,i1), rows [j0,j1), quads [q0,q1)) with a thread per quad; start[b] is the
// first thread of box b.  Quads run fastest, then the box's shorter one of planes and
// rows, so that the i+1 / i-1 and j+1 / j-1 neighbour rows a quad reads are read by
// nearby threads too (L2 hits).
struct ShellBoxes {
    int n;
    int i0[6], i1[6], j0[6], j1[6], q0[6], q1[6];
    int i_inner[6];  // planes before rows in the thread order
    unsigned start[7];
};

__device__ __forceinline__ void shell_quad(const ShellBoxes &s, unsigned f, int &i, int &j, int &k0)
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
    i = s.i0[b] + (ii ? a : o);
    j = s.j0[b] + (ii ? o : a);
    k0 = 4 * (s.q0[b] + (int)q);
}

// CE: CEM as in the sweep (kCeTable stages the 3 x 256 table in LDS per block).
template <int CEM>
__global__ __launch_bounds__(256) void k_shell_e(SGeom g, ShellBoxes sb, unsigned total, const float *ex, const float *ey,
                                                 const float *ez, float *oex, float *oey, float *oez,
                                                 const float *__restrict__ hx, const float *__restrict__ hy,
                                                 const float *__restrict__ hz, const float *__restrict__ cex,
                                                 const float *__restrict__ cey, const float *__restrict__ cez,
                                                 const unsigned char *__restrict__ mx,
                                                 const unsigned char *__restrict__ my,
                                                 const unsigned char *__restrict__ mz, const float *__restrict__ lut,
                                                 SCpml cp, SSrc src, SProbe prb)
{
    __shared__ float tab[CEM == kCeTable ? 768 : 1];
    if (CEM == kCeTable) {
        for (int t = threadIdx.x; t < 768; t += 256) tab[t] = lut[t];
        __syncthreads();
    }
    const unsigned f = blockIdx.x * 256u + threadIdx.x;
    if (f >= total) return;
    int i, j, k0;
    shell_quad(sb, f, i, j, k0);
    e_quad_at<CEM, true>(g, i, j, k0, ex, ey, ez, oex, oey, oez, hx, hy, hz, cex, cey, cez, cp, src, prb, mx, my, mz,
                         tab);
}

__global__ __launch_bounds__(256) void k_shell_h(SGeom g, ShellBoxes sb, unsigned total, const float *__restrict__ oex,
                                                 const float *__restrict__ oey, const float *__restrict__ oez,
                                                 const float *hx, const float *hy, const float *hz, float *ohx,
                                                 float *ohy, float *ohz, SCpml cp, SSrc src, SProbe prb)
{
    const unsigned f = blockIdx.x * 256u + threadIdx.x;
    if (f >= total) return;
    int i, j, k0;
    shell_quad(sb, f, i, j, k0);
    h_quad_at<true>(g, i, j, k0, oex, oey, oez, hx, hy, hz, ohx, ohy, ohz, cp, src, prb);
}

}  // namespace mifdtd

// =========================================================================== C ABI
using namespace mifdtd;

extern "C" {

int mifdtd_fused_run_timed(const mifdtd_fused_desc *d, int nsteps, const float *tvals, float *probe_series,
                           int *parity, mifdtd_stream_t stream, float *sweep_ms)
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
    const bool tbl = d->ce_mat[0] != nullptr;
    if (tbl) {
        if (!d->ce_mat[1] || !d->ce_mat[2] || !d->ce_lut)
            return fail(MIFDTD_ENULL, "CE material indices need all three index arrays and the table");
        for (int q = 0; q < 3; q++)
            if (reinterpret_cast<uintptr_t>(d->ce_mat[q]) & 3u)
                return fail(MIFDTD_EALIGN, "CE material index arrays not 4-byte aligned");
    }
    const int cem = tbl ? kCeTable : hc ? kCeArrays : kCeConst;
    const int nz = nzp - pad;
    if (d->nsrc < 0 || d->nsrc > MIFDTD_FUSED_MAX_SRC) return fail(MIFDTD_EINVAL, "nsrc outside [0,%d]", MIFDTD_FUSED_MAX_SRC);
    if (d->nprobe < 0 || d->nprobe > MIFDTD_FUSED_MAX_PROBE) return fail(MIFDTD_EINVAL, "nprobe outside [0,%d]", MIFDTD_FUSED_MAX_PROBE);
    if (d->nprobe && !probe_series) return fail(MIFDTD_ENULL, "probe_series is NULL");
    if (d->nsrc && !tvals) return fail(MIFDTD_ENULL, "tvals is NULL");

    const bool pml = d->npml > 0;
    SCpml scp{};
    if (pml) {
        const int np_ = d->npml;
        if (np_ > MIFDTD_MAX_NPML) return fail(MIFDTD_EINVAL, "npml %d > %d", np_, MIFDTD_MAX_NPML);
        if (nx < 2 * np_ + 2 || ny < 2 * np_ + 2 || nz < 2 * np_ + 2)
            return fail(MIFDTD_EINVAL, "grid too small for npml %d", np_);
        if (!d->bE || !d->aE || !d->bH || !d->aH) return fail(MIFDTD_ENULL, "CPML tables are NULL");
        if (!d->psiE || !d->psiE_alt || !d->psiH) return fail(MIFDTD_ENULL, "packed psi buffers are NULL");
        if (!aligned16(d->psiE) || !aligned16(d->psiE_alt) || !aligned16(d->psiH))
            return fail(MIFDTD_EALIGN, "packed psi buffers not 16-byte aligned");
        scp.npml = np_;
        scp.bE = d->bE; scp.aE = d->aE; scp.bH = d->bH; scp.aH = d->aH;
        scp.ph = d->psiH;
        scp.sX = (size_t)np_ * ny * nzp;
        scp.sY = (size_t)nx * np_ * nzp;
        scp.sZ = (size_t)nx * ny * np_;
    }

    // interior box: no CPML term in its E and H updates nor in its E halo
    FGeom g;
    g.nx = nx; g.ny = ny; g.nzp = nzp; g.nz = nz;
    g.T = T;
    g.cx = CX;
    if (pml) {
        const int np_ = d->npml;
        g.xb = np_ + 1; g.xe = nx - 1 - np_;
        g.yb = np_ + 1; g.ye = ny - 1 - np_;
        g.kb = 32 * ((np_ + 1 + 31) / 32); g.ke = 32 * ((nz - 1 - np_) / 32);
        if (g.xb >= g.xe || g.yb >= g.ye || g.kb >= g.ke) {  // no interior: the shell is the grid
            g.xb = g.xe = nx;
            g.yb = g.ye = ny;
            g.kb = g.ke = nzp;
        }
    } else {
        g.xb = 0; g.xe = nx;
        g.yb = 0; g.ye = ny;
        g.kb = 0; g.ke = nzp;
    }
    const bool sweep = g.xe > g.xb;
    g.nrb = sweep ? (g.ye - g.yb + T - 1) / T : 0;
    g.nchunks = sweep ? (g.xe - g.xb + CX - 1) / CX : 0;
    g.zb0 = 256 * (g.kb / 256);
    g.nseg = sweep ? (g.ke - g.zb0 + 255) / 256 : 0;
    const long long nbl = (long long)g.nrb * g.nchunks;
    if (nbl >= (1ll << 31)) return fail(MIFDTD_ERANGE, "too many blocks");
    g.nblocks = (int)nbl;
    const size_t lds_bytes = sizeof(float) * ((cem == kCeTable ? 768 : 0) + (size_t)2 * 2 * T * 2 * 256 +
                                              (size_t)(CX + 1) * (T + 1) * 2);
    if (lds_bytes + sizeof(FSrc) + sizeof(FProbe) > kMaxLds)
        return fail(MIFDTD_EINVAL, "rows_per_thread %d / planes_per_chunk %d need %zu B of LDS (> %zu)", T, CX, lds_bytes,
                    kMaxLds);

    // shell boxes (only with CPML)
    ShellBoxes sb{};
    unsigned long long total = 0;
    if (pml) {
        const int Q = nzp / 4;
        auto add = [&](int i0, int i1, int j0, int j1, int q0, int q1) {
            if (i1 <= i0 || j1 <= j0 || q1 <= q0) return;
            const int b = sb.n++;
            sb.i0[b] = i0; sb.i1[b] = i1; sb.j0[b] = j0; sb.j1[b] = j1; sb.q0[b] = q0; sb.q1[b] = q1;
            sb.i_inner[b] = (i1 - i0) < (j1 - j0);
            sb.start[b] = (unsigned)total;
            total += (unsigned long long)(i1 - i0) * (j1 - j0) * (q1 - q0);
        };
        add(0, g.xb, 0, ny, 0, Q);
        add(g.xe, nx, 0, ny, 0, Q);
        add(g.xb, g.xe, 0, g.yb, 0, Q);
        add(g.xb, g.xe, g.ye, ny, 0, Q);
        add(g.xb, g.xe, g.yb, g.ye, 0, g.kb / 4);
        add(g.xb, g.xe, g.yb, g.ye, g.ke / 4, Q);
        if (total >= (1ull << 32) - 256) return fail(MIFDTD_ERANGE, "CPML shell too large");
        sb.start[sb.n] = (unsigned)total;
    }
    SGeom sg{};
    sg.nx = nx; sg.ny = ny; sg.nzp = nzp; sg.nz = nz;

    FSrc src{};
    SSrc ssrc{};
    src.n = ssrc.n = d->nsrc;
    for (int n = 0; n < d->nsrc; n++) {
        const int *b = d->src_box[n];
        if (d->src_comp[n] < 0 || d->src_comp[n] > 5) return fail(MIFDTD_EINVAL, "source component %d", d->src_comp[n]);
        if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] >= nx || b[3] >= ny || b[5] >= nzp || b[0] > b[1] || b[2] > b[3] || b[4] > b[5])
            return fail(MIFDTD_EINVAL, "source box %d outside the grid", n);
        src.comp[n] = ssrc.comp[n] = d->src_comp[n];
        src.i0[n] = ssrc.i0[n] = b[0]; src.i1[n] = ssrc.i1[n] = b[1];
        src.j0[n] = ssrc.j0[n] = b[2]; src.j1[n] = ssrc.j1[n] = b[3];
        src.k0[n] = ssrc.k0[n] = b[4]; src.k1[n] = ssrc.k1[n] = b[5];
        src.sv[n] = ssrc.sv[n] = d->src_sv[n];
        src.svs[n] = ssrc.svs[n] = d->src_svs[n];
        src.ow[n] = ssrc.ow[n] = d->src_overwrite[n];
    }
    FProbe prb{};
    SProbe sprb{};
    prb.n = sprb.n = d->nprobe;
    for (int n = 0; n < d->nprobe; n++) {
        const int *p = d->probe_ijk[n];
        if (d->probe_comp[n] < 0 || d->probe_comp[n] > 5 || p[0] < 0 || p[0] >= nx || p[1] < 0 || p[1] >= ny || p[2] < 0 || p[2] >= nzp)
            return fail(MIFDTD_EINVAL, "probe %d outside the grid", n);
        prb.comp[n] = sprb.comp[n] = d->probe_comp[n];
        prb.i[n] = sprb.i[n] = p[0];
        prb.j[n] = sprb.j[n] = p[1];
        prb.k[n] = sprb.k[n] = p[2];
        sprb.slot[n] = n;
    }

    const float *in[6];
    float *out[6];
    for (int q = 0; q < 6; q++) { in[q] = d->e[q]; out[q] = d->e_alt[q]; }
    float *pin = d->psiE, *pout = d->psiE_alt;

    hipStream_t st = to_stream(stream);
    const unsigned nthr = 64u * (unsigned)(T + 1);
    // optional HIP events around every sweep launch (the dominant kernel of the step),
    // destroyed on every return path
    struct Marks {
        std::vector<hipEvent_t> v;
        ~Marks() { for (hipEvent_t x : v) (void)hipEventDestroy(x); }
    } mk;
    std::vector<hipEvent_t> &marks = mk.v;
    if (sweep_ms) {
        *sweep_ms = 0.f;
        marks.reserve(2 * (size_t)nsteps);
        for (int q = 0; q < 2 * nsteps; q++) {
            hipEvent_t e;  // timing only: no system-scope fence (no cache writeback) on record
            if (hipError_t err = hipEventCreateWithFlags(&e, hipEventDisableSystemFence))
                return fail((int)err, "hipEventCreate failed");
            marks.push_back(e);
        }
    }
    const unsigned shell_blocks = (unsigned)((total + 255) / 256);
    for (int s = 0; s < nsteps; s++) {
        for (int n = 0; n < src.n; n++) ssrc.tval[n] = src.tval[n] = tvals[(size_t)s * src.n + n];
        sprb.out = prb.out = d->nprobe ? probe_series + (size_t)s * d->nprobe : nullptr;
        if (pml) {
            scp.pe = pin;
            scp.pe_out = pout;
#define MIFDTD_SHELL_E_LAUNCH(CEV)                                                                               \
    k_shell_e<CEV><<<shell_blocks, 256, 0, st>>>(sg, sb, (unsigned)total, in[0], in[1], in[2], out[0], out[1],   \
                                                 out[2], in[3], in[4], in[5], d->ce[0], d->ce[1], d->ce[2],      \
                                                 d->ce_mat[0], d->ce_mat[1], d->ce_mat[2], d->ce_lut, scp, ssrc, \
                                                 sprb)
            if (cem == kCeTable) MIFDTD_SHELL_E_LAUNCH(kCeTable);
            else if (cem == kCeArrays) MIFDTD_SHELL_E_LAUNCH(kCeArrays);
            else MIFDTD_SHELL_E_LAUNCH(kCeConst);
#undef MIFDTD_SHELL_E_LAUNCH
            MIFDTD_CHECK_LAUNCH("CPML shell E launch");
        }
        if (sweep) {
            if (sweep_ms) (void)hipEventRecord(marks[2 * (size_t)s], st);
#define MIFDTD_FUSED_LAUNCH(CEV)                                                                                 \
    k_fused<CEV><<<g.nblocks, nthr, lds_bytes, st>>>(g, in[0], in[1], in[2], in[3], in[4], in[5], out[0], out[1], \
                                                     out[2], out[3], out[4], out[5], d->ce[0], d->ce[1], d->ce[2], \
                                                     d->ce_mat[0], d->ce_mat[1], d->ce_mat[2], d->ce_lut, src, prb)
            if (cem == kCeTable) MIFDTD_FUSED_LAUNCH(kCeTable);
            else if (cem == kCeArrays) MIFDTD_FUSED_LAUNCH(kCeArrays);
            else MIFDTD_FUSED_LAUNCH(kCeConst);
#undef MIFDTD_FUSED_LAUNCH
            if (sweep_ms) (void)hipEventRecord(marks[2 * (size_t)s + 1], st);
            MIFDTD_CHECK_LAUNCH("fused sweep launch");
        }
        if (pml) {
            k_shell_h<<<shell_blocks, 256, 0, st>>>(sg, sb, (unsigned)total, out[0], out[1], out[2], in[3], in[4], in[5],
                                                    out[3], out[4], out[5], scp, ssrc, sprb);
            MIFDTD_CHECK_LAUNCH("CPML shell H launch");
        }
        for (int q = 0; q < 6; q++) { const float *x = in[q]; in[q] = out[q]; out[q] = const_cast<float *>(x); }
        float *tp = pin;
        pin = pout;
        pout = tp;
    }
    if (parity) *parity = nsteps & 1;
    if (sweep_ms) {
        float total_ms = 0.f;
        hipError_t err = hipStreamSynchronize(st);
        for (int q = 0; sweep && err == hipSuccess && q < nsteps; q++) {
            float t = 0.f;
            err = hipEventElapsedTime(&t, marks[2 * (size_t)q], marks[2 * (size_t)q + 1]);
            total_ms += t;
        }
        if (err != hipSuccess) return fail((int)err, "sweep timing events failed");
        *sweep_ms = nsteps > 0 ? total_ms / nsteps : 0.f;
    }
    return 0;
}

int mifdtd_fused_run(const mifdtd_fused_desc *d, int nsteps, const float *tvals, float *probe_series,
                     int *parity, mifdtd_stream_t stream)
{
    return mifdtd_fused_run_timed(d, nsteps, tvals, probe_series, parity, stream, nullptr);
}

}  // extern "C"
