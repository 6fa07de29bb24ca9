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
// line boundaries of the field arrays, so that the shell's z pieces and the sweep share
// no line; when the arrays start a multiple of 16 bytes past a line, the boundaries sit
// at k = 32m - phase, and the caller can place them close to the CPML).  Its
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
// Step pairs (a third buffer set given): the two-step sweep k_fused2 below advances a box
// two cells further from the CPML slabs by two steps per pass, between shell passes of
// one step each (host code of mifdtd_fused_run_timed).
// Nothing a launch reads from memory is written by another thread of the same launch.
// Per-cell arithmetic order equals the in-place path (core, x, y, z CPML, sources)
// under -ffp-contract=off: values are bit-identical to it; only the sign of zeros
// stored in masked lanes (value*0 in the in-place path) may differ.

#include "common.hpp"
#include "step_quad.hpp"
#include "shell_boxes.hpp"

#include <vector>

namespace mifdtd {

constexpr int kMaxRows = 15;  // owned rows per block (a block has rows+1 waves, at most 1024 threads)
// Up to this many rows the sweep is compiled for its real block of rows+1 <= 10 waves:
// one such work-group per CU either way, and the register budget of 3 waves per SIMD
// (168 VGPRs) holds the kernel without the scratch spill of a 1024-thread bound.
constexpr int kNarrowRows = 9;
constexpr int kDrudeMaxRows = 7;  // with a Drude box: rows+1 waves at 2 per SIMD (up to 256 VGPRs)
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


// Floats of the CE table at the start of the sweep's LDS (a multiple of 4: the row
// exchange buffers after it take 16-byte stores).
__host__ __device__ constexpr int ce_tab_floats(int cem, int ntri)
{
    return cem == kCeTable ? 768 : cem == kCeTriple ? 4 * ((3 * ntri + 3) / 4) : 0;
}

struct Ctx {  // buffers of one launch (restrict-qualified at the kernel boundary)
    const float *ex, *ey, *ez, *hx, *hy, *hz;   // step n
    float *oex, *oey, *oez, *ohx, *ohy, *ohz;   // step n+1
    const float *cex, *cey, *cez;               // CE arrays (kCeArrays)
    const unsigned char *mx, *my, *mz;          // CE material indices (kCeTable)
    const unsigned short *mt;                   // CE triple indices (kCeTriple)
};

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
template <int CEM, int FEAT = 0>
__device__ __forceinline__ void compute_E(const FGeom &g, const Ctx &c, const float *tab, const FSrc *src, int i, int j,
                                          int m, unsigned l4, unsigned lc, bool end_lane, const Q4 hc[3],
                                          const Q4 &hux, const Q4 &huz, const Q4 hN[3], float hx_nx, float hy_nx,
                                          Q4 &exn, Q4 &eyn, Q4 &ezn, const SDrude *dr = nullptr, Q4 *jn = nullptr,
                                          const STfsf *tf = nullptr, unsigned tmask = 0)
{
    constexpr bool DR = (FEAT & kFeatDrude) != 0;
    constexpr bool TF = (FEAT & kFeatTfsf) != 0;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int k0 = g.zb0 + m * 256 + (int)l4;
    const size_t rb = row_base(g, i, j, m);
    const Q4 e0 = ld4(c.ex + rb + lc);
    const Q4 e1 = ld4(c.ey + rb + lc);
    const Q4 e2 = ld4(c.ez + rb + lc);
    Q4 c0, c1, c2;
    if (CEM == kCeTriple) {  // one uint16 per cell into (cex, cey, cez) triples
        const uint2 u = *reinterpret_cast<const uint2 *>(c.mt + rb + lc);
        const unsigned t[4] = {u.x & 0xffffu, u.x >> 16, u.y & 0xffffu, u.y >> 16};
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            c0.v[l] = tab[3 * t[l]];
            c1.v[l] = tab[3 * t[l] + 1];
            c2.v[l] = tab[3 * t[l] + 2];
        }
    } else {
        c0 = load_ce<CEM>(c.cex, c.mx, tab, rb, lc);
        c1 = load_ce<CEM>(c.cey, c.my, tab + 256, rb, lc);
        c2 = load_ce<CEM>(c.cez, c.mz, tab + 512, rb, lc);
    }
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
    if (DR && dr->row_in(i, j)) {  // Drude box: after the core update, before the sources
        drude_quad(*dr, 0, i, j, k0, c0, exn, jn[0]);
        drude_quad(*dr, 1, i, j, k0, c1, eyn, jn[1]);
        drude_quad(*dr, 2, i, j, k0, c2, ezn, jn[2]);
    }
    if (src) {
        apply_src(*src, 0, i, j, k0, exn);
        apply_src(*src, 1, i, j, k0, eyn);
        apply_src(*src, 2, i, j, k0, ezn);
    }
    if (TF && tmask) tfsf_quad(*tf, 0, i, j, k0, exn, eyn, ezn, tmask);  // TFSF E patches after the sources
}

// H_new of row j of plane i for this lane's quad: hc = H(i,j) of step n, exn/eyn/ezn =
// E_new(i,j), exb/ezb = E_new(i,j-1) (ex, ez), eyp/ezp = E_new(i-1,j) (ey, ez),
// exm/eym = E_new(i,j) at k0-1 (ex, ey).
template <bool TF = false>
__device__ __forceinline__ void compute_H(const FGeom &g, const FSrc *src, int i, int j, int m, unsigned l4,
                                          const Q4 hc[3], const Q4 &exn, const Q4 &eyn, const Q4 &ezn, const Q4 &exb,
                                          const Q4 &ezb, const Q4 &eyp, const Q4 &ezp, float exm, float eym, Q4 &hxo,
                                          Q4 &hyo, Q4 &hzo, const STfsf *tf = nullptr, unsigned tmask = 0)
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
    if (TF && tmask) tfsf_quad(*tf, 3, i, j, k0, hxo, hyo, hzo, tmask);  // TFSF H patches after the sources
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
template <int CEM, int FEAT, bool WIDE>
__global__ __launch_bounds__((FEAT & kFeatDrude) ? 64 * (kDrudeMaxRows + 1) : WIDE ? 1024 : 64 * (kNarrowRows + 1),
                             (FEAT & kFeatDrude) ? 2 : WIDE ? 4 : 3) void k_fused(
    FGeom g,
    const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    float *__restrict__ oex, float *__restrict__ oey, float *__restrict__ oez,
    float *__restrict__ ohx, float *__restrict__ ohy, float *__restrict__ ohz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    const unsigned char *__restrict__ mx, const unsigned char *__restrict__ my,
    const unsigned char *__restrict__ mz, const float *__restrict__ lut, FSrc src, FProbe prb, SDrude dr,
    STfsf tfe, STfsf tfh, const unsigned short *__restrict__ mt)
{
    constexpr bool DR = (FEAT & kFeatDrude) != 0;
    constexpr bool TF = (FEAT & kFeatTfsf) != 0;
    const Ctx c{ex, ey, ez, hx, hy, hz, oex, oey, oez, ohx, ohy, ohz, cex, cey, cez, mx, my, mz, mt};
    extern __shared__ float lds[];
    __shared__ FSrc s_src;
    __shared__ FProbe s_prb;
    __shared__ STfsf s_tf[TF ? 2 : 1];  // TFSF E and H patch tables (read from LDS, not kernel arguments)
    const int T = g.T;
    float *tab = lds;  // [3][256] CE table (kCeTable) or [ntri][3] triples (kCeTriple)
    float *ebuf = lds + ce_tab_floats(CEM, g.ntri);  // [2][T][2][256]: ex, ez of rows w = 0..T-1
    float *hbuf = ebuf + 2 * T * 512;                 // [2][T][2][256]: hx, hz of H(i+1) of rows w = 1..T
    float *cbuf = hbuf + 2 * T * 512;                 // [cx+1][T+1][2]: ex, ey at the segment's last cell
    float *tline = cbuf + (size_t)(g.cx + 1) * (T + 1) * 2;  // TFSF: [2][line_n] the incident line's H, E

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
        if (TF) {
            s_tf[0] = tfe;
            s_tf[TF ? 1 : 0] = tfh;
            s_tf[0].line = tline;  // the patches read the block's LDS copy of the line
            s_tf[TF ? 1 : 0].line = tline + tfe.line_n;
        }
    }
    if (TF)
        for (int t = threadIdx.x; t < 2 * tfe.line_n; t += blockDim.x)
            tline[t] = t < tfe.line_n ? tfe.line[t] : tfh.line[t - tfe.line_n];
    if (CEM == kCeTable || CEM == kCeTriple)
        for (int t = threadIdx.x; t < (CEM == kCeTable ? 768 : 3 * g.ntri); t += blockDim.x) tab[t] = lut[t];
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
        // TFSF patches that meet this row's planes and segment (the plane loop visits only these)
        const unsigned tmE = TF ? tfsf_mask(s_tf[0], ib, i1 - 1, j, base, base + 255) : 0u;
        const unsigned tmH = TF ? tfsf_mask(s_tf[TF ? 1 : 0], ib, i1 - 1, j, base, base + 255) : 0u;
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
            Q4 jn[3] = {zero4(), zero4(), zero4()};  // J' of a Drude row (stored by the owner)
            if (live)
                compute_E<CEM, FEAT>(g, c, tab, sp, i, j, m, l4, lc, end_lane, hc, hux, huz, hN, hx_nx, hy_nx, exn, eyn,
                                     ezn, &dr, jn, &s_tf[0], tmE);
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
                compute_H<TF>(g, sp, i, j, m, l4, hc, exn, eyn, ezn, exb, ezb, eyp, ezp, exm, eym, hxo, hyo, hzo,
                              &s_tf[TF ? 1 : 0], tmH);
                const size_t rbase = row_base(g, i, j, m);
                if (act) {
                    st4(oex + rbase + l4, exn);
                    st4(oey + rbase + l4, eyn);
                    st4(oez + rbase + l4, ezn);
                    st4(ohx + rbase + l4, hxo);
                    st4(ohy + rbase + l4, hyo);
                    st4(ohz + rbase + l4, hzo);
                    if (DR && dr.row_in(i, j)) {
                        const int k0 = g.zb0 + m * 256 + (int)l4;
                        drude_store(dr, 0, i, j, k0, jn[0]);
                        drude_store(dr, 1, i, j, k0, jn[1]);
                        drude_store(dr, 2, i, j, k0, jn[2]);
                    }
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

// ------------------------------------------------------------------ two-step sweep
// Two time steps of the interior box in one pass (state n read from one buffer set, state
// n+2 written to another; state n+1 of the box never leaves the chip).  Used with CPML,
// between shell passes that advance the CPML shell one step at a time (see the host code):
// the box keeps two planes / rows of CPML-free cells to the CPML slabs in x and y, so the
// halo cells it recomputes at step n+1 need no CPML term, and its z bounds are the 128-byte
// line bounds of the one-step box (at the box's z faces the step n+1 values of the cell
// below come from the shell's step-n+1 buffer, those of step n+2 from the output).
//
// Block: T owned rows of an x chunk [i0, i1); wave w streams row j = j0-2+w:
//   E1 (step n+1 E) on rows j0-2 .. j0+T, H1 on j0-1 .. j0+T, E2 (step n+2) on j0-1 .. j0+T-1,
//   H2 on the owned rows j0 .. j0+T-1.
// March over planes p = i0-2 .. i1 within a 256-float z segment; at plane p the block forms
// E1(p) and H1(p), then E2(p-1) and H2(p-1) (level 2 lags one plane, needing H1(p) as its
// i+1 neighbour).  Row neighbours go through LDS: H0(p+1) (hx, hz) for the row below,
// published for the next plane, H1(p-1) (hx, hz) for the row below, E1(p) (ex, ez) and
// E2(p-1) (ex, ez) for the row above, within the plane (two barriers per plane).  z
// neighbours: DPP wave shifts (lane i takes lane i+1's or i-1's value); the
// segment's first cell takes k-1 from the previous segment's last cell (LDS carry per plane
// and row) or, on the box's first segment, from memory; the k+1 neighbour of the segment's
// last cell is the first cell of the next segment, whose E1 and H1 (hx, hy) the wave forms
// as an extra element from step-n data.  H1(p-1) is published at plane p (before the first
// barrier) for the row below's E2(p-1); every exchange buffer is single: each is rewritten
// only after a barrier that follows its last read.
// Per-cell arithmetic is the in-place path's (core curl, then sources; no masks apply in
// the box or its halo), so the values equal two in-place steps bit for bit.
struct F2Lds {
    float *tab, *x1, *x2, *x3, *x4, *xe, *cb;
};

__device__ __forceinline__ Q4 lds_q(const float *p, unsigned l4)
{
    const float4 t = *reinterpret_cast<const float4 *>(p + l4);
    return Q4{{t.x, t.y, t.z, t.w}};
}

// Whole-wave lane shifts (DPP wave_shl:1 / wave_shr:1 of the 64-lane wave): lane i takes
// lane i+1's (next) or lane i-1's (prev) value; the lane at the end takes 0.
__device__ __forceinline__ float lane_next(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_prev(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_at(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ void lds_put(float *p, unsigned l4, const Q4 &q)
{
    *reinterpret_cast<float4 *>(p + l4) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
}

// E update of a quad from H: e += c * curl (KEMP order), k+1 neighbours of the last element
// given (hy1 = hy[k+1], hx1 = hx[k+1]); hu = (hx, hz) of the row above, hN = H of the next plane.
__device__ __forceinline__ void e_curl(const Q4 e[3], const Q4 c[3], const Q4 h[3], const Q4 &hux, const Q4 &huz,
                                       const Q4 hN[3], float hy1, float hx1, Q4 o[3])
{
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const float y1 = l < 3 ? h[1].v[l + 1] : hy1;
        const float x1 = l < 3 ? h[0].v[l + 1] : hx1;
        o[0].v[l] = e[0].v[l] + c[0].v[l] * ((huz.v[l] - h[2].v[l]) - (y1 - h[1].v[l]));
        o[1].v[l] = e[1].v[l] + c[1].v[l] * ((x1 - h[0].v[l]) - (hN[2].v[l] - h[2].v[l]));
        o[2].v[l] = e[2].v[l] + c[2].v[l] * ((hN[1].v[l] - h[1].v[l]) - (hux.v[l] - h[0].v[l]));
    }
}

// H update of a quad from E: h -= 0.5 * curl; eb = (ex, ez) of the row below, ep = (ey, ez)
// of the previous plane, xm / ym = ex / ey at k-1 of the first element.
__device__ __forceinline__ void h_curl(const Q4 h[3], const Q4 e[3], const Q4 &exb, const Q4 &ezb, const Q4 &eyp,
                                       const Q4 &ezp, float xm, float ym, Q4 o[3])
{
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const float ymv = l > 0 ? e[1].v[l - 1] : ym;
        const float xmv = l > 0 ? e[0].v[l - 1] : xm;
        o[0].v[l] = h[0].v[l] - 0.5f * ((e[2].v[l] - ezb.v[l]) - (e[1].v[l] - ymv));
        o[1].v[l] = h[1].v[l] - 0.5f * ((e[0].v[l] - xmv) - (e[2].v[l] - ezp.v[l]));
        o[2].v[l] = h[2].v[l] - 0.5f * ((e[1].v[l] - eyp.v[l]) - (e[0].v[l] - exb.v[l]));
    }
}

// Probe records of one quad (component c, values v) at step-table p.
__device__ __forceinline__ void probe_quad(const FProbe &p, int c, int i, int j, int k0, const Q4 &v)
{
    for (int n = 0; n < p.n; ++n) {
        if (p.comp[n] == c && p.i[n] == i && p.j[n] == j) {
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (p.k[n] == k0 + l) p.out[n] = v.v[l];
        }
    }
}

// Raw CE of a quad as the sweep carries it from one plane to the next: nothing (the
// constant 0.5), four floats, or four packed uint8 material indices into the LDS table.
template <int CEM>
struct CeRaw {
    Q4 f[3];
    unsigned t[3];
};

template <int CEM>
__device__ __forceinline__ void ce_raw_load(CeRaw<CEM> &r, const float *cex, const float *cey, const float *cez,
                                            const unsigned char *mx, const unsigned char *my, const unsigned char *mz,
                                            size_t off)
{
    if constexpr (CEM == kCeArrays) {
        r.f[0] = ld4(cex + off);
        r.f[1] = ld4(cey + off);
        r.f[2] = ld4(cez + off);
    } else if constexpr (CEM == kCeTable) {
        r.t[0] = *reinterpret_cast<const unsigned *>(mx + off);
        r.t[1] = *reinterpret_cast<const unsigned *>(my + off);
        r.t[2] = *reinterpret_cast<const unsigned *>(mz + off);
    }
}

template <int CEM>
__device__ __forceinline__ Q4 ce_raw_get(const CeRaw<CEM> &r, const float *tab, int c)
{
    if constexpr (CEM == kCeArrays) return r.f[c];
    if constexpr (CEM == kCeTable) {
        const float *t = tab + 256 * c;
        const unsigned v = r.t[c];
        return Q4{{t[v & 255u], t[(v >> 8) & 255u], t[(v >> 16) & 255u], t[v >> 24]}};
    }
    return splat4(0.5f);
}

template <int CEM>
__global__ __launch_bounds__(768) void k_fused2(
    FGeom g,
    const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    const float *__restrict__ mex, const float *__restrict__ mey,
    float *__restrict__ oex, float *__restrict__ oey, float *__restrict__ oez,
    float *__restrict__ ohx, float *__restrict__ ohy, float *__restrict__ ohz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    const unsigned char *__restrict__ mx, const unsigned char *__restrict__ my,
    const unsigned char *__restrict__ mz, const float *__restrict__ lut, FSrc src1, FSrc src2, FProbe prb1,
    FProbe prb2)
{
    extern __shared__ float lds[];
    __shared__ FSrc s_src[2];
    __shared__ FProbe s_prb[2];
    const int T = g.T;
    const int R = T + 3;  // rows streamed per block
    F2Lds L;
    L.tab = lds;
    L.x1 = lds + (CEM == kCeTable ? 768 : 0);  // [R][2][256] H0(p+1): hx, hz
    L.x2 = L.x1 + (size_t)R * 512;             // [R][2][256] E1(p): ex, ez
    L.x3 = L.x2 + (size_t)R * 512;             // [R][2][256] H1(p-1): hx, hz
    L.x4 = L.x3 + (size_t)R * 512;             // [R][2][256] E2(p-1): ex, ez
    L.xe = L.x4 + (size_t)R * 512;             // [R] ez of the extra element, step n+1
    L.cb = L.xe + R;                           // [cx+3][R][4] z carries: E1 ex, ey; E2 ex, ey

    int bid = blockIdx.x;
    {
        const int nb = g.nblocks, nb8 = nb & ~7;
        if (bid < nb8) bid = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    }
    const int rb = bid % g.nrb;
    const int chunk = bid / g.nrb;
    const int j0 = g.yb + rb * T;
    const int nown = min(T, g.ye - j0);
    const int i0 = g.xb + chunk * g.cx;
    const int i1 = min(g.xe, i0 + g.cx);
    const int ib = i0 - 2;
    bool src_hit = false, prb_hit = false;
    for (int n = 0; n < kMaxSrc; ++n)
        src_hit |= n < src1.n && src1.i1[n] >= ib && src1.i0[n] <= i1 && src1.j1[n] >= j0 - 2 && src1.j0[n] <= j0 + nown;
    for (int n = 0; n < kMaxProbe; ++n)
        prb_hit |= n < prb1.n && prb1.i[n] >= i0 && prb1.i[n] < i1 && prb1.j[n] >= j0 && prb1.j[n] <= j0 + nown - 1;
    if (threadIdx.x == 0) {
        if (src_hit) {
            s_src[0] = src1;
            s_src[1] = src2;
        }
        if (prb_hit) {
            s_prb[0] = prb1;
            s_prb[1] = prb2;
        }
    }
    if (CEM == kCeTable)
        for (int t = threadIdx.x; t < 768; t += blockDim.x) L.tab[t] = lut[t];
    __syncthreads();
    const FSrc *sp1 = src_hit ? &s_src[0] : nullptr;
    const FSrc *sp2 = src_hit ? &s_src[1] : nullptr;

    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const unsigned l4 = 4u * (unsigned)lane;
    const int j = j0 - 2 + w;
    const bool live = w <= nown + 2;           // E1 rows j0-2 .. j0+nown
    const bool h1row = live && w >= 1;         // H1 rows j0-1 .. j0+nown
    const bool e2row = w >= 1 && w <= nown + 1;  // E2 rows j0-1 .. j0+nown-1
    const bool own = w >= 2 && w <= nown + 1;  // H2 and stored rows j0 .. j0+nown-1
    const bool top = w == nown + 2;            // loads H0 of the row above itself
    const Q4 z4 = zero4();
    const float *tab = L.tab;
    const float *ctab[3] = {L.tab, L.tab + 256, L.tab + 512};
    float *cbw = L.cb + (size_t)w * 4;         // this row's carries; plane q at + q*R*4

    for (int m = 0; m < g.nseg; ++m) {
        const int base = g.zb0 + m * 256;
        const int lo = max(g.kb - base, 0), hi = min(g.ke - base, 256);
        const bool act = (int)l4 >= lo && (int)l4 < hi;
        const unsigned lc = (int)l4 < lo ? (unsigned)lo : (int)l4 >= hi ? (unsigned)(hi - 4) : l4;
        const bool start_lane = (int)l4 == lo;
        const bool end_lane = (int)l4 + 4 == hi;
        const int endl = (hi >> 2) - 1;  // index of the end lane
        const int kend = base + hi;      // the extra element: first cell after the segment's box part
        const int k0 = base + (int)l4;
        // state at plane ib (live rows)
        Q4 hc[3], hux = z4, huz = z4;
        float hcx[3] = {0.f, 0.f, 0.f};  // H0 of the extra element (hx, hy, hz)
        hc[0] = hc[1] = hc[2] = z4;
        if (live) {
            const size_t r = row_base(g, ib, j, m);
            hc[0] = ld4(hx + r + lc);
            hc[1] = ld4(hy + r + lc);
            hc[2] = ld4(hz + r + lc);
            const size_t rx_ = ((size_t)ib * g.ny + j) * g.nzp + kend;
            hcx[0] = hx[rx_];
            hcx[1] = hy[rx_];
            hcx[2] = hz[rx_];
            const size_t ru = row_base(g, ib, j + 1, m);
            hux = ld4(hx + ru + lc);
            huz = ld4(hz + ru + lc);
        }
        Q4 e1q[3] = {z4, z4, z4};      // E1(p-1)
        CeRaw<CEM> crq;                // CE of plane p-1
        Q4 h1q[3] = {z4, z4, z4};      // H1(p-1)
        float h1xq[2] = {0.f, 0.f};    // H1(p-1) of the extra element: hx, hy
        Q4 e2p[2] = {z4, z4};          // E2(p-2): ey, ez
        float ez1xq = 0.f;             // ez of the extra element at p-1
        for (int p = ib; p <= i1; ++p) {
            const int q = p - ib;
            const bool do_h1 = h1row && p > ib;
            const bool do_e2 = e2row && p >= ib + 2;
            const bool do_h2 = own && p >= ib + 3;
            // ---------------------------------------------- phase 1: loads, E1(p)
            Q4 hN[3];
            hN[0] = hN[1] = hN[2] = z4;
            float hNx[3] = {0.f, 0.f, 0.f};
            Q4 e1[3];
            CeRaw<CEM> cr;
            float e1x[3] = {0.f, 0.f, 0.f};
            e1[0] = e1[1] = e1[2] = z4;
            // k-1 values of the segment's first cell: previous segment's carries (read before
            // this plane overwrites them) or, on the first segment, the memory values
            float c1x = 0.f, c1y = 0.f, c2x = 0.f, c2y = 0.f;
            if (m == 0) {
                if (do_h1) {
                    const size_t r = ((size_t)p * g.ny + j) * g.nzp + (g.kb - 1);
                    c1x = mex[r];
                    c1y = mey[r];
                }
                if (do_h2) {
                    const size_t r = ((size_t)(p - 1) * g.ny + j) * g.nzp + (g.kb - 1);
                    c2x = oex[r];
                    c2y = oey[r];
                }
            } else if (start_lane) {
                const float *cs = cbw + (size_t)q * R * 4;
                c1x = cs[0];
                c1y = cs[1];
                if (q >= 1) {
                    const float *cs2 = cbw + (size_t)(q - 1) * R * 4;
                    c2x = cs2[2];
                    c2y = cs2[3];
                }
            }
            if (live) {
                const size_t rn = row_base(g, p + 1, j, m);
                hN[0] = ld4(hx + rn + lc);
                hN[1] = ld4(hy + rn + lc);
                hN[2] = ld4(hz + rn + lc);
                const size_t cxn = ((size_t)(p + 1) * g.ny + j) * g.nzp + kend;
                hNx[0] = hx[cxn];
                hNx[1] = hy[cxn];
                hNx[2] = hz[cxn];
                const size_t r = row_base(g, p, j, m);
                Q4 e0[3];
                e0[0] = ld4(ex + r + lc);
                e0[1] = ld4(ey + r + lc);
                e0[2] = ld4(ez + r + lc);
                ce_raw_load<CEM>(cr, cex, cey, cez, mx, my, mz, r + lc);
                // the extra element's step-n data
                const size_t cx0 = ((size_t)p * g.ny + j) * g.nzp + kend;
                float e0x[3] = {ex[cx0], ey[cx0], ez[cx0]};
                float cxx[3];
                if (CEM == kCeConst) {
                    cxx[0] = cxx[1] = cxx[2] = 0.5f;
                } else if (CEM == kCeArrays) {
                    cxx[0] = cex[cx0];
                    cxx[1] = cey[cx0];
                    cxx[2] = cez[cx0];
                } else {
                    cxx[0] = ctab[0][mx[cx0]];
                    cxx[1] = ctab[1][my[cx0]];
                    cxx[2] = ctab[2][mz[cx0]];
                }
                const float hy_k1 = hy[cx0 + 1], hx_k1 = hx[cx0 + 1];
                const float hz_u = hz[cx0 + g.nzp], hx_u = hx[cx0 + g.nzp];
                if (p > ib) {  // H0(p, j+1): from the row above (LDS) or memory (top row)
                    if (!top) {
                        const float *xb = L.x1 + (size_t)((w + 1) * 2) * 256;
                        hux = lds_q(xb, l4);
                        huz = lds_q(xb + 256, l4);
                    } else {
                        const size_t ru = row_base(g, p, j + 1, m);
                        hux = ld4(hx + ru + lc);
                        huz = ld4(hz + ru + lc);
                    }
                }
                float hy1 = lane_next(hc[1].v[0]), hx1 = lane_next(hc[0].v[0]);
                if (end_lane) {
                    hy1 = hcx[1];
                    hx1 = hcx[0];
                }
                const Q4 ce[3] = {ce_raw_get<CEM>(cr, tab, 0), ce_raw_get<CEM>(cr, tab, 1), ce_raw_get<CEM>(cr, tab, 2)};
                e_curl(e0, ce, hc, hux, huz, hN, hy1, hx1, e1);
                e1x[0] = e0x[0] + cxx[0] * ((hz_u - hcx[2]) - (hy_k1 - hcx[1]));
                e1x[1] = e0x[1] + cxx[1] * ((hx_k1 - hcx[0]) - (hNx[2] - hcx[2]));
                e1x[2] = e0x[2] + cxx[2] * ((hNx[1] - hcx[1]) - (hx_u - hcx[0]));
                if (sp1) {
                    apply_src(*sp1, 0, p, j, k0, e1[0]);
                    apply_src(*sp1, 1, p, j, k0, e1[1]);
                    apply_src(*sp1, 2, p, j, k0, e1[2]);
                    Q4 t[3] = {splat4(e1x[0]), splat4(e1x[1]), splat4(e1x[2])};
                    for (int c = 0; c < 3; ++c) {
                        apply_src(*sp1, c, p, j, kend, t[c]);
                        e1x[c] = t[c].v[0];
                    }
                }
                float *xs = L.x2 + (size_t)(w * 2) * 256;
                lds_put(xs, l4, e1[0]);
                lds_put(xs + 256, l4, e1[2]);
                if (lane == 0) L.xe[w] = e1x[2];
                if (end_lane) {
                    float *cs = cbw + (size_t)q * R * 4;
                    cs[0] = e1[0].v[3];
                    cs[1] = e1[1].v[3];
                }
            }
            if (h1row) {  // H1(p-1) (hx, hz) for the row below's E2(p-1), this plane
                float *xs = L.x3 + (size_t)(w * 2) * 256;
                lds_put(xs, l4, h1q[0]);
                lds_put(xs + 256, l4, h1q[2]);
            }
            __syncthreads();
            // ---------------------------------------------- phase 2: H1(p)
            Q4 h1[3];
            h1[0] = h1[1] = h1[2] = z4;
            float h1x[2] = {0.f, 0.f};
            if (do_h1) {
                const float *xb = L.x2 + (size_t)((w - 1) * 2) * 256;
                const Q4 exb = lds_q(xb, l4), ezb = lds_q(xb + 256, l4);
                float xm = lane_prev(e1[0].v[3]), ym = lane_prev(e1[1].v[3]);
                if (start_lane) {
                    xm = c1x;
                    ym = c1y;
                }
                const Q4 e1p[2] = {e1q[1], e1q[2]};
                h_curl(hc, e1, exb, ezb, e1p[0], e1p[1], xm, ym, h1);
                const float ey_end = lane_at(e1[1].v[3], endl), ex_end = lane_at(e1[0].v[3], endl);
                h1x[0] = hcx[0] - 0.5f * ((e1x[2] - L.xe[w - 1]) - (e1x[1] - ey_end));
                h1x[1] = hcx[1] - 0.5f * ((e1x[0] - ex_end) - (e1x[2] - ez1xq));
                if (sp1) {
                    apply_src(*sp1, 3, p, j, k0, h1[0]);
                    apply_src(*sp1, 4, p, j, k0, h1[1]);
                    apply_src(*sp1, 5, p, j, k0, h1[2]);
                    Q4 t0 = splat4(h1x[0]), t1 = splat4(h1x[1]);
                    apply_src(*sp1, 3, p, j, kend, t0);
                    apply_src(*sp1, 4, p, j, kend, t1);
                    h1x[0] = t0.v[0];
                    h1x[1] = t1.v[0];
                }
                if (prb_hit && own && act && p >= i0 && p < i1) {
                    for (int c = 0; c < 3; ++c) probe_quad(s_prb[0], c, p, j, k0, e1[c]);
                    for (int c = 0; c < 3; ++c) probe_quad(s_prb[0], 3 + c, p, j, k0, h1[c]);
                }
            }
            if (live) {  // H0(p+1) for the row below (read at plane p+1, after this plane's second barrier)
                float *xs = L.x1 + (size_t)(w * 2) * 256;
                lds_put(xs, l4, hN[0]);
                lds_put(xs + 256, l4, hN[2]);
            }
            // ---------------------------------------------- phase 3: E2(p-1)
            Q4 e2[3];
            e2[0] = e2[1] = e2[2] = z4;
            if (do_e2) {
                const float *xu = L.x3 + (size_t)((w + 1) * 2) * 256;
                const Q4 h1ux = lds_q(xu, l4), h1uz = lds_q(xu + 256, l4);
                float hy1 = lane_next(h1q[1].v[0]), hx1 = lane_next(h1q[0].v[0]);
                if (end_lane) {
                    hy1 = h1xq[1];
                    hx1 = h1xq[0];
                }
                const Q4 cq[3] = {ce_raw_get<CEM>(crq, tab, 0), ce_raw_get<CEM>(crq, tab, 1),
                                  ce_raw_get<CEM>(crq, tab, 2)};
                e_curl(e1q, cq, h1q, h1ux, h1uz, h1, hy1, hx1, e2);
                if (sp2) {
                    apply_src(*sp2, 0, p - 1, j, k0, e2[0]);
                    apply_src(*sp2, 1, p - 1, j, k0, e2[1]);
                    apply_src(*sp2, 2, p - 1, j, k0, e2[2]);
                }
                float *xs = L.x4 + (size_t)(w * 2) * 256;
                lds_put(xs, l4, e2[0]);
                lds_put(xs + 256, l4, e2[2]);
                if (end_lane) {
                    float *cs = cbw + (size_t)(q - 1) * R * 4;
                    cs[2] = e2[0].v[3];
                    cs[3] = e2[1].v[3];
                }
            }
            __syncthreads();
            // ---------------------------------------------- phase 4: H2(p-1), stores
            if (do_h2) {
                const float *xb = L.x4 + (size_t)((w - 1) * 2) * 256;
                const Q4 exb = lds_q(xb, l4), ezb = lds_q(xb + 256, l4);
                float xm = lane_prev(e2[0].v[3]), ym = lane_prev(e2[1].v[3]);
                if (start_lane) {
                    xm = c2x;
                    ym = c2y;
                }
                Q4 h2[3];
                h_curl(h1q, e2, exb, ezb, e2p[0], e2p[1], xm, ym, h2);
                if (sp2) {
                    apply_src(*sp2, 3, p - 1, j, k0, h2[0]);
                    apply_src(*sp2, 4, p - 1, j, k0, h2[1]);
                    apply_src(*sp2, 5, p - 1, j, k0, h2[2]);
                }
                if (act) {
                    const size_t r = row_base(g, p - 1, j, m) + l4;
                    st4(oex + r, e2[0]);
                    st4(oey + r, e2[1]);
                    st4(oez + r, e2[2]);
                    st4(ohx + r, h2[0]);
                    st4(ohy + r, h2[1]);
                    st4(ohz + r, h2[2]);
                    if (prb_hit) {
                        for (int c = 0; c < 3; ++c) probe_quad(s_prb[1], c, p - 1, j, k0, e2[c]);
                        for (int c = 0; c < 3; ++c) probe_quad(s_prb[1], 3 + c, p - 1, j, k0, h2[c]);
                    }
                }
            }
            // shift the plane registers
            e1q[0] = e1[0];
            e1q[1] = e1[1];
            e1q[2] = e1[2];
            crq = cr;
            h1q[0] = h1[0];
            h1q[1] = h1[1];
            h1q[2] = h1[2];
            h1xq[0] = h1x[0];
            h1xq[1] = h1x[1];
            e2p[0] = e2[1];
            e2p[1] = e2[2];
            ez1xq = e1x[2];
            hc[0] = hN[0];
            hc[1] = hN[1];
            hc[2] = hN[2];
            hcx[0] = hNx[0];
            hcx[1] = hNx[1];
            hcx[2] = hNx[2];
        }
    }
}

// CE: CEM as in the sweep (kCeTable reads the 3 x 256 table from global memory).
template <int CEM, int FEAT>
__global__ __launch_bounds__(256) void k_shell_e(SGeom g, ShellBoxes sb, unsigned total, const float *ex, const float *ey,
                                                 const float *ez, float *oex, float *oey, float *oez,
                                                 const float *__restrict__ hx, const float *__restrict__ hy,
                                                 const float *__restrict__ hz, const float *__restrict__ cex,
                                                 const float *__restrict__ cey, const float *__restrict__ cez,
                                                 const unsigned char *__restrict__ mx,
                                                 const unsigned char *__restrict__ my,
                                                 const unsigned char *__restrict__ mz, const float *__restrict__ lut,
                                                 SCpml cp, SSrc src, SProbe prb, SDrude dr, STfsf tfe)
{
    const float *tab = lut;  // read from global memory (L1/L2 resident)
    const unsigned f = blockIdx.x * 256u + threadIdx.x;
    if (f >= total) return;
    int i, j, k0;
    shell_quad(sb, f, i, j, k0);
    e_quad_at<CEM, true, FEAT>(g, i, j, k0, ex, ey, ez, oex, oey, oez, hx, hy, hz, cex, cey, cez, cp, src, prb, mx, my,
                               mz, tab, &dr, &tfe);
}

template <bool TF>
__global__ __launch_bounds__(256) void k_shell_h(SGeom g, ShellBoxes sb, unsigned total, const float *__restrict__ oex,
                                                 const float *__restrict__ oey, const float *__restrict__ oez,
                                                 const float *hx, const float *hy, const float *hz, float *ohx,
                                                 float *ohy, float *ohz, SCpml cp, SSrc src, SProbe prb, STfsf tfh)
{
    const unsigned f = blockIdx.x * 256u + threadIdx.x;
    if (f >= total) return;
    int i, j, k0;
    shell_quad(sb, f, i, j, k0);
    h_quad_at<true>(g, i, j, k0, oex, oey, oez, hx, hy, hz, ohx, ohy, ohz, cp, src, prb, TF ? &tfh : nullptr);
}

}  // namespace mifdtd

// =========================================================================== C ABI
using namespace mifdtd;

// Launches M(CEV, FEAT) for the run-time feature set f (kFeatDrude | kFeatTfsf bits).
#define MIFDTD_BY_FEAT(M, CEV, f)                  \
    do {                                           \
        switch (f) {                               \
        case 0: M(CEV, 0); break;                  \
        case kFeatDrude: M(CEV, kFeatDrude); break; \
        case kFeatTfsf: M(CEV, kFeatTfsf); break;  \
        default: M(CEV, kFeatDrude | kFeatTfsf);   \
        }                                          \
    } while (0)

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
    if (d->cpml_axes < 0 || d->cpml_axes > 7) return fail(MIFDTD_EINVAL, "cpml_axes %d outside [0,7]", d->cpml_axes);
    const int pml_axes = pml ? (d->cpml_axes ? d->cpml_axes : 7) : 0;
    // periodic axes (in the order the copies run)
    if (d->npbc < 0 || d->npbc > 3) return fail(MIFDTD_EINVAL, "npbc %d outside [0,3]", d->npbc);
    int periodic = 0;
    for (int q = 0; q < d->npbc; q++) {
        const int a = d->pbc_axes[q];
        if (a < 0 || a > 2 || ((periodic >> a) & 1)) return fail(MIFDTD_EINVAL, "periodic axis %d", a);
        if ((pml_axes >> a) & 1) return fail(MIFDTD_EINVAL, "axis %d is both periodic and CPML", a);
        periodic |= 1 << a;
    }
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
        scp.axes = pml_axes;
        scp.bE = d->bE; scp.aE = d->aE; scp.bH = d->bH; scp.aH = d->aH;
        scp.ph = d->psiH;
        scp.sX = (size_t)np_ * ny * nzp;
        scp.sY = (size_t)nx * np_ * nzp;
        scp.sZ = (size_t)nx * ny * np_;
    }

    // interior box: no CPML term in its E and H updates nor in its E halo
    const int zph = line_phase(d->e[0], nzp);
    FGeom g;
    g.nx = nx; g.ny = ny; g.nzp = nzp; g.nz = nz;
    g.T = T;
    g.cx = CX;
    {
        // per axis: inside the CPML slabs; on a periodic axis all but the last plane (whose
        // E the periodic copy replaces before H reads it); else the whole extent
        const int np_ = d->npml;
        auto bounds = [&](int a, int n, int &lo, int &hi) {
            if ((pml_axes >> a) & 1) { lo = np_ + 1; hi = n - 1 - np_; }
            else if ((periodic >> a) & 1) { lo = 0; hi = n - 1; }
            else { lo = 0; hi = n; }
        };
        bounds(0, nx, g.xb, g.xe);
        bounds(1, ny, g.yb, g.ye);
        if ((pml_axes >> 2) & 1) { g.kb = line_up(np_ + 1, zph); g.ke = line_down(nz - 1 - np_, zph); }
        else if ((periodic >> 2) & 1) { g.kb = 0; g.ke = 32 * ((nz - 1) / 32); }
        else { g.kb = 0; g.ke = nzp; }
        if (g.xb >= g.xe || g.yb >= g.ye || g.kb >= g.ke) {  // no interior: the shell is the grid
            g.xb = g.xe = nx;
            g.yb = g.ye = ny;
            g.kb = g.ke = nzp;
        }
    }
    // shell passes run whenever the box is not the whole grid
    const bool shell_on = !(g.xb == 0 && g.xe == nx && g.yb == 0 && g.ye == ny && g.kb == 0 && g.ke == nzp);
    const bool sweep = g.xe > g.xb;
    g.nrb = sweep ? (g.ye - g.yb + T - 1) / T : 0;
    g.nchunks = sweep ? (g.xe - g.xb + CX - 1) / CX : 0;
    g.zb0 = ((pml_axes >> 2) & 1) && zph ? seg_origin(g.kb, d->e[0]) : 256 * (g.kb / 256);
    g.nseg = sweep ? (g.ke - g.zb0 + 255) / 256 : 0;
    const long long nbl = (long long)g.nrb * g.nchunks;
    if (nbl >= (1ll << 31)) return fail(MIFDTD_ERANGE, "too many blocks");
    g.nblocks = (int)nbl;
    // CE of the plain sweep as one uint16 triple index per cell (the shell keeps cem)
    if (d->ce_tri) {
        if (!d->ce_tri_lut) return fail(MIFDTD_ENULL, "CE triple table is NULL");
        if (d->ce_tri_n < 1 || d->ce_tri_n > MIFDTD_FUSED_MAX_CE_TRI)
            return fail(MIFDTD_EINVAL, "ce_tri_n %d outside [1,%d]", d->ce_tri_n, MIFDTD_FUSED_MAX_CE_TRI);
        if (reinterpret_cast<uintptr_t>(d->ce_tri) & 7u) return fail(MIFDTD_EALIGN, "CE triple indices not 8-byte aligned");
        if (cem != kCeTable) return fail(MIFDTD_EINVAL, "CE triple indices need the material indices (ce_mat) too");
    }
    const int sweep_cem = d->ce_tri ? (int)kCeTriple : cem;
    g.ntri = sweep_cem == kCeTriple ? d->ce_tri_n : 0;
    const size_t lds_bytes = sizeof(float) * ((size_t)ce_tab_floats(sweep_cem, g.ntri) + (size_t)2 * 2 * T * 2 * 256 +
                                              (size_t)(CX + 1) * (T + 1) * 2 +
                                              (d->tfsf_line_n > 0 ? 2 * (size_t)d->tfsf_line_n : 0));
    if (lds_bytes + sizeof(FSrc) + sizeof(FProbe) + 2 * sizeof(STfsf) > kMaxLds)
        return fail(MIFDTD_EINVAL, "rows_per_thread %d / planes_per_chunk %d need %zu B of LDS (> %zu)", T, CX, lds_bytes,
                    kMaxLds);

    // shell boxes (only with CPML)
    ShellBoxes sb{};
    unsigned long long total = 0;
    const int Q = nzp / 4;
    if (shell_on) {
        add_shell(sb, total, g, Q);
        if (total >= (1ull << 32) - 256) return fail(MIFDTD_ERANGE, "CPML shell too large");
        sb.start[sb.n] = (unsigned)total;
    }

    // step pairs (third buffer set given, CPML on): the two-step sweep's box keeps two
    // planes / rows to the CPML slabs in x and y (the step n+1 values it forms on its
    // halo need no CPML term) and the one-step box's z bounds less the cells next to the
    // z slabs; per pair, shell launches on: [0] the shell, [1] the shell and the box's low
    // fringe, [2] the shell and both fringes.
    FGeom g2{};
    ShellBoxes sb2[3]{};
    unsigned long long tot2[3] = {0, 0, 0};
    size_t lds2 = 0;
    bool pairs = false;
    const int R2 = T + 3;
    if (pml && d->e_mid[0]) {
        if (pml_axes != 7 || periodic) return fail(MIFDTD_EINVAL, "step pairs need CPML on every face and no periodic axis");
        for (int q = 0; q < 6; q++) {
            if (!d->e_mid[q]) return fail(MIFDTD_ENULL, "e_mid buffer is NULL");
            if (!aligned16(d->e_mid[q])) return fail(MIFDTD_EALIGN, "e_mid buffer not 16-byte aligned");
        }
        if (T > 9) return fail(MIFDTD_EINVAL, "rows_per_thread must be in [1,9] with step pairs (rows+3 waves)");
        const int np_ = d->npml;
        g2 = g;
        g2.xb = np_ + 2; g2.xe = nx - 2 - np_;
        g2.yb = np_ + 2; g2.ye = ny - 2 - np_;
        g2.kb = line_up(np_ + 2, zph); g2.ke = line_down(nz - 2 - np_, zph);
        pairs = g2.xe - g2.xb >= 4 && g2.ye - g2.yb >= 4 && g2.ke > g2.kb;  // else single steps only
        if (pairs) {
            g2.nrb = (g2.ye - g2.yb + T - 1) / T;
            g2.nchunks = (g2.xe - g2.xb + CX - 1) / CX;
            g2.zb0 = zph ? seg_origin(g2.kb, d->e[0]) : 256 * (g2.kb / 256);
            g2.nseg = (g2.ke - g2.zb0 + 255) / 256;
            const long long nb2 = (long long)g2.nrb * g2.nchunks;
            if (nb2 >= (1ll << 31)) return fail(MIFDTD_ERANGE, "too many blocks");
            g2.nblocks = (int)nb2;
            lds2 = sizeof(float) * ((cem == kCeTable ? 768 : 0) + (size_t)4 * R2 * 512 + (size_t)4 * ((R2 + 3) / 4) +
                                    (size_t)(CX + 3) * R2 * 4);
            if (lds2 + 2 * sizeof(FSrc) + 2 * sizeof(FProbe) > kMaxLds)
                return fail(MIFDTD_EINVAL, "rows_per_thread %d / planes_per_chunk %d need %zu B of LDS with step pairs (> %zu)",
                            T, CX, lds2, kMaxLds);
            for (int v = 0; v < 3; v++) {
                add_shell(sb2[v], tot2[v], g2, Q);
                if (v >= 1) add_low_fringe(sb2[v], tot2[v], g2);
                if (v == 2) add_high_fringe(sb2[v], tot2[v], g2);
                if (tot2[v] >= (1ull << 32) - 256) return fail(MIFDTD_ERANGE, "CPML shell too large");
                sb2[v].start[sb2[v].n] = (unsigned)tot2[v];
            }
        }
    }
    // Drude box (single steps only)
    const bool drude = d->drude_j[0] != nullptr;
    SDrude sdr{};
    if (drude) {
        if (pairs) return fail(MIFDTD_EINVAL, "step pairs (e_mid) do not support a Drude box");
        if (T > kDrudeMaxRows)
            return fail(MIFDTD_EINVAL, "rows_per_thread must be in [1,%d] with a Drude box", kDrudeMaxRows);
        const int *b = d->drude_box;
        if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] >= nx || b[3] >= ny || b[5] >= nz || b[0] > b[1] || b[2] > b[3] ||
            b[4] > b[5])
            return fail(MIFDTD_EINVAL, "Drude box outside the grid");
        for (int q = 0; q < 3; q++) {
            if (!d->drude_a[q] || !d->drude_b[q] || !d->drude_j[q] || !d->drude_j_alt[q])
                return fail(MIFDTD_ENULL, "Drude arrays must all be set");
            if (!aligned16(d->drude_a[q]) || !aligned16(d->drude_b[q]) || !aligned16(d->drude_j[q]) ||
                !aligned16(d->drude_j_alt[q]))
                return fail(MIFDTD_EALIGN, "Drude arrays not 16-byte aligned");
        }
        sdr.i0 = b[0]; sdr.i1 = b[1]; sdr.j0 = b[2]; sdr.j1 = b[3]; sdr.k0 = b[4]; sdr.k1 = b[5];
        sdr.lj = b[3] - b[2] + 1;
        sdr.lk = b[5] - b[4] + 1;
        for (int q = 0; q < 3; q++) {
            sdr.a[q] = d->drude_a[q];
            sdr.b[q] = d->drude_b[q];
            sdr.jin[q] = d->drude_j[q];
            sdr.jout[q] = d->drude_j_alt[q];
        }
    } else {
        sdr.i0 = sdr.j0 = sdr.k0 = 1;  // an empty box
        sdr.i1 = sdr.j1 = sdr.k1 = 0;
    }

    // TFSF plane wave (single steps only)
    const bool tfsf = d->tfsf_line_n > 0;
    STfsf tfe{}, tfh{};
    static_assert(kTfMaxPatch == MIFDTD_TFSF_MAX_PATCH, "TFSF patch tables differ");
    if (tfsf) {
        if (pairs) return fail(MIFDTD_EINVAL, "step pairs (e_mid) do not support a TFSF source");
        if (!d->tfsf_line_e || !d->tfsf_line_h || !d->tfsf_abc) return fail(MIFDTD_ENULL, "TFSF line pointer is NULL");
        if (d->tfsf_line_n < 6) return fail(MIFDTD_EINVAL, "TFSF line of %d points (at least 6)", d->tfsf_line_n);
        if (nsteps > 0 && !d->tfsf_src) return fail(MIFDTD_ENULL, "tfsf_src is NULL");
        const mifdtd_tfsf_patches *P[2] = {&d->tfsf_e, &d->tfsf_h};
        STfsf *Q[2] = {&tfe, &tfh};
        for (int h = 0; h < 2; h++) {
            const mifdtd_tfsf_patches &pp = *P[h];
            STfsf &q = *Q[h];
            if (pp.n < 0 || pp.n > kTfMaxPatch) return fail(MIFDTD_EINVAL, "TFSF patch count %d", pp.n);
            if (pp.axis < 0 || pp.axis > 2 || pp.axis != d->tfsf_e.axis) return fail(MIFDTD_EINVAL, "TFSF axis %d", pp.axis);
            q.n = pp.n;
            q.axis = pp.axis;
            q.line_n = d->tfsf_line_n;
            q.line = h == 0 ? d->tfsf_line_h : d->tfsf_line_e;  // E patches read the line's H and vice versa
            for (int r = 0; r < pp.n; r++) {
                const int *b = pp.box[r];
                if (pp.comp[r] < 3 * h || pp.comp[r] > 3 * h + 2)
                    return fail(MIFDTD_EINVAL, "TFSF %s patch %d of component %d", h ? "H" : "E", r, pp.comp[r]);
                if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] >= nx || b[3] >= ny || b[5] >= nz || b[0] > b[1] ||
                    b[2] > b[3] || b[4] > b[5])
                    return fail(MIFDTD_EINVAL, "TFSF patch %d outside the grid", r);
                const int ext = b[2 * pp.axis + 1] - b[2 * pp.axis];
                if (pp.base[r] < 0 || pp.base[r] + ext >= d->tfsf_line_n)
                    return fail(MIFDTD_EINVAL, "TFSF patch %d reads outside the line", r);
                q.comp[r] = pp.comp[r];
                for (int m = 0; m < 6; m++) q.box[r][m] = b[m];
                q.base[r] = pp.base[r];
                q.coef[r] = pp.coef[r];
            }
        }
    }
    const int feat = (drude ? kFeatDrude : 0) | (tfsf ? kFeatTfsf : 0);

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
    float *out[6], *mid[6];
    for (int q = 0; q < 6; q++) { in[q] = d->e[q]; out[q] = d->e_alt[q]; mid[q] = d->e_mid[q]; }
    float *pin = d->psiE, *pout = d->psiE_alt;

    hipStream_t st = to_stream(stream);
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
    int nsweeps = 0;  // sweep launches so far (event pairs used)

    // the sources' time values and probe rows of step s
    auto step_tables = [&](int s, FSrc &fs, SSrc &ss, FProbe &fp, SProbe &sp) {
        fs = src;
        ss = ssrc;
        fp = prb;
        sp = sprb;
        for (int n = 0; n < src.n; n++) ss.tval[n] = fs.tval[n] = tvals[(size_t)s * src.n + n];
        sp.out = fp.out = d->nprobe ? probe_series + (size_t)s * d->nprobe : nullptr;
    };
    // shell E pass over boxes b: E from ei (H from hi) to eo, psi E from pe to pe_out
    auto shell_e = [&](const ShellBoxes &b, unsigned long long tot, const float *const *ei, const float *const *hi,
                       float *const *eo, float *pe, float *pe_out, const SSrc &ss, const SProbe &sp, int fe) -> int {
        SCpml c = scp;
        c.pe = pe;
        c.pe_out = pe_out;
        const unsigned nbk = (unsigned)((tot + 255) / 256);
        if (!nbk) return 0;
#define MIFDTD_SHELL_E_LAUNCH(CEV, FV)                                                                            \
    k_shell_e<CEV, FV><<<nbk, 256, 0, st>>>(sg, b, (unsigned)tot, ei[0], ei[1], ei[2], eo[0], eo[1], eo[2], hi[3],    \
                                            hi[4], hi[5], d->ce[0], d->ce[1], d->ce[2], d->ce_mat[0], d->ce_mat[1],  \
                                            d->ce_mat[2], d->ce_lut, c, ss, sp, sdr, tfe)
        if (cem == kCeTable) MIFDTD_BY_FEAT(MIFDTD_SHELL_E_LAUNCH, kCeTable, fe);
        else if (cem == kCeArrays) MIFDTD_BY_FEAT(MIFDTD_SHELL_E_LAUNCH, kCeArrays, fe);
        else MIFDTD_BY_FEAT(MIFDTD_SHELL_E_LAUNCH, kCeConst, fe);
#undef MIFDTD_SHELL_E_LAUNCH
        MIFDTD_CHECK_LAUNCH("CPML shell E launch");
        return 0;
    };
    // shell H pass over boxes b: H from hi to ho with the new E of en (psi H in place)
    auto shell_h = [&](const ShellBoxes &b, unsigned long long tot, const float *const *en, const float *const *hi,
                       float *const *ho, const SSrc &ss, const SProbe &sp) -> int {
        const unsigned nbk = (unsigned)((tot + 255) / 256);
        if (!nbk) return 0;
        if (tfsf)
            k_shell_h<true><<<nbk, 256, 0, st>>>(sg, b, (unsigned)tot, en[0], en[1], en[2], hi[3], hi[4], hi[5], ho[3],
                                                 ho[4], ho[5], scp, ss, sp, tfh);
        else
            k_shell_h<false><<<nbk, 256, 0, st>>>(sg, b, (unsigned)tot, en[0], en[1], en[2], hi[3], hi[4], hi[5], ho[3],
                                                  ho[4], ho[5], scp, ss, sp, tfh);
        MIFDTD_CHECK_LAUNCH("CPML shell H launch");
        return 0;
    };
    auto mark = [&](int q) {
        if (sweep_ms) (void)hipEventRecord(marks[2 * (size_t)nsweeps + q], st);
    };
    auto swap_sets = [&]() {
        for (int q = 0; q < 6; q++) { const float *x = in[q]; in[q] = out[q]; out[q] = const_cast<float *>(x); }
        float *tp = pin;
        pin = pout;
        pout = tp;
        for (int q = 0; q < 3; q++) {  // Drude current sets
            const float *x = sdr.jin[q];
            sdr.jin[q] = sdr.jout[q];
            sdr.jout[q] = const_cast<float *>(x);
        }
    };
    const unsigned nthr = 64u * (unsigned)(T + 1);
    const bool wide = T > kNarrowRows;  // blocks of more than 10 waves: the 1024-thread build
    const unsigned nthr2 = 64u * (unsigned)R2;
    int swaps = 0;
    int s = 0;
    // step pairs: shell E of step s (shell and both fringes) and shell H of step s (shell
    // and low fringe) into the middle set, shell E of step s+1, the two-step sweep, then
    // shell H of step s+1; psi E goes pin -> pout (step s) and is updated in place in pout
    // (step s+1).
    for (; pairs && s + 2 <= nsteps; s += 2) {
        FSrc fs1, fs2;
        SSrc ss1, ss2;
        FProbe fp1, fp2;
        SProbe sp1, sp2;
        step_tables(s, fs1, ss1, fp1, sp1);
        step_tables(s + 1, fs2, ss2, fp2, sp2);
        if (int rc = shell_e(sb2[2], tot2[2], in, in, mid, pin, pout, ss1, sp1, 0)) return rc;
        if (int rc = shell_h(sb2[1], tot2[1], mid, in, mid, ss1, sp1)) return rc;
        if (int rc = shell_e(sb2[0], tot2[0], mid, mid, out, pout, pout, ss2, sp2, 0)) return rc;
        mark(0);
#define MIFDTD_FUSED2_LAUNCH(CEV)                                                                                   \
    k_fused2<CEV><<<g2.nblocks, nthr2, lds2, st>>>(g2, in[0], in[1], in[2], in[3], in[4], in[5], mid[0], mid[1],    \
                                                   out[0], out[1], out[2], out[3], out[4], out[5], d->ce[0], d->ce[1], \
                                                   d->ce[2], d->ce_mat[0], d->ce_mat[1], d->ce_mat[2], d->ce_lut, fs1, \
                                                   fs2, fp1, fp2)
        if (cem == kCeTable) MIFDTD_FUSED2_LAUNCH(kCeTable);
        else if (cem == kCeArrays) MIFDTD_FUSED2_LAUNCH(kCeArrays);
        else MIFDTD_FUSED2_LAUNCH(kCeConst);
#undef MIFDTD_FUSED2_LAUNCH
        mark(1);
        ++nsweeps;
        MIFDTD_CHECK_LAUNCH("two-step sweep launch");
        if (int rc = shell_h(sb2[0], tot2[0], out, mid, out, ss2, sp2)) return rc;
        swap_sets();
        ++swaps;
    }
    // single steps: shell E, the one-step sweep, shell H
    for (; s < nsteps; s++) {
        FSrc fs;
        SSrc ss;
        FProbe fp;
        SProbe sp;
        step_tables(s, fs, ss, fp, sp);
        if (tfsf)  // the incident line's E half: the E patches of this step read its H
            if (int rc = mifdtd_tfsf_line(d->tfsf_line_e, d->tfsf_line_h, d->tfsf_abc, d->tfsf_line_n, 1, d->tfsf_src[s],
                                          stream))
                return rc;
        if (shell_on)
            if (int rc = shell_e(sb, total, in, in, out, pin, pout, ss, sp, feat)) return rc;
        if (sweep) {
            mark(0);
#define MIFDTD_FUSED_LAUNCH_W(CEV, FV, W)                                                                            \
    k_fused<CEV, FV, W><<<g.nblocks, nthr, lds_bytes, st>>>(g, in[0], in[1], in[2], in[3], in[4], in[5], out[0], out[1], \
                                                            out[2], out[3], out[4], out[5], d->ce[0], d->ce[1], d->ce[2], \
                                                            d->ce_mat[0], d->ce_mat[1], d->ce_mat[2],                     \
                                                            sweep_cem == kCeTriple ? d->ce_tri_lut : d->ce_lut, fs, fp,   \
                                                            sdr, tfe, tfh, d->ce_tri)
#define MIFDTD_FUSED_LAUNCH(CEV, FV)                                                                              \
    do {                                                                                                          \
        if (((FV) & kFeatDrude) || !wide) MIFDTD_FUSED_LAUNCH_W(CEV, FV, false);                                  \
        else MIFDTD_FUSED_LAUNCH_W(CEV, FV, !((FV) & kFeatDrude));                                                \
    } while (0)
            if (sweep_cem == kCeTriple) MIFDTD_BY_FEAT(MIFDTD_FUSED_LAUNCH, kCeTriple, feat);
            else if (cem == kCeTable) MIFDTD_BY_FEAT(MIFDTD_FUSED_LAUNCH, kCeTable, feat);
            else if (cem == kCeArrays) MIFDTD_BY_FEAT(MIFDTD_FUSED_LAUNCH, kCeArrays, feat);
            else MIFDTD_BY_FEAT(MIFDTD_FUSED_LAUNCH, kCeConst, feat);
#undef MIFDTD_FUSED_LAUNCH
#undef MIFDTD_FUSED_LAUNCH_W
            mark(1);
            ++nsweeps;
            MIFDTD_CHECK_LAUNCH("fused sweep launch");
        }
        for (int q = 0; q < d->npbc; q++) {  // periodic E copies (priority after the incident corrections)
            const int a = d->pbc_axes[q];
            const int c1 = a == 0 ? 1 : 0, c2 = a == 2 ? 1 : 2;  // the E components tangential to axis a
            if (int rc = mifdtd_pbc(out[c1], out[c2], nx, ny, nzp, pad, a, 1, stream)) return rc;
        }
        if (shell_on)
            if (int rc = shell_h(sb, total, out, in, out, ss, sp)) return rc;
        for (int q = 0; q < d->npbc; q++) {  // periodic H copies
            const int a = d->pbc_axes[q];
            const int c1 = a == 0 ? 4 : 3, c2 = a == 2 ? 4 : 5;
            if (int rc = mifdtd_pbc(out[c1], out[c2], nx, ny, nzp, pad, a, 0, stream)) return rc;
        }
        if (tfsf)  // the line's H half (the H patches of this step read its E)
            if (int rc = mifdtd_tfsf_line(d->tfsf_line_e, d->tfsf_line_h, d->tfsf_abc, d->tfsf_line_n, 0, 0.f, stream))
                return rc;
        swap_sets();
        ++swaps;
    }
    if (parity) *parity = swaps & 1;
    if (sweep_ms) {
        float total_ms = 0.f;
        hipError_t err = hipStreamSynchronize(st);
        for (int q = 0; err == hipSuccess && q < nsweeps; q++) {
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
