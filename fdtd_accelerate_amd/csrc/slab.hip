// Slab-ordered step: per time step, for each slab of consecutive x planes, one E pass
// (core E + CPML-E + E sources + E probes) then one H pass (core H + CPML-H + H sources
// + H probes) over that slab, in place; by default the slab is the whole grid (two
// launches per step).  Smaller slabs keep a slab's E/H planes in the last-level cache
// between its two passes, which on MI355X was measured to gain less than the extra
// launches cost.
//
// Ordering is the reference's (E of the slab before H of the slab; slabs in increasing
// x): H of plane i needs E_new of planes i and i-1 (done by this or the previous slab's
// E pass) and E of plane i needs H_old of planes i and i+1 (not yet touched: the slab's
// H pass runs after its E pass, the next slab's later).

#include "step_quad.hpp"

namespace mifdtd {

// CM: CeMode (kCeTable: CE from the material indices and the 3 x 256 table).
template <int CM, bool PML>
__global__ __launch_bounds__(SB) void k_slab_e(SGeom g, SFields F, SCpml cp, SSrc src, SProbe prb)
{
    const float *tab = F.lut;
    const unsigned f = (g.remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x) * SB + threadIdx.x;
    if (f < g.nquads)
        e_quad<CM, PML>(g, f, F.e[0], F.e[1], F.e[2], F.e[3], F.e[4], F.e[5], F.ce[0], F.ce[1], F.ce[2], cp, src, prb,
                        F.cm[0], F.cm[1], F.cm[2], tab);
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
        cp.axes = 7;
        cp.bE = d->bE;
        cp.aE = d->aE;
        cp.bH = d->bH;
        cp.aH = d->aH;
        cp.pe = d->psiE;
        cp.pe_out = d->psiE;
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
    SFields F{};
    for (int c = 0; c < 6; c++) F.e[c] = d->e[c];
    for (int c = 0; c < 3; c++) F.ce[c] = d->ce[c];
    if (tbl) {
        for (int c = 0; c < 3; c++) F.cm[c] = d->ce_mat[c];
        F.lut = d->ce_lut;
    }
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
#define MIFDTD_SLAB_E(CM)                                                   \
    do {                                                                    \
        if (pml) k_slab_e<CM, true><<<grid, SB, 0, st>>>(g, F, cp, se, pen);  \
        else k_slab_e<CM, false><<<grid, SB, 0, st>>>(g, F, cp, se, pen);    \
    } while (0)
            if (cem == kCeTable) MIFDTD_SLAB_E(kCeTable);
            else if (cem == kCeArrays) MIFDTD_SLAB_E(kCeArrays);
            else MIFDTD_SLAB_E(kCeConst);
#undef MIFDTD_SLAB_E
            MIFDTD_CHECK_LAUNCH("slab E launch");
            if (pml) k_slab_h<true><<<grid, SB, 0, st>>>(g, F, cp, sh, phn);
            else k_slab_h<false><<<grid, SB, 0, st>>>(g, F, cp, sh, phn);
            MIFDTD_CHECK_LAUNCH("slab H launch");
        }
    }
    return 0;
}

}  // extern "C"
