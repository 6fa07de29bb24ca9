// Batched whole-step engine: many independent Yee + CPML simulations of uneven sizes
// advanced together, two launches per step for the whole batch (E pass of every
// instance, then H pass of every instance).
//
// A launch's blocks are the concatenation of each instance's blocks (instance order);
// block b finds its instance by binary search in the block-prefix table and runs the
// same per-quad E/H update as the slab engine (core + CPML + sources + probes), so small
// instances no longer leave the GPU idle between launches and a batch of N instances
// costs 2 launches per step instead of 2N.  Per-instance descriptors live in device
// memory (built once by mifdtd_batch_create); the per-step source values and probe
// outputs are rows of device arrays indexed by the step.  Optionally the launches of a
// run are captured once into a hipGraph and replayed as one graph launch.

#include "step_quad.hpp"

#include <mutex>
#include <vector>

namespace mifdtd {

struct BInst {
    SGeom g;          // whole instance: i0 = 0, nquads = nx*ny*nz_pitch/4
    SFields F;
    SCpml cp;
    SSrc se, sh;      // E / H sources with tval unset
    int tv_e[kSMaxSrc], tv_h[kSMaxSrc];  // column of each source's value in a tvals row
    SProbe pe, ph;    // probes with out unset; slot = column in a probe row
};

// Renumber instance-local block b so each XCD takes R consecutive blocks in turn.
__device__ __forceinline__ unsigned xcd_rows(unsigned b, unsigned nb, unsigned R)
{
    if (R == 0) return b;
    const unsigned span = 8u * R;
    const unsigned full = nb / span * span;
    if (b >= full) return b;
    const unsigned x = b & 7u, sl = b >> 3;
    return (sl / R) * span + x * R + sl % R;
}

// CM: CeMode of the E pass (kCeTable: the instance's material indices and table).
template <int CM, bool PML, bool EHALF>
__global__ __launch_bounds__(SB) void k_batch(const BInst *__restrict__ inst, const unsigned *__restrict__ first_block,
                                             int ninst, const float *__restrict__ tv_row, float *probe_row,
                                             unsigned remap_rows)
{
    const unsigned b = blockIdx.x;
    int lo = 0, hi = ninst;  // first_block[lo] <= b < first_block[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first_block[mid] <= b) lo = mid;
        else hi = mid;
    }
    const BInst &I = inst[lo];
    const float *tab = I.F.lut;
    const unsigned nb = first_block[lo + 1] - first_block[lo];
    const unsigned f = xcd_rows(b - first_block[lo], nb, remap_rows) * SB + threadIdx.x;
    if (f >= I.g.nquads) return;
    SSrc src = EHALF ? I.se : I.sh;
#pragma unroll
    for (int m = 0; m < kSMaxSrc; ++m)
        if (m < src.n) src.tval[m] = tv_row[EHALF ? I.tv_e[m] : I.tv_h[m]];
    SProbe prb = EHALF ? I.pe : I.ph;
    prb.out = probe_row;
    if (EHALF)
        e_quad<CM, PML>(I.g, f, I.F.e[0], I.F.e[1], I.F.e[2], I.F.e[3], I.F.e[4], I.F.e[5], I.F.ce[0], I.F.ce[1],
                        I.F.ce[2], I.cp, src, prb, I.F.cm[0], I.F.cm[1], I.F.cm[2], tab);
    else
        h_quad<PML>(I.g, f, I.F.e[0], I.F.e[1], I.F.e[2], I.F.e[3], I.F.e[4], I.F.e[5], I.cp, src, prb);
}

}  // namespace mifdtd

using namespace mifdtd;

// What the launches of a batch need besides the device descriptors.
struct BMeta {
    int ninst = 0;
    int cem = kCeConst;
    bool pml = false;
    unsigned total_blocks = 0;
    int total_src = 0, total_probe = 0;
};

// Validate the instances and fill h[0..ninst) and first[0..ninst] (host arrays).
static int build_batch(const mifdtd_instance *inst, int ninst, const float *bE, const float *aE, const float *bH,
                       const float *aH, BInst *h, unsigned *first, BMeta &m)
{
    if (!inst || ninst < 1) return fail(MIFDTD_EINVAL, "need at least one instance");
    const bool hc = inst[0].ce[0] != nullptr;
    const bool tbl = inst[0].ce_mat[0] != nullptr;
    bool pml = false;
    int batch_npml = 0;  // the one npml of the shared 2*(npml+1)-entry tables
    int tsrc = 0, tprobe = 0;
    unsigned long long blocks = 0;
    for (int n = 0; n < ninst; n++) {
        const mifdtd_instance &d = inst[n];
        const int nx = d.nx, ny = d.ny, nzp = d.nz_pitch, pad = d.pad;
        if (int rc = check_grid(nx, ny, nzp, pad)) return fail(rc, "instance %d: %s", n, mifdtd_last_error());
        for (int c = 0; c < 6; c++)
            if (!d.e[c] || !aligned16(d.e[c])) return fail(MIFDTD_EALIGN, "instance %d: field %d NULL or misaligned", n, c);
        if ((d.ce[0] != nullptr) != hc || (hc && (!d.ce[1] || !d.ce[2])))
            return fail(MIFDTD_ENULL, "instance %d: CE arrays must be set on every instance or on none", n);
        if ((d.ce_mat[0] != nullptr) != tbl)
            return fail(MIFDTD_ENULL, "instance %d: CE material indices must be set on every instance or on none", n);
        if (tbl) {
            if (!d.ce_mat[1] || !d.ce_mat[2] || !d.ce_lut)
                return fail(MIFDTD_ENULL, "instance %d: CE material indices need all three index arrays and the table", n);
            for (int q = 0; q < 3; q++)
                if (reinterpret_cast<uintptr_t>(d.ce_mat[q]) & 3u)
                    return fail(MIFDTD_EALIGN, "instance %d: CE material index arrays not 4-byte aligned", n);
        }
        const int nz = nzp - pad;
        BInst &B = h[n];
        B = BInst{};
        B.g.nx = nx;
        B.g.ny = ny;
        B.g.nzp = nzp;
        B.g.nz = nz;
        B.g.Q = (unsigned)(nzp / 4);
        B.g.i0 = 0;
        B.g.nquads = (unsigned)((unsigned long long)nx * ny * (nzp / 4));
        B.g.remap = 0;
        for (int c = 0; c < 6; c++) B.F.e[c] = d.e[c];
        for (int c = 0; c < 3; c++) B.F.ce[c] = d.ce[c];
        if (tbl) {
            for (int c = 0; c < 3; c++) B.F.cm[c] = d.ce_mat[c];
            B.F.lut = d.ce_lut;
        }
        if (d.npml > 0) {
            const int np_ = d.npml;
            if (np_ > MIFDTD_MAX_NPML) return fail(MIFDTD_EINVAL, "instance %d: npml %d too large", n, np_);
            if (batch_npml && np_ != batch_npml)
                return fail(MIFDTD_EINVAL, "instance %d: npml %d differs from the batch's %d (the tables are shared)", n,
                            np_, batch_npml);
            batch_npml = np_;
            if (nx < 2 * np_ + 2 || ny < 2 * np_ + 2 || nz < 2 * np_ + 2)
                return fail(MIFDTD_EINVAL, "instance %d: grid too small for npml %d", n, np_);
            if (!bE || !aE || !bH || !aH) return fail(MIFDTD_ENULL, "CPML tables are NULL");
            if (!d.psiE || !d.psiH || !aligned16(d.psiE) || !aligned16(d.psiH))
                return fail(MIFDTD_EALIGN, "instance %d: psi buffers NULL or misaligned", n);
            pml = true;
            B.cp.npml = np_;
            B.cp.axes = 7;
            B.cp.bE = bE;
            B.cp.aE = aE;
            B.cp.bH = bH;
            B.cp.aH = aH;
            B.cp.pe = d.psiE;
            B.cp.pe_out = d.psiE;
            B.cp.ph = d.psiH;
            B.cp.sX = (size_t)np_ * ny * nzp;
            B.cp.sY = (size_t)nx * np_ * nzp;
            B.cp.sZ = (size_t)nx * ny * np_;
        }
        if (d.nsrc < 0 || d.nsrc > MIFDTD_FUSED_MAX_SRC) return fail(MIFDTD_EINVAL, "instance %d: nsrc", n);
        for (int s = 0; s < d.nsrc; s++) {
            const int *bx = d.src_box[s];
            const int c = d.src_comp[s];
            if (c < 0 || c > 5 || bx[0] < 0 || bx[2] < 0 || bx[4] < 0 || bx[1] >= nx || bx[3] >= ny || bx[5] >= nzp ||
                bx[0] > bx[1] || bx[2] > bx[3] || bx[4] > bx[5])
                return fail(MIFDTD_EINVAL, "instance %d: source %d outside the grid", n, s);
            SSrc &X = c < 3 ? B.se : B.sh;
            int *tv = c < 3 ? B.tv_e : B.tv_h;
            const int m = X.n++;
            X.comp[m] = c;
            X.i0[m] = bx[0];
            X.i1[m] = bx[1];
            X.j0[m] = bx[2];
            X.j1[m] = bx[3];
            X.k0[m] = bx[4];
            X.k1[m] = bx[5];
            X.sv[m] = d.src_sv[s];
            X.svs[m] = d.src_svs[s];
            X.ow[m] = d.src_overwrite[s];
            tv[m] = tsrc + s;
        }
        tsrc += d.nsrc;
        if (d.nprobe < 0 || d.nprobe > MIFDTD_FUSED_MAX_PROBE) return fail(MIFDTD_EINVAL, "instance %d: nprobe", n);
        for (int q = 0; q < d.nprobe; q++) {
            const int *p = d.probe_ijk[q];
            const int c = d.probe_comp[q];
            if (c < 0 || c > 5 || p[0] < 0 || p[0] >= nx || p[1] < 0 || p[1] >= ny || p[2] < 0 || p[2] >= nzp)
                return fail(MIFDTD_EINVAL, "instance %d: probe %d outside the grid", n, q);
            SProbe &P = c < 3 ? B.pe : B.ph;
            const int m = P.n++;
            P.comp[m] = c;
            P.i[m] = p[0];
            P.j[m] = p[1];
            P.k[m] = p[2];
            P.slot[m] = tprobe + q;
        }
        tprobe += d.nprobe;
        first[n] = (unsigned)blocks;
        blocks += (B.g.nquads + SB - 1) / SB;
        if (blocks >= (1ull << 31)) return fail(MIFDTD_ERANGE, "batch exceeds 2^31 blocks");
    }
    first[ninst] = (unsigned)blocks;
    m.ninst = ninst;
    m.cem = tbl ? kCeTable : hc ? kCeArrays : kCeConst;
    m.pml = pml;
    m.total_blocks = (unsigned)blocks;
    m.total_src = tsrc;
    m.total_probe = tprobe;
    return 0;
}

static int batch_launches(const BMeta &b, const BInst *d_inst, const unsigned *d_first, int nsteps, const float *tvals,
                          float *probe_series, unsigned R, hipStream_t st)
{
    const unsigned grid = b.total_blocks;
    for (int s = 0; s < nsteps; s++) {
        const float *tv = b.total_src ? tvals + (size_t)s * b.total_src : nullptr;
        float *pr = b.total_probe ? probe_series + (size_t)s * b.total_probe : nullptr;
#define MIFDTD_BATCH_LAUNCH(CM, PML, EH) \
    k_batch<CM, PML, EH><<<grid, SB, 0, st>>>(d_inst, d_first, b.ninst, tv, pr, R)
#define MIFDTD_BATCH_E(CM)                                   \
    do {                                                     \
        if (b.pml) MIFDTD_BATCH_LAUNCH(CM, true, true);      \
        else MIFDTD_BATCH_LAUNCH(CM, false, true);           \
    } while (0)
        if (b.cem == kCeTable) MIFDTD_BATCH_E(kCeTable);
        else if (b.cem == kCeArrays) MIFDTD_BATCH_E(kCeArrays);
        else MIFDTD_BATCH_E(kCeConst);
        MIFDTD_CHECK_LAUNCH("batch E launch");
        if (b.pml) MIFDTD_BATCH_LAUNCH(kCeConst, true, false);
        else MIFDTD_BATCH_LAUNCH(kCeConst, false, false);
#undef MIFDTD_BATCH_E
        MIFDTD_CHECK_LAUNCH("batch H launch");
#undef MIFDTD_BATCH_LAUNCH
    }
    return 0;
}

static int check_run_args(const BMeta &b, int nsteps, const float *tvals, float *probe_series, int remap_rows)
{
    if (nsteps < 0) return fail(MIFDTD_EINVAL, "nsteps < 0");
    if (b.total_src && !tvals) return fail(MIFDTD_ENULL, "tvals is NULL");
    if (b.total_probe && !probe_series) return fail(MIFDTD_ENULL, "probe_series is NULL");
    if (remap_rows < 0 || remap_rows > 4096) return fail(MIFDTD_EINVAL, "remap_rows %d outside [0,4096]", remap_rows);
    return 0;
}

// ------------------------------------------------------------ allocation-free run
// mifdtd_batch_run_static keeps the descriptors in this file-scope device array instead
// of device memory from hipMalloc: one copy per GPU, held until the process ends, shared
// by every call on that GPU (overlapping calls on one GPU race on it).  Host code reaches
// it only through the address hipGetSymbolAddress returns for the current device (the
// array's name is a host-side stand-in, not a device address).
constexpr int kStaticMaxInst = MIFDTD_BATCH_STATIC_MAX_INST;
constexpr size_t kStaticFirstOff = ((size_t)kStaticMaxInst * sizeof(BInst) + 255) & ~(size_t)255;
constexpr size_t kStaticBytes = kStaticFirstOff + sizeof(unsigned) * (kStaticMaxInst + 1);
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// The current device's address of `scratch` (looked up once per device).
static int scratch_ptr(unsigned char **out)
{
    int dev = 0;
    if (hipError_t e = hipGetDevice(&dev)) return hip_status(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(MIFDTD_ERANGE, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) return hip_status(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(MIFDTD_ENULL, "hipGetSymbolAddress(scratch) gave NULL");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    return 0;
}

// Host bytes reach the scratch array as kernel arguments of k_put (2 KiB per launch), so
// no host buffer has to outlive the call and no staging memory is involved.
constexpr int kPutWords = 512;
struct PutChunk {
    unsigned w[kPutWords];
};

__global__ __launch_bounds__(256) void k_put(unsigned *__restrict__ dst, PutChunk c, int nw)
{
    for (int t = threadIdx.x; t < nw; t += 256) dst[t] = c.w[t];
}

static int put_bytes(unsigned char *dst, const void *src, size_t bytes, hipStream_t st)
{
    const unsigned *w = static_cast<const unsigned *>(src);
    const size_t nw = (bytes + 3) / 4;  // whole words: the sources are word-sized structs and arrays
    for (size_t o = 0; o < nw; o += kPutWords) {
        PutChunk c;
        const int n = (int)(nw - o < (size_t)kPutWords ? nw - o : (size_t)kPutWords);
        for (int t = 0; t < n; t++) c.w[t] = w[o + t];
        k_put<<<1, 256, 0, st>>>(reinterpret_cast<unsigned *>(dst) + o, c, n);
        MIFDTD_CHECK_LAUNCH("descriptor put");
    }
    return 0;
}

struct mifdtd_batch {
    BMeta m;
    BInst *d_inst = nullptr;
    unsigned *d_first = nullptr;
};

extern "C" {

int mifdtd_batch_create(const mifdtd_instance *inst, int ninst, const float *bE, const float *aE, const float *bH,
                        const float *aH, mifdtd_batch **out)
{
    if (!out) return fail(MIFDTD_ENULL, "out is NULL");
    *out = nullptr;
    if (!inst || ninst < 1) return fail(MIFDTD_EINVAL, "need at least one instance");
    std::vector<BInst> h(ninst);
    std::vector<unsigned> first(ninst + 1, 0);
    BMeta m;
    if (int rc = build_batch(inst, ninst, bE, aE, bH, aH, h.data(), first.data(), m)) return rc;
    auto *bt = new mifdtd_batch;
    bt->m = m;
    int rc = hip_status(hipMalloc(&bt->d_inst, sizeof(BInst) * ninst), "batch descriptor alloc");
    if (!rc) rc = hip_status(hipMalloc(&bt->d_first, sizeof(unsigned) * (ninst + 1)), "batch prefix alloc");
    if (!rc) rc = hip_status(hipMemcpy(bt->d_inst, h.data(), sizeof(BInst) * ninst, hipMemcpyHostToDevice), "batch upload");
    if (!rc)
        rc = hip_status(hipMemcpy(bt->d_first, first.data(), sizeof(unsigned) * (ninst + 1), hipMemcpyHostToDevice),
                        "batch upload");
    if (rc) {
        if (bt->d_inst) (void)hipFree(bt->d_inst);
        if (bt->d_first) (void)hipFree(bt->d_first);
        delete bt;
        return rc;
    }
    *out = bt;
    return 0;
}

int mifdtd_batch_sizes(const mifdtd_batch *b, int *total_src, int *total_probe, unsigned *total_blocks)
{
    if (!b) return fail(MIFDTD_ENULL, "batch is NULL");
    if (total_src) *total_src = b->m.total_src;
    if (total_probe) *total_probe = b->m.total_probe;
    if (total_blocks) *total_blocks = b->m.total_blocks;
    return 0;
}

int mifdtd_batch_run(mifdtd_batch *b, int nsteps, const float *tvals, float *probe_series, int use_graph,
                     int remap_rows, mifdtd_stream_t stream)
{
    if (!b) return fail(MIFDTD_ENULL, "batch is NULL");
    if (int rc = check_run_args(b->m, nsteps, tvals, probe_series, remap_rows)) return rc;
    if (nsteps == 0) return 0;
    hipStream_t st = to_stream(stream);
    if (!use_graph) return batch_launches(b->m, b->d_inst, b->d_first, nsteps, tvals, probe_series, (unsigned)remap_rows, st);
    if (!st) return fail(MIFDTD_EINVAL, "graph capture needs a non-null stream");
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = hip_status(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    const int lrc = batch_launches(b->m, b->d_inst, b->d_first, nsteps, tvals, probe_series, (unsigned)remap_rows, st);
    rc = hip_status(hipStreamEndCapture(st, &graph), "end capture");
    if (lrc) {
        if (graph) (void)hipGraphDestroy(graph);
        return lrc;
    }
    if (!rc) rc = hip_status(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    if (!rc) rc = hip_status(hipGraphLaunch(exec, st), "graph launch");
    // the executable graph is released once its launch has completed
    if (!rc) rc = hip_status(hipStreamSynchronize(st), "graph wait");
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

int mifdtd_batch_destroy(mifdtd_batch *b)
{
    if (!b) return 0;
    int rc = 0;
    if (b->d_inst) rc = hip_status(hipFree(b->d_inst), "batch free");
    if (b->d_first && !rc) rc = hip_status(hipFree(b->d_first), "batch free");
    delete b;
    return rc;
}

int mifdtd_batch_run_static(const mifdtd_instance *inst, int ninst, const float *bE, const float *aE, const float *bH,
                            const float *aH, int nsteps, const float *tvals, float *probe_series, int remap_rows,
                            mifdtd_stream_t stream)
{
    static_assert(sizeof(BInst) % 4 == 0, "descriptors are copied as words");
    if (ninst > kStaticMaxInst)
        return fail(MIFDTD_ERANGE, "%d instances: the allocation-free run holds at most %d (MIFDTD_BATCH_STATIC_MAX_INST)",
                    ninst, kStaticMaxInst);
    static thread_local BInst h[kStaticMaxInst];
    static thread_local unsigned first[kStaticMaxInst + 1];
    BMeta m;
    if (int rc = build_batch(inst, ninst, bE, aE, bH, aH, h, first, m)) return rc;
    if (int rc = check_run_args(m, nsteps, tvals, probe_series, remap_rows)) return rc;
    if (nsteps == 0) return 0;
    unsigned char *base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    hipStream_t st = to_stream(stream);
    if (int rc = put_bytes(base, h, sizeof(BInst) * ninst, st)) return rc;
    if (int rc = put_bytes(base + kStaticFirstOff, first, sizeof(unsigned) * (ninst + 1), st)) return rc;
    return batch_launches(m, reinterpret_cast<const BInst *>(base), reinterpret_cast<const unsigned *>(base + kStaticFirstOff),
                          nsteps, tvals, probe_series, (unsigned)remap_rows, st);
}

}  // extern "C"
