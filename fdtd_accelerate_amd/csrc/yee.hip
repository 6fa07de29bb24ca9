// In-place Yee kernels behind the reference's own call surface: the KEMP core
// update_e/update_h (K7/cpu/src/core.c), kuFDTD's CPML face updates
// (KU/cpu/core/cpml_non_kappa.c), the IncidentDirect source add and the probe gather.
//
// Layout: C-order (nx, ny, nzp) float32, z contiguous.  One thread owns one aligned
// z-quad (4 floats, one 16-byte load/store per array), consecutive lanes own
// consecutive quads so a 64-lane wave streams 1 KiB contiguous per array.  The
// k+1 / k-1 neighbour that falls outside a thread's quad is a scalar load that hits
// L1/L2 (the neighbouring lane loaded it); as in the reference's flat loop, the k+1
// neighbour of a row's last element is the next row's first (k-1 of the first: the
// previous row's last), which only feeds masked lanes.  Masked lanes store value*0
// like the reference's SSE mask multiply, and arithmetic keeps the reference's
// evaluation order under -ffp-contract=off, so results are bit-identical (including
// the sign of masked zeros) to the reference's code.

#include <cstdarg>
#include <mutex>

#include "common.hpp"

namespace mifdtd {

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_status(hipError_t e, const char *where)
{
    if (e == hipSuccess) return 0;
    return fail(static_cast<int>(e), "%s: %s", where, hipGetErrorString(e));
}

int check_grid(int nx, int ny, int nzp, int pad)
{
    if (nx < 2 || ny < 2 || nzp < 4) return fail(MIFDTD_EINVAL, "grid (%d,%d,%d) too small", nx, ny, nzp);
    if (nzp % 4) return fail(MIFDTD_EALIGN, "nz_pitch %d is not a multiple of 4", nzp);
    if (pad < 0 || pad > nzp - 2) return fail(MIFDTD_EINVAL, "pad %d invalid for nz_pitch %d", pad, nzp);
    const unsigned long long quads = (unsigned long long)nx * ny * (nzp / 4);
    if (quads >= (1ull << 31)) return fail(MIFDTD_ERANGE, "grid (%d,%d,%d) exceeds 2^31 quads", nx, ny, nzp);
    return 0;
}

// ------------------------------------------------------------------ core kernels
constexpr int kBlock = 256;

template <bool HAS_C>
__global__ __launch_bounds__(kBlock) void k_update_e(
    int nx, int ny, int nzp, int nz, unsigned q0, unsigned q1, unsigned Q,
    float *__restrict__ ex, float *__restrict__ ey, float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez)
{
    const unsigned f = q0 + blockIdx.x * kBlock + threadIdx.x;  // quads [q0, q1) of the grid
    if (f >= q1) return;
    const unsigned row = f / Q;
    const int k0 = (int)(f - row * Q) * 4;
    const int i = (int)(row / (unsigned)ny);
    const int j = (int)row - i * ny;
    const size_t id = (size_t)row * nzp + k0;
    const size_t sx = (size_t)ny * nzp;

    const Q4 hx0 = ld4(hx + id), hy0 = ld4(hy + id), hz0 = ld4(hz + id);

    if (j < ny - 1) {
        const Q4 e = ld4(ex + id);
        const Q4 c = HAS_C ? ld4(cex + id) : splat4(0.5f);
        const Q4 h1 = ld4(hz + id + nzp);
        const float hyn = hy[id + 4];  // next row's first element on the last quad (j+1 exists)
        Q4 r;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const float h2 = l < 3 ? hy0.v[l + 1] : hyn;
            const float v = e.v[l] + c.v[l] * ((h1.v[l] - hz0.v[l]) - (h2 - hy0.v[l]));
            r.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
        st4(ex + id, r);
    }
    if (i < nx - 1) {
        const Q4 e = ld4(ey + id);
        const Q4 c = HAS_C ? ld4(cey + id) : splat4(0.5f);
        const Q4 h2 = ld4(hz + id + sx);
        const float hxn = hx[id + 4];  // next row (or plane) start on the last quad (i+1 exists)
        Q4 r;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const float h1 = l < 3 ? hx0.v[l + 1] : hxn;
            const float v = e.v[l] + c.v[l] * ((h1 - hx0.v[l]) - (h2.v[l] - hz0.v[l]));
            r.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
        st4(ey + id, r);
        if (j < ny - 1) {
            const Q4 e3 = ld4(ez + id);
            const Q4 c3 = HAS_C ? ld4(cez + id) : splat4(0.5f);
            const Q4 g1 = ld4(hy + id + sx);
            const Q4 g2 = ld4(hx + id + nzp);
            Q4 r3;
#pragma unroll
            for (int l = 0; l < 4; l++) {
                const float v = e3.v[l] + c3.v[l] * ((g1.v[l] - hy0.v[l]) - (g2.v[l] - hx0.v[l]));
                r3.v[l] = (k0 + l < nz) ? v : v * 0.f;
            }
            st4(ez + id, r3);
        }
    }
}

// update_e with the CE arrays held as per-component uint8 material indices into a
// 3 x 256 float table (the exact float32 CE values; staged in LDS): 39 instead of 48
// bytes per cell, same arithmetic and masks as k_update_e<true>.
__device__ __forceinline__ Q4 lut4(const float *tab, const unsigned char *m)
{
    const uchar4 t = *reinterpret_cast<const uchar4 *>(m);
    Q4 q;
    q.v[0] = tab[t.x];
    q.v[1] = tab[t.y];
    q.v[2] = tab[t.z];
    q.v[3] = tab[t.w];
    return q;
}

__global__ __launch_bounds__(kBlock) void k_update_e_lut(
    int nx, int ny, int nzp, int nz, unsigned nquads, unsigned Q,
    float *__restrict__ ex, float *__restrict__ ey, float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    const unsigned char *__restrict__ mx, const unsigned char *__restrict__ my, const unsigned char *__restrict__ mz,
    const float *__restrict__ lut)
{
    __shared__ float tab[3 * 256];
    for (int t = threadIdx.x; t < 3 * 256; t += kBlock) tab[t] = lut[t];
    __syncthreads();
    const unsigned f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= nquads) return;
    const unsigned row = f / Q;
    const int k0 = (int)(f - row * Q) * 4;
    const int i = (int)(row / (unsigned)ny);
    const int j = (int)row - i * ny;
    const size_t id = (size_t)row * nzp + k0;
    const size_t sx = (size_t)ny * nzp;

    const Q4 hx0 = ld4(hx + id), hy0 = ld4(hy + id), hz0 = ld4(hz + id);

    if (j < ny - 1) {
        const Q4 e = ld4(ex + id);
        const Q4 c = lut4(tab, mx + id);
        const Q4 h1 = ld4(hz + id + nzp);
        const float hyn = hy[id + 4];
        Q4 r;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const float h2 = l < 3 ? hy0.v[l + 1] : hyn;
            const float v = e.v[l] + c.v[l] * ((h1.v[l] - hz0.v[l]) - (h2 - hy0.v[l]));
            r.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
        st4(ex + id, r);
    }
    if (i < nx - 1) {
        const Q4 e = ld4(ey + id);
        const Q4 c = lut4(tab + 256, my + id);
        const Q4 h2 = ld4(hz + id + sx);
        const float hxn = hx[id + 4];
        Q4 r;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const float h1 = l < 3 ? hx0.v[l + 1] : hxn;
            const float v = e.v[l] + c.v[l] * ((h1 - hx0.v[l]) - (h2.v[l] - hz0.v[l]));
            r.v[l] = (k0 + l < nz - 1) ? v : v * 0.f;
        }
        st4(ey + id, r);
        if (j < ny - 1) {
            const Q4 e3 = ld4(ez + id);
            const Q4 c3 = lut4(tab + 512, mz + id);
            const Q4 g1 = ld4(hy + id + sx);
            const Q4 g2 = ld4(hx + id + nzp);
            Q4 r3;
#pragma unroll
            for (int l = 0; l < 4; l++) {
                const float v = e3.v[l] + c3.v[l] * ((g1.v[l] - hy0.v[l]) - (g2.v[l] - hx0.v[l]));
                r3.v[l] = (k0 + l < nz) ? v : v * 0.f;
            }
            st4(ez + id, r3);
        }
    }
}

template <bool HAS_C>
__global__ __launch_bounds__(kBlock) void k_update_h(
    int nx, int ny, int nzp, int nz, unsigned q0, unsigned q1, unsigned Q,
    const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
    const float *__restrict__ chx, const float *__restrict__ chy, const float *__restrict__ chz)
{
    (void)nx;
    const unsigned f = q0 + blockIdx.x * kBlock + threadIdx.x;  // quads [q0, q1) of the grid
    if (f >= q1) return;
    const unsigned row = f / Q;
    const int k0 = (int)(f - row * Q) * 4;
    const int i = (int)(row / (unsigned)ny);
    const int j = (int)row - i * ny;
    const size_t id = (size_t)row * nzp + k0;
    const size_t sx = (size_t)ny * nzp;

    const Q4 ex0 = ld4(ex + id), ey0 = ld4(ey + id), ez0 = ld4(ez + id);

    if (j > 0) {
        const Q4 h = ld4(hx + id);
        const Q4 c = HAS_C ? ld4(chx + id) : splat4(0.5f);
        const Q4 e1 = ld4(ez + id - nzp);
        const float eyp = ey[id - 1];  // previous row's last element when k0 == 0 (j-1 exists)
        Q4 r;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const float e2 = l > 0 ? ey0.v[l - 1] : eyp;
            const float v = h.v[l] - c.v[l] * ((ez0.v[l] - e1.v[l]) - (ey0.v[l] - e2));
            const int k = k0 + l;
            r.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
        }
        st4(hx + id, r);
    }
    if (i > 0) {
        const Q4 h = ld4(hy + id);
        const Q4 c = HAS_C ? ld4(chy + id) : splat4(0.5f);
        const Q4 e2 = ld4(ez + id - sx);
        const float exp_ = ex[id - 1];  // previous row's last element when k0 == 0 (i-1 exists)
        Q4 r;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const float e1 = l > 0 ? ex0.v[l - 1] : exp_;
            const float v = h.v[l] - c.v[l] * ((ex0.v[l] - e1) - (ez0.v[l] - e2.v[l]));
            const int k = k0 + l;
            r.v[l] = (k > 0 && k < nz) ? v : v * 0.f;
        }
        st4(hy + id, r);
        if (j > 0) {
            const Q4 h3 = ld4(hz + id);
            const Q4 c3 = HAS_C ? ld4(chz + id) : splat4(0.5f);
            const Q4 g1 = ld4(ey + id - sx);
            const Q4 g2 = ld4(ex + id - nzp);
            Q4 r3;
#pragma unroll
            for (int l = 0; l < 4; l++) {
                const float v = h3.v[l] - c3.v[l] * ((ey0.v[l] - g1.v[l]) - (ex0.v[l] - g2.v[l]));
                r3.v[l] = (k0 + l < nz) ? v : v * 0.f;
            }
            st4(hz + id, r3);
        }
    }
}

// idx0 / nmax: the element range [idx0, nmax) of the split form (nmax < 0: the whole
// grid); a quad is updated when its first element lies in the range, as the reference's
// SSE loop does, so idx0 must start a quad.
static int launch_core(bool e_half, int nx, int ny, int nzp, int pad,
                       float *ex, float *ey, float *ez, float *hx, float *hy, float *hz,
                       const float *c0, const float *c1, const float *c2, hipStream_t st,
                       long long idx0 = 0, long long nmax = -1)
{
    if (int rc = check_grid(nx, ny, nzp, pad)) return rc;
    const float *ptrs[6] = {ex, ey, ez, hx, hy, hz};
    for (const float *p : ptrs) {
        if (!p) return fail(MIFDTD_ENULL, "field pointer is NULL");
        if (!aligned16(p)) return fail(MIFDTD_EALIGN, "field pointer %p is not 16-byte aligned", (const void *)p);
    }
    const bool has_c = c0 || c1 || c2;
    if (has_c) {
        if (!(c0 && c1 && c2)) return fail(MIFDTD_ENULL, "coefficient arrays must be all set or all NULL");
        if (!aligned16(c0) || !aligned16(c1) || !aligned16(c2))
            return fail(MIFDTD_EALIGN, "coefficient pointer is not 16-byte aligned");
    }
    const unsigned Q = (unsigned)(nzp / 4);
    const unsigned nquads = (unsigned)nx * ny * Q;
    unsigned q0 = 0, q1 = nquads;
    if (nmax >= 0) {
        const long long n = 4ll * nquads;
        if (idx0 < 0 || idx0 > nmax || nmax > n)
            return fail(MIFDTD_EINVAL, "element range [%lld, %lld) outside the grid's %lld", idx0, nmax, n);
        if (idx0 % 4) return fail(MIFDTD_EALIGN, "idx0 %lld does not start a 4-element vector", idx0);
        q0 = (unsigned)(idx0 / 4);
        q1 = (unsigned)((nmax + 3) / 4);
    }
    if (q1 == q0) return 0;
    const unsigned grid = (q1 - q0 + kBlock - 1) / kBlock;
    const int nz = nzp - pad;
    if (e_half) {
        if (has_c)
            k_update_e<true><<<grid, kBlock, 0, st>>>(nx, ny, nzp, nz, q0, q1, Q, ex, ey, ez, hx, hy, hz, c0, c1, c2);
        else
            k_update_e<false><<<grid, kBlock, 0, st>>>(nx, ny, nzp, nz, q0, q1, Q, ex, ey, ez, hx, hy, hz, c0, c1, c2);
        MIFDTD_CHECK_LAUNCH("update_e launch");
    } else {
        if (has_c)
            k_update_h<true><<<grid, kBlock, 0, st>>>(nx, ny, nzp, nz, q0, q1, Q, ex, ey, ez, hx, hy, hz, c0, c1, c2);
        else
            k_update_h<false><<<grid, kBlock, 0, st>>>(nx, ny, nzp, nz, q0, q1, Q, ex, ey, ez, hx, hy, hz, c0, c1, c2);
        MIFDTD_CHECK_LAUNCH("update_h launch");
    }
    return 0;
}

// ------------------------------------------------------------------ CPML kernels
// Validity masks of the core update (a CPML correction is applied only where the
// core writes): component c = 0,1,2 for x,y,z.
__device__ __forceinline__ bool valid_e(int c, int i, int j, int k, int nx, int ny, int nz)
{
    if (c == 0) return j < ny - 1 && k < nz - 1;
    if (c == 1) return i < nx - 1 && k < nz - 1;
    return i < nx - 1 && j < ny - 1 && k < nz;
}
__device__ __forceinline__ bool valid_h(int c, int i, int j, int k, int nz)
{
    if (c == 0) return j > 0 && k > 0 && k < nz;
    if (c == 1) return i > 0 && k > 0 && k < nz;
    return i > 0 && j > 0 && k < nz;
}

struct FaceSet {
    float *psi1[2], *psi2[2];  // [side]
};

// E half, x or y axis (AXIS 0/1): thread per slab quad, both sides in one grid.
// Work items: side-major, then p, then the in-plane index.
template <int AXIS, bool HAS_C>
__global__ __launch_bounds__(kBlock) void k_cpml_e_xy(
    int nx, int ny, int nzp, int nz, int npml, unsigned per_side, unsigned Q,
    float *__restrict__ ex, float *__restrict__ ey, float *__restrict__ ez,
    const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
    const float *__restrict__ cex, const float *__restrict__ cey, const float *__restrict__ cez,
    FaceSet fs, const float *__restrict__ tb, const float *__restrict__ ta)
{
    unsigned f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= 2 * per_side) return;
    const int side = f >= per_side;
    if (side) f -= per_side;
    float *psi1 = fs.psi1[side], *psi2 = fs.psi2[side];
    if (!psi1) return;
    // per_side = npml * (other extent) * Q
    const unsigned rowq = f / Q;
    const int k0 = (int)(f - rowq * Q) * 4;
    int i, j, p;
    const int n_e = AXIS == 0 ? nx : ny;
    if (AXIS == 0) {
        p = (int)(rowq / (unsigned)ny);
        j = (int)rowq - p * ny;
        i = side ? n_e - 1 - npml + p : p;
    } else {
        i = (int)(rowq / (unsigned)npml);
        p = (int)rowq - i * npml;
        j = side ? n_e - 1 - npml + p : p;
    }
    const int t = side ? npml + 1 + p : p + 1;
    const float b = tb[t], a = ta[t];
    const size_t id = ((size_t)i * ny + j) * nzp + k0;
    const size_t pid = (size_t)rowq * nzp + k0;
    const size_t s = AXIS == 0 ? (size_t)ny * nzp : (size_t)nzp;
    // x: psi1 -> ey from hz, psi2 -> ez from hy ; y: psi1 -> ez from hx, psi2 -> ex from hz
    float *e1 = AXIS == 0 ? ey : ez;
    float *e2 = AXIS == 0 ? ez : ex;
    const float *g1 = AXIS == 0 ? hz : hx;
    const float *g2 = AXIS == 0 ? hy : hz;
    const float *c1 = AXIS == 0 ? cey : cez;
    const float *c2 = AXIS == 0 ? cez : cex;
    const int comp1 = AXIS == 0 ? 1 : 2, comp2 = AXIS == 0 ? 2 : 0;

    {
        const Q4 hp = ld4(g1 + id + s), h0 = ld4(g1 + id), ps = ld4(psi1 + pid);
        Q4 e = ld4(e1 + id);
        const Q4 c = HAS_C ? ld4(c1 + id) : splat4(0.5f);
        Q4 np = ps;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            if (valid_e(comp1, i, j, k0 + l, nx, ny, nz)) {
                const float tt = b * ps.v[l] + a * (hp.v[l] - h0.v[l]);
                np.v[l] = tt;
                e.v[l] = e.v[l] - c.v[l] * tt;
            }
        }
        st4(psi1 + pid, np);
        st4(e1 + id, e);
    }
    {
        const Q4 hp = ld4(g2 + id + s), h0 = ld4(g2 + id), ps = ld4(psi2 + pid);
        Q4 e = ld4(e2 + id);
        const Q4 c = HAS_C ? ld4(c2 + id) : splat4(0.5f);
        Q4 np = ps;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            if (valid_e(comp2, i, j, k0 + l, nx, ny, nz)) {
                const float tt = b * ps.v[l] + a * (hp.v[l] - h0.v[l]);
                np.v[l] = tt;
                e.v[l] = e.v[l] + c.v[l] * tt;
            }
        }
        st4(psi2 + pid, np);
        st4(e2 + id, e);
    }
}

// H half, x or y axis.
template <int AXIS, bool HAS_C>
__global__ __launch_bounds__(kBlock) void k_cpml_h_xy(
    int nx, int ny, int nzp, int nz, int npml, unsigned per_side, unsigned Q,
    const float *__restrict__ ex, const float *__restrict__ ey, const float *__restrict__ ez,
    float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
    const float *__restrict__ chx, const float *__restrict__ chy, const float *__restrict__ chz,
    FaceSet fs, const float *__restrict__ tb, const float *__restrict__ ta)
{
    unsigned f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= 2 * per_side) return;
    const int side = f >= per_side;
    if (side) f -= per_side;
    float *psi1 = fs.psi1[side], *psi2 = fs.psi2[side];
    if (!psi1) return;
    const unsigned rowq = f / Q;
    const int k0 = (int)(f - rowq * Q) * 4;
    int i, j, p;
    const int n_e = AXIS == 0 ? nx : ny;
    if (AXIS == 0) {
        p = (int)(rowq / (unsigned)ny);
        j = (int)rowq - p * ny;
        i = side ? n_e - npml + p : p;
    } else {
        i = (int)(rowq / (unsigned)npml);
        p = (int)rowq - i * npml;
        j = side ? n_e - npml + p : p;
    }
    const int t = side ? npml + 2 + p : p + 1;
    const float b = tb[t], a = ta[t];
    const size_t id = ((size_t)i * ny + j) * nzp + k0;
    const size_t pid = (size_t)rowq * nzp + k0;
    const size_t s = AXIS == 0 ? (size_t)ny * nzp : (size_t)nzp;
    // x: psi1 -> hy from ez, psi2 -> hz from ey ; y: psi1 -> hz from ex, psi2 -> hx from ez
    float *h1 = AXIS == 0 ? hy : hz;
    float *h2 = AXIS == 0 ? hz : hx;
    const float *g1 = AXIS == 0 ? ez : ex;
    const float *g2 = AXIS == 0 ? ey : ez;
    const float *c1 = AXIS == 0 ? chy : chz;
    const float *c2 = AXIS == 0 ? chz : chx;
    const int comp1 = AXIS == 0 ? 1 : 2, comp2 = AXIS == 0 ? 2 : 0;
    // Front planes p = 0 are the wall: valid_h() is false there, so id - s is never read.
    const bool inner = (AXIS == 0 ? i : j) > 0;

    if (inner) {
        const Q4 g0 = ld4(g1 + id), gm = ld4(g1 + id - s), ps = ld4(psi1 + pid);
        Q4 h = ld4(h1 + id);
        const Q4 c = HAS_C ? ld4(c1 + id) : splat4(0.5f);
        Q4 np = ps;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            if (valid_h(comp1, i, j, k0 + l, nz)) {
                const float tt = b * ps.v[l] + a * (g0.v[l] - gm.v[l]);
                np.v[l] = tt;
                h.v[l] = h.v[l] + c.v[l] * tt;
            }
        }
        st4(psi1 + pid, np);
        st4(h1 + id, h);

        const Q4 q0 = ld4(g2 + id), qm = ld4(g2 + id - s), ps2 = ld4(psi2 + pid);
        Q4 h_ = ld4(h2 + id);
        const Q4 c_ = HAS_C ? ld4(c2 + id) : splat4(0.5f);
        Q4 np2 = ps2;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            if (valid_h(comp2, i, j, k0 + l, nz)) {
                const float tt = b * ps2.v[l] + a * (q0.v[l] - qm.v[l]);
                np2.v[l] = tt;
                h_.v[l] = h_.v[l] - c_.v[l] * tt;
            }
        }
        st4(psi2 + pid, np2);
        st4(h2 + id, h_);
    }
}

// z axis: thread per slab element (psi layout (nx, ny, npml)).
template <bool HAS_C>
__global__ __launch_bounds__(kBlock) void k_cpml_e_z(
    int nx, int ny, int nzp, int nz, int npml, unsigned per_side,
    float *__restrict__ ex, float *__restrict__ ey,
    const float *__restrict__ hx, const float *__restrict__ hy,
    const float *__restrict__ cex, const float *__restrict__ cey,
    FaceSet fs, const float *__restrict__ tb, const float *__restrict__ ta)
{
    unsigned f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= 2 * per_side) return;
    const int side = f >= per_side;
    if (side) f -= per_side;
    float *psi1 = fs.psi1[side], *psi2 = fs.psi2[side];
    if (!psi1) return;
    const unsigned row = f / (unsigned)npml;
    const int p = (int)(f - row * npml);
    const int i = (int)(row / (unsigned)ny);
    const int j = (int)row - i * ny;
    const int k = side ? nz - 1 - npml + p : p;
    const int t = side ? npml + 1 + p : p + 1;
    const float b = tb[t], a = ta[t];
    const size_t id = (size_t)row * nzp + k;
    // psi1 -> ex from hy, psi2 -> ey from hx
    if (valid_e(0, i, j, k, nx, ny, nz)) {
        const float tt = b * psi1[f] + a * (hy[id + 1] - hy[id]);
        psi1[f] = tt;
        ex[id] = ex[id] - (HAS_C ? cex[id] : 0.5f) * tt;
    }
    if (valid_e(1, i, j, k, nx, ny, nz)) {
        const float tt = b * psi2[f] + a * (hx[id + 1] - hx[id]);
        psi2[f] = tt;
        ey[id] = ey[id] + (HAS_C ? cey[id] : 0.5f) * tt;
    }
}

template <bool HAS_C>
__global__ __launch_bounds__(kBlock) void k_cpml_h_z(
    int nx, int ny, int nzp, int nz, int npml, unsigned per_side,
    const float *__restrict__ ex, const float *__restrict__ ey,
    float *__restrict__ hx, float *__restrict__ hy,
    const float *__restrict__ chx, const float *__restrict__ chy,
    FaceSet fs, const float *__restrict__ tb, const float *__restrict__ ta)
{
    (void)nx;
    unsigned f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= 2 * per_side) return;
    const int side = f >= per_side;
    if (side) f -= per_side;
    float *psi1 = fs.psi1[side], *psi2 = fs.psi2[side];
    if (!psi1) return;
    const unsigned row = f / (unsigned)npml;
    const int p = (int)(f - row * npml);
    const int i = (int)(row / (unsigned)ny);
    const int j = (int)row - i * ny;
    const int k = side ? nz - npml + p : p;
    const int t = side ? npml + 2 + p : p + 1;
    const float b = tb[t], a = ta[t];
    const size_t id = (size_t)row * nzp + k;
    // psi1 -> hx from ey, psi2 -> hy from ex  (k > 0 on every valid cell)
    if (valid_h(0, i, j, k, nz)) {
        const float tt = b * psi1[f] + a * (ey[id] - ey[id - 1]);
        psi1[f] = tt;
        hx[id] = hx[id] + (HAS_C ? chx[id] : 0.5f) * tt;
    }
    if (valid_h(1, i, j, k, nz)) {
        const float tt = b * psi2[f] + a * (ex[id] - ex[id - 1]);
        psi2[f] = tt;
        hy[id] = hy[id] - (HAS_C ? chy[id] : 0.5f) * tt;
    }
}

static int check_cpml(int nx, int ny, int nzp, int pad, int npml, const float *b, const float *a)
{
    if (int rc = check_grid(nx, ny, nzp, pad)) return rc;
    if (npml < 1 || npml > MIFDTD_MAX_NPML) return fail(MIFDTD_EINVAL, "npml %d outside [1,%d]", npml, MIFDTD_MAX_NPML);
    const int nz = nzp - pad;
    if (nx < 2 * npml + 2 || ny < 2 * npml + 2 || nz < 2 * npml + 2)
        return fail(MIFDTD_EINVAL, "grid (%d,%d,%d) too small for npml %d (need >= 2*npml+2 per axis)", nx, ny, nz, npml);
    if (!b || !a) return fail(MIFDTD_ENULL, "CPML coefficient table is NULL");
    return 0;
}

// One axis of the E half, both sides.
static int cpml_e_axis(int axis, int nx, int ny, int nzp, int pad, int npml, FaceSet fs,
                       float *ex, float *ey, float *ez, const float *hx, const float *hy, const float *hz,
                       bool hc, const float *cex, const float *cey, const float *cez,
                       const float *b, const float *a, hipStream_t st)
{
    if (!fs.psi1[0] && !fs.psi1[1]) return 0;
    for (int s = 0; s < 2; s++) {
        if ((fs.psi1[s] == nullptr) != (fs.psi2[s] == nullptr))
            return fail(MIFDTD_ENULL, "psi1/psi2 of one face must be both set or both NULL");
        if (axis < 2 && fs.psi1[s] && (!aligned16(fs.psi1[s]) || !aligned16(fs.psi2[s])))
            return fail(MIFDTD_EALIGN, "psi pointer is not 16-byte aligned");
    }
    const int nz = nzp - pad;
    const unsigned Q = nzp / 4;
    if (axis == 0) {
        const unsigned per = (unsigned)npml * ny * Q;
        const unsigned g = (2 * per + kBlock - 1) / kBlock;
        if (hc) k_cpml_e_xy<0, true><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, cex, cey, cez, fs, b, a);
        else k_cpml_e_xy<0, false><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, cex, cey, cez, fs, b, a);
    } else if (axis == 1) {
        const unsigned per = (unsigned)nx * npml * Q;
        const unsigned g = (2 * per + kBlock - 1) / kBlock;
        if (hc) k_cpml_e_xy<1, true><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, cex, cey, cez, fs, b, a);
        else k_cpml_e_xy<1, false><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, cex, cey, cez, fs, b, a);
    } else {
        const unsigned per = (unsigned)nx * ny * npml;
        const unsigned g = (2 * per + kBlock - 1) / kBlock;
        if (hc) k_cpml_e_z<true><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, ex, ey, hx, hy, cex, cey, fs, b, a);
        else k_cpml_e_z<false><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, ex, ey, hx, hy, cex, cey, fs, b, a);
    }
    MIFDTD_CHECK_LAUNCH("cpml E launch");
    return 0;
}

static int cpml_h_axis(int axis, int nx, int ny, int nzp, int pad, int npml, FaceSet fs,
                       const float *ex, const float *ey, const float *ez, float *hx, float *hy, float *hz,
                       bool hc, const float *chx, const float *chy, const float *chz,
                       const float *b, const float *a, hipStream_t st)
{
    if (!fs.psi1[0] && !fs.psi1[1]) return 0;
    for (int s = 0; s < 2; s++) {
        if ((fs.psi1[s] == nullptr) != (fs.psi2[s] == nullptr))
            return fail(MIFDTD_ENULL, "psi1/psi2 of one face must be both set or both NULL");
        if (axis < 2 && fs.psi1[s] && (!aligned16(fs.psi1[s]) || !aligned16(fs.psi2[s])))
            return fail(MIFDTD_EALIGN, "psi pointer is not 16-byte aligned");
    }
    const int nz = nzp - pad;
    const unsigned Q = nzp / 4;
    if (axis == 0) {
        const unsigned per = (unsigned)npml * ny * Q;
        const unsigned g = (2 * per + kBlock - 1) / kBlock;
        if (hc) k_cpml_h_xy<0, true><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, chx, chy, chz, fs, b, a);
        else k_cpml_h_xy<0, false><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, chx, chy, chz, fs, b, a);
    } else if (axis == 1) {
        const unsigned per = (unsigned)nx * npml * Q;
        const unsigned g = (2 * per + kBlock - 1) / kBlock;
        if (hc) k_cpml_h_xy<1, true><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, chx, chy, chz, fs, b, a);
        else k_cpml_h_xy<1, false><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, Q, ex, ey, ez, hx, hy, hz, chx, chy, chz, fs, b, a);
    } else {
        const unsigned per = (unsigned)nx * ny * npml;
        const unsigned g = (2 * per + kBlock - 1) / kBlock;
        if (hc) k_cpml_h_z<true><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, ex, ey, hx, hy, chx, chy, fs, b, a);
        else k_cpml_h_z<false><<<g, kBlock, 0, st>>>(nx, ny, nzp, nz, npml, per, ex, ey, hx, hy, chx, chy, fs, b, a);
    }
    MIFDTD_CHECK_LAUNCH("cpml H launch");
    return 0;
}

// ------------------------------------------------------------------ source / gather
__global__ __launch_bounds__(kBlock) void k_source(
    float *__restrict__ f, int ny, int nzp, int i0, int j0, int k0, int li, int lj, int lk,
    float tval, const float *__restrict__ sv, float svs, int overwrite)
{
    const unsigned n = blockIdx.x * kBlock + threadIdx.x;
    const unsigned tot = (unsigned)li * lj * lk;
    if (n >= tot) return;
    const int dk = (int)(n % (unsigned)lk);
    const unsigned r = n / (unsigned)lk;
    const int dj = (int)(r % (unsigned)lj);
    const int di = (int)(r / (unsigned)lj);
    const size_t id = ((size_t)(i0 + di) * ny + (j0 + dj)) * nzp + (k0 + dk);
    const float v = tval * (sv ? sv[n] : svs);
    f[id] = overwrite ? v : f[id] + v;
}

__global__ __launch_bounds__(kBlock) void k_gather(const float *__restrict__ f, const long long *__restrict__ off,
                                                   int count, float *__restrict__ out)
{
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n < count) out[n] = f[off[n]];
}

// ------------------------------------------------------------- periodic boundary
// Copy plane `src` of two components onto plane `dst` across axis `axis`: element n of
// the plane is (j, k) of plane i for x (ny * nzp elements, contiguous), (i, k) of row j
// for y (nx rows of nzp), (i, j) at z = plane for z (nx * ny elements, stride nzp).
__global__ __launch_bounds__(kBlock) void k_pbc(float *__restrict__ f1, float *__restrict__ f2, int nx, int ny,
                                                int nzp, int axis, int src, int dst, unsigned count)
{
    const unsigned n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= count) return;
    size_t s, d;
    if (axis == 0) {
        const size_t plane = (size_t)ny * nzp;
        s = (size_t)src * plane + n;
        d = (size_t)dst * plane + n;
    } else if (axis == 1) {
        const size_t i = n / (unsigned)nzp, k = n % (unsigned)nzp;
        s = (i * ny + src) * nzp + k;
        d = (i * ny + dst) * nzp + k;
    } else {
        const size_t row = n;  // i * ny + j
        s = row * nzp + src;
        d = row * nzp + dst;
    }
    f1[d] = f1[s];
    f2[d] = f2[s];
}

// ------------------------------------------------------------------ Drude dispersion
// Auxiliary-current update of the region [i0, i0+li) x [j0, j0+lj) x [k0, k0+lk): per
// component e' = e - ce J, J' = cja J + cjb e' (J, cja, cjb region-sized, C order; ce
// the grid's E coefficient, 0.5 where the pointer is null).
__global__ __launch_bounds__(kBlock) void k_drude_e(float *ex, float *ey, float *ez, const float *cex, const float *cey,
                                                    const float *cez, float *jx, float *jy, float *jz,
                                                    const float *__restrict__ ax, const float *__restrict__ ay,
                                                    const float *__restrict__ az, const float *__restrict__ bx,
                                                    const float *__restrict__ by, const float *__restrict__ bz, int ny,
                                                    int nzp, int i0, int j0, int k0, int lj, int lk, unsigned count)
{
    const unsigned n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= count) return;
    const int dk = (int)(n % (unsigned)lk);
    const unsigned r = n / (unsigned)lk;
    const int dj = (int)(r % (unsigned)lj), di = (int)(r / (unsigned)lj);
    const size_t id = ((size_t)(i0 + di) * ny + (j0 + dj)) * nzp + (k0 + dk);
    float *const e[3] = {ex, ey, ez};
    const float *const ce[3] = {cex, cey, cez};
    float *const J[3] = {jx, jy, jz};
    const float *const A[3] = {ax, ay, az};
    const float *const B[3] = {bx, by, bz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float f = J[c][n];
        const float c0 = ce[c] ? ce[c][id] : 0.5f;
        const float et = e[c][id] - c0 * f;
        e[c][id] = et;
        J[c][n] = A[c][n] * f + B[c][n] * et;
    }
}

// ------------------------------------------------------------------ TFSF plane wave
// The 1-D incident line (one work-group): E half: E[1:n-1] -= 0.5 (H[2:] -
// H[1:n-1]), the two-step-delay ends (abc[0..1] front, abc[2..3] back), E[3] = src;
// H half: H[1:] -= 0.5 (E[1:] - E[:-1]).  Each half reads only the other array.
__global__ __launch_bounds__(1024) void k_tfsf_line(float *__restrict__ e, float *__restrict__ h, float *__restrict__ abc,
                                                    int n, int e_half, float src)
{
    if (e_half) {
        for (int m = 1 + threadIdx.x; m < n - 1; m += blockDim.x) e[m] = e[m] - 0.5f * (h[m + 1] - h[m]);
        __syncthreads();
        if (threadIdx.x == 0) {
            e[0] = abc[0];
            abc[0] = abc[1];
            abc[1] = e[1];
            e[n - 1] = abc[2];
            abc[2] = abc[3];
            abc[3] = e[n - 2];
            e[3] = src;
        }
    } else {
        for (int m = 1 + threadIdx.x; m < n; m += blockDim.x) h[m] = h[m] - 0.5f * (e[m] - e[m - 1]);
    }
}

// Face correction of the patches in P: the box [i0..i1, j0..j1, k0..k1] of component
// comp takes f += 0.5 * (coef * line[base + (coordinate along axis - box start)]).
// Patches of one component meet on the box edges, where a cell takes two corrections in
// table order, so the patches of one launch are of distinct components.
__global__ __launch_bounds__(kBlock) void k_tfsf_faces(mifdtd_tfsf_patches P, float *ex, float *ey, float *ez,
                                                       float *hx, float *hy, float *hz,
                                                       const float *__restrict__ line, int ny, int nzp)
{
    const unsigned t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= P.start[P.n]) return;
    int p = 0;
    while (p + 1 < P.n && t >= P.start[p + 1]) ++p;
    const int *b = P.box[p];
    const unsigned r = t - P.start[p];
    const unsigned lk = (unsigned)(b[5] - b[4] + 1), lj = (unsigned)(b[3] - b[2] + 1);
    const int k = b[4] + (int)(r % lk);
    const int j = b[2] + (int)((r / lk) % lj);
    const int i = b[0] + (int)(r / lk / lj);
    const int along = P.axis == 0 ? i - b[0] : P.axis == 1 ? j - b[2] : k - b[4];
    const float v = 0.5f * (P.coef[p] * line[P.base[p] + along]);
    float *f = P.comp[p] == 0 ? ex : P.comp[p] == 1 ? ey : P.comp[p] == 2 ? ez : P.comp[p] == 3 ? hx : P.comp[p] == 4 ? hy : hz;
    const size_t id = ((size_t)i * ny + j) * nzp + k;
    f[id] = f[id] + v;
}

}  // namespace mifdtd

// =========================================================================== C ABI
using namespace mifdtd;

extern "C" {

const char *mifdtd_last_error(void) { return g_err.c_str(); }

int mifdtd_version(void) { return 1; }

int mifdtd_update_e(int nx, int ny, int nz_pitch, int pad, float *ex, float *ey, float *ez,
                    const float *hx, const float *hy, const float *hz,
                    const float *cex, const float *cey, const float *cez, mifdtd_stream_t stream)
{
    return launch_core(true, nx, ny, nz_pitch, pad, ex, ey, ez, const_cast<float *>(hx), const_cast<float *>(hy),
                       const_cast<float *>(hz), cex, cey, cez, to_stream(stream));
}

int mifdtd_update_e_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax, float *ex, float *ey,
                          float *ez, const float *hx, const float *hy, const float *hz, const float *cex,
                          const float *cey, const float *cez, mifdtd_stream_t stream)
{
    if (nmax < 0) return fail(MIFDTD_EINVAL, "nmax %lld < 0", nmax);
    return launch_core(true, nx, ny, nz_pitch, pad, ex, ey, ez, const_cast<float *>(hx), const_cast<float *>(hy),
                       const_cast<float *>(hz), cex, cey, cez, to_stream(stream), idx0, nmax);
}

int mifdtd_update_h_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax, const float *ex,
                          const float *ey, const float *ez, float *hx, float *hy, float *hz, const float *chx,
                          const float *chy, const float *chz, mifdtd_stream_t stream)
{
    if (nmax < 0) return fail(MIFDTD_EINVAL, "nmax %lld < 0", nmax);
    return launch_core(false, nx, ny, nz_pitch, pad, const_cast<float *>(ex), const_cast<float *>(ey),
                       const_cast<float *>(ez), hx, hy, hz, chx, chy, chz, to_stream(stream), idx0, nmax);
}

int mifdtd_update_e_lut(int nx, int ny, int nz_pitch, int pad, float *ex, float *ey, float *ez, const float *hx,
                        const float *hy, const float *hz, const unsigned char *mat_x, const unsigned char *mat_y,
                        const unsigned char *mat_z, const float *lut, mifdtd_stream_t stream)
{
    if (int rc = check_grid(nx, ny, nz_pitch, pad)) return rc;
    const float *ptrs[6] = {ex, ey, ez, hx, hy, hz};
    for (const float *p : ptrs) {
        if (!p) return fail(MIFDTD_ENULL, "field pointer is NULL");
        if (!aligned16(p)) return fail(MIFDTD_EALIGN, "field pointer %p is not 16-byte aligned", (const void *)p);
    }
    if (!mat_x || !mat_y || !mat_z || !lut) return fail(MIFDTD_ENULL, "material index or table is NULL");
    const unsigned char *mats[3] = {mat_x, mat_y, mat_z};
    for (const unsigned char *m : mats)
        if (reinterpret_cast<uintptr_t>(m) & 3u) return fail(MIFDTD_EALIGN, "material index array not 4-byte aligned");
    const unsigned Q = (unsigned)(nz_pitch / 4);
    const unsigned nquads = (unsigned)nx * ny * Q;
    const unsigned grid = (nquads + kBlock - 1) / kBlock;
    k_update_e_lut<<<grid, kBlock, 0, to_stream(stream)>>>(nx, ny, nz_pitch, nz_pitch - pad, nquads, Q, ex, ey, ez, hx,
                                                           hy, hz, mat_x, mat_y, mat_z, lut);
    MIFDTD_CHECK_LAUNCH("update_e_lut launch");
    return 0;
}

int mifdtd_update_h(int nx, int ny, int nz_pitch, int pad, const float *ex, const float *ey, const float *ez,
                    float *hx, float *hy, float *hz,
                    const float *chx, const float *chy, const float *chz, mifdtd_stream_t stream)
{
    return launch_core(false, nx, ny, nz_pitch, pad, const_cast<float *>(ex), const_cast<float *>(ey),
                       const_cast<float *>(ez), hx, hy, hz, chx, chy, chz, to_stream(stream));
}

int mifdtd_cpml_e(int nx, int ny, int nzp, int pad, const mifdtd_cpml_half *cp,
                  float *ex, float *ey, float *ez, const float *hx, const float *hy, const float *hz,
                  const float *cex, const float *cey, const float *cez, mifdtd_stream_t stream)
{
    if (!cp) return fail(MIFDTD_ENULL, "cpml descriptor is NULL");
    if (int rc = check_cpml(nx, ny, nzp, pad, cp->npml, cp->b, cp->a)) return rc;
    if ((cex || cey || cez) && !(cex && cey && cez)) return fail(MIFDTD_ENULL, "coefficient arrays must be all set or all NULL");
    for (int ax = 0; ax < 3; ax++) {
        FaceSet fs{{cp->psi1[2 * ax], cp->psi1[2 * ax + 1]}, {cp->psi2[2 * ax], cp->psi2[2 * ax + 1]}};
        if (int rc = cpml_e_axis(ax, nx, ny, nzp, pad, cp->npml, fs, ex, ey, ez, hx, hy, hz, cex != nullptr, cex, cey, cez,
                                 cp->b, cp->a, to_stream(stream)))
            return rc;
    }
    return 0;
}

int mifdtd_cpml_h(int nx, int ny, int nzp, int pad, const mifdtd_cpml_half *cp,
                  const float *ex, const float *ey, const float *ez, float *hx, float *hy, float *hz,
                  const float *chx, const float *chy, const float *chz, mifdtd_stream_t stream)
{
    if (!cp) return fail(MIFDTD_ENULL, "cpml descriptor is NULL");
    if (int rc = check_cpml(nx, ny, nzp, pad, cp->npml, cp->b, cp->a)) return rc;
    if ((chx || chy || chz) && !(chx && chy && chz)) return fail(MIFDTD_ENULL, "coefficient arrays must be all set or all NULL");
    for (int ax = 0; ax < 3; ax++) {
        FaceSet fs{{cp->psi1[2 * ax], cp->psi1[2 * ax + 1]}, {cp->psi2[2 * ax], cp->psi2[2 * ax + 1]}};
        if (int rc = cpml_h_axis(ax, nx, ny, nzp, pad, cp->npml, fs, ex, ey, ez, hx, hy, hz, chx != nullptr, chx, chy, chz,
                                 cp->b, cp->a, to_stream(stream)))
            return rc;
    }
    return 0;
}

// Per-face forms (kuFDTD's argument order, Ncore dropped; pad and stream explicit).
static int face_e(int axis, int nx, int ny, int nz, int pad, int npml, int backward, float *f1, float *f2,
                  const float *g1, const float *g2, const float *c1, const float *c2, float *psi1, float *psi2,
                  const float *b, const float *a, mifdtd_stream_t stream)
{
    if (int rc = check_cpml(nx, ny, nz, pad, npml, b, a)) return rc;
    if (backward != 0 && backward != 1) return fail(MIFDTD_EINVAL, "backward must be 0 or 1");
    if (!psi1 || !psi2) return fail(MIFDTD_ENULL, "psi is NULL");
    FaceSet fs{{nullptr, nullptr}, {nullptr, nullptr}};
    fs.psi1[backward] = psi1;
    fs.psi2[backward] = psi2;
    float *ex = nullptr, *ey = nullptr, *ez = nullptr;
    const float *cex = nullptr, *cey = nullptr, *cez = nullptr;
    const float *hx = nullptr, *hy = nullptr, *hz = nullptr;
    if (axis == 0) { ey = f1; ez = f2; hy = g1; hz = g2; cey = c1; cez = c2; }
    if (axis == 1) { ex = f1; ez = f2; hx = g1; hz = g2; cex = c1; cez = c2; }
    if (axis == 2) { ex = f1; ey = f2; hx = g1; hy = g2; cex = c1; cey = c2; }
    if ((c1 == nullptr) != (c2 == nullptr)) return fail(MIFDTD_ENULL, "CE arrays must be both set or both NULL");
    return cpml_e_axis(axis, nx, ny, nz, pad, npml, fs, ex, ey, ez, hx, hy, hz, c1 != nullptr, cex, cey, cez, b, a,
                       to_stream(stream));
}

static int face_h(int axis, int nx, int ny, int nz, int pad, int npml, int backward, const float *f1,
                  const float *f2, float *g1, float *g2, float *psi1, float *psi2, const float *b, const float *a,
                  mifdtd_stream_t stream)
{
    if (int rc = check_cpml(nx, ny, nz, pad, npml, b, a)) return rc;
    if (backward != 0 && backward != 1) return fail(MIFDTD_EINVAL, "backward must be 0 or 1");
    if (!psi1 || !psi2) return fail(MIFDTD_ENULL, "psi is NULL");
    FaceSet fs{{nullptr, nullptr}, {nullptr, nullptr}};
    fs.psi1[backward] = psi1;
    fs.psi2[backward] = psi2;
    const float *ex = nullptr, *ey = nullptr, *ez = nullptr;
    float *hx = nullptr, *hy = nullptr, *hz = nullptr;
    if (axis == 0) { ey = f1; ez = f2; hy = g1; hz = g2; }
    if (axis == 1) { ex = f1; ez = f2; hx = g1; hz = g2; }
    if (axis == 2) { ex = f1; ey = f2; hx = g1; hy = g2; }
    return cpml_h_axis(axis, nx, ny, nz, pad, npml, fs, ex, ey, ez, hx, hy, hz, false, nullptr, nullptr, nullptr, b, a,
                       to_stream(stream));
}

int mifdtd_cpml_face_x_e(int nx, int ny, int nzp, int pad, int npml, int backward, float *ey, float *ez,
                         const float *hy, const float *hz, const float *cey, const float *cez, float *psi1,
                         float *psi2, const float *rcmb, const float *rcma, mifdtd_stream_t stream)
{
    return face_e(0, nx, ny, nzp, pad, npml, backward, ey, ez, hy, hz, cey, cez, psi1, psi2, rcmb, rcma, stream);
}
int mifdtd_cpml_face_x_h(int nx, int ny, int nzp, int pad, int npml, int backward, const float *ey, const float *ez,
                         float *hy, float *hz, float *psi1, float *psi2, const float *rcmb, const float *rcma,
                         mifdtd_stream_t stream)
{
    return face_h(0, nx, ny, nzp, pad, npml, backward, ey, ez, hy, hz, psi1, psi2, rcmb, rcma, stream);
}
int mifdtd_cpml_face_y_e(int nx, int ny, int nzp, int pad, int npml, int backward, float *ex, float *ez,
                         const float *hx, const float *hz, const float *cex, const float *cez, float *psi1,
                         float *psi2, const float *rcmb, const float *rcma, mifdtd_stream_t stream)
{
    return face_e(1, nx, ny, nzp, pad, npml, backward, ex, ez, hx, hz, cex, cez, psi1, psi2, rcmb, rcma, stream);
}
int mifdtd_cpml_face_y_h(int nx, int ny, int nzp, int pad, int npml, int backward, const float *ex, const float *ez,
                         float *hx, float *hz, float *psi1, float *psi2, const float *rcmb, const float *rcma,
                         mifdtd_stream_t stream)
{
    return face_h(1, nx, ny, nzp, pad, npml, backward, ex, ez, hx, hz, psi1, psi2, rcmb, rcma, stream);
}
int mifdtd_cpml_face_z_e(int nx, int ny, int nzp, int pad, int npml, int backward, float *ex, float *ey,
                         const float *hx, const float *hy, const float *cex, const float *cey, float *psi1,
                         float *psi2, const float *rcmb, const float *rcma, mifdtd_stream_t stream)
{
    return face_e(2, nx, ny, nzp, pad, npml, backward, ex, ey, hx, hy, cex, cey, psi1, psi2, rcmb, rcma, stream);
}
int mifdtd_cpml_face_z_h(int nx, int ny, int nzp, int pad, int npml, int backward, const float *ex, const float *ey,
                         float *hx, float *hy, float *psi1, float *psi2, const float *rcmb, const float *rcma,
                         mifdtd_stream_t stream)
{
    return face_h(2, nx, ny, nzp, pad, npml, backward, ex, ey, hx, hy, psi1, psi2, rcmb, rcma, stream);
}

int mifdtd_source(float *f, int nx, int ny, int nzp, int i0, int i1, int j0, int j1, int k0, int k1,
                  float tval, const float *sv, float svs, int overwrite, mifdtd_stream_t stream)
{
    if (!f) return fail(MIFDTD_ENULL, "field pointer is NULL");
    if (i0 < 0 || j0 < 0 || k0 < 0 || i1 >= nx || j1 >= ny || k1 >= nzp || i0 > i1 || j0 > j1 || k0 > k1)
        return fail(MIFDTD_EINVAL, "source box [%d:%d,%d:%d,%d:%d] outside (%d,%d,%d)", i0, i1, j0, j1, k0, k1, nx, ny, nzp);
    const int li = i1 - i0 + 1, lj = j1 - j0 + 1, lk = k1 - k0 + 1;
    const unsigned long long tot = (unsigned long long)li * lj * lk;
    if (tot >= (1ull << 31)) return fail(MIFDTD_ERANGE, "source box too large");
    const unsigned g = (unsigned)((tot + kBlock - 1) / kBlock);
    k_source<<<g, kBlock, 0, to_stream(stream)>>>(f, ny, nzp, i0, j0, k0, li, lj, lk, tval, sv, svs, overwrite);
    MIFDTD_CHECK_LAUNCH("source launch");
    return 0;
}

int mifdtd_pbc(float *f1, float *f2, int nx, int ny, int nz_pitch, int pad, int axis, int e_half,
               mifdtd_stream_t stream)
{
    if (!f1 || !f2) return fail(MIFDTD_ENULL, "field pointer is NULL");
    if (int rc = check_grid(nx, ny, nz_pitch, pad)) return rc;
    if (axis < 0 || axis > 2) return fail(MIFDTD_EINVAL, "axis %d outside [0,2]", axis);
    if (f1 == f2) return fail(MIFDTD_EINVAL, "the two components of a periodic copy must differ");
    const int n = axis == 0 ? nx : axis == 1 ? ny : nz_pitch - pad;
    if (n < 2) return fail(MIFDTD_EINVAL, "a periodic axis needs at least 2 cells");
    const int src = e_half ? 0 : n - 1, dst = e_half ? n - 1 : 0;
    const unsigned long long count = axis == 0 ? (unsigned long long)ny * nz_pitch
                                   : axis == 1 ? (unsigned long long)nx * nz_pitch
                                               : (unsigned long long)nx * ny;
    if (count >= (1ull << 32) - kBlock) return fail(MIFDTD_ERANGE, "periodic plane too large");
    const unsigned g = (unsigned)((count + kBlock - 1) / kBlock);
    k_pbc<<<g, kBlock, 0, to_stream(stream)>>>(f1, f2, nx, ny, nz_pitch, axis, src, dst, (unsigned)count);
    MIFDTD_CHECK_LAUNCH("periodic copy launch");
    return 0;
}

int mifdtd_drude_e(float *const e[3], const float *const ce[3], float *const j[3], const float *const cja[3],
                   const float *const cjb[3], int nx, int ny, int nz_pitch, int i0, int j0, int k0, int li, int lj,
                   int lk, mifdtd_stream_t stream)
{
    if (!e || !j || !cja || !cjb) return fail(MIFDTD_ENULL, "Drude argument is NULL");
    for (int c = 0; c < 3; c++)
        if (!e[c] || !j[c] || !cja[c] || !cjb[c]) return fail(MIFDTD_ENULL, "Drude array is NULL");
    if (li < 1 || lj < 1 || lk < 1 || i0 < 0 || j0 < 0 || k0 < 0 || i0 + li > nx || j0 + lj > ny || k0 + lk > nz_pitch)
        return fail(MIFDTD_EINVAL, "Drude region outside the grid");
    const unsigned long long tot = (unsigned long long)li * lj * lk;
    if (tot >= (1ull << 32) - kBlock) return fail(MIFDTD_ERANGE, "Drude region too large");
    const float *c0 = ce ? ce[0] : nullptr, *c1 = ce ? ce[1] : nullptr, *c2 = ce ? ce[2] : nullptr;
    const unsigned g = (unsigned)((tot + kBlock - 1) / kBlock);
    k_drude_e<<<g, kBlock, 0, to_stream(stream)>>>(e[0], e[1], e[2], c0, c1, c2, j[0], j[1], j[2], cja[0], cja[1], cja[2],
                                                   cjb[0], cjb[1], cjb[2], ny, nz_pitch, i0, j0, k0, lj, lk,
                                                   (unsigned)tot);
    MIFDTD_CHECK_LAUNCH("Drude launch");
    return 0;
}

int mifdtd_tfsf_line(float *e, float *h, float *abc, int n, int e_half, float src, mifdtd_stream_t stream)
{
    if (!e || !h || !abc) return fail(MIFDTD_ENULL, "TFSF line pointer is NULL");
    if (n < 6) return fail(MIFDTD_EINVAL, "TFSF line of %d points (at least 6)", n);
    k_tfsf_line<<<1, 1024, 0, to_stream(stream)>>>(e, h, abc, n, e_half, src);
    MIFDTD_CHECK_LAUNCH("TFSF line launch");
    return 0;
}

int mifdtd_tfsf_faces(const mifdtd_tfsf_patches *p, float *const f[6], const float *line, int line_n, int nx, int ny,
                      int nz_pitch, mifdtd_stream_t stream)
{
    if (!p || !f || !line) return fail(MIFDTD_ENULL, "TFSF argument is NULL");
    for (int q = 0; q < 6; q++)
        if (!f[q]) return fail(MIFDTD_ENULL, "field pointer is NULL");
    if (p->n < 0 || p->n > MIFDTD_TFSF_MAX_PATCH) return fail(MIFDTD_EINVAL, "patch count %d", p->n);
    if (p->axis < 0 || p->axis > 2) return fail(MIFDTD_EINVAL, "axis %d", p->axis);
    mifdtd_tfsf_patches P = *p;
    unsigned long long tot = 0;
    for (int q = 0; q < P.n; q++) {
        const int *b = P.box[q];
        if (P.comp[q] < 0 || P.comp[q] > 5) return fail(MIFDTD_EINVAL, "patch %d component %d", q, P.comp[q]);
        if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] >= nx || b[3] >= ny || b[5] >= nz_pitch || b[0] > b[1] ||
            b[2] > b[3] || b[4] > b[5])
            return fail(MIFDTD_EINVAL, "patch %d box outside the grid", q);
        const int len = P.axis == 0 ? b[1] - b[0] : P.axis == 1 ? b[3] - b[2] : b[5] - b[4];
        if (P.base[q] < 0 || P.base[q] + len >= line_n) return fail(MIFDTD_EINVAL, "patch %d reads past the line", q);
        P.start[q] = (unsigned)tot;
        tot += (unsigned long long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
        if (tot >= (1ull << 31)) return fail(MIFDTD_ERANGE, "TFSF faces too large");
    }
    P.start[P.n] = (unsigned)tot;
    // Launch r holds the r-th patch (in table order) of every component: a cell that
    // two patches correct (a box edge) takes them in table order, and the patches of
    // one launch write different arrays.
    int rank[MIFDTD_TFSF_MAX_PATCH], seen[6] = {0, 0, 0, 0, 0, 0}, nrank = 0;
    for (int q = 0; q < P.n; q++) {
        rank[q] = seen[P.comp[q]]++;
        if (rank[q] + 1 > nrank) nrank = rank[q] + 1;
    }
    for (int r = 0; r < nrank; r++) {
        mifdtd_tfsf_patches Q{};
        Q.axis = P.axis;
        unsigned cells = 0;
        for (int q = 0; q < P.n; q++) {
            if (rank[q] != r) continue;
            const int m = Q.n++;
            Q.comp[m] = P.comp[q];
            for (int c = 0; c < 6; c++) Q.box[m][c] = P.box[q][c];
            Q.base[m] = P.base[q];
            Q.coef[m] = P.coef[q];
            Q.start[m] = cells;
            cells += P.start[q + 1] - P.start[q];  // at least 1 per patch
        }
        Q.start[Q.n] = cells;
        const unsigned g = (cells + kBlock - 1) / kBlock;
        k_tfsf_faces<<<g, kBlock, 0, to_stream(stream)>>>(Q, f[0], f[1], f[2], f[3], f[4], f[5], line, ny, nz_pitch);
        MIFDTD_CHECK_LAUNCH("TFSF faces launch");
    }
    return 0;
}

int mifdtd_gather(const float *f, const long long *offsets, int count, float *out, mifdtd_stream_t stream)
{
    if (count <= 0) return 0;
    if (!f || !offsets || !out) return fail(MIFDTD_ENULL, "gather pointer is NULL");
    const unsigned g = (unsigned)((count + kBlock - 1) / kBlock);
    k_gather<<<g, kBlock, 0, to_stream(stream)>>>(f, offsets, count, out);
    MIFDTD_CHECK_LAUNCH("gather launch");
    return 0;
}

}  // extern "C"
