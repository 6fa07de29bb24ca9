// Double-precision KEMP core: update_e / update_h of the reference's 'double' programs
// (K7/cpu/src/core.c:11,59 substituted with DTYPE=double, INCRE=2 by K7/cpu/core.py:52-54;
// fields of K7/cpu/fields.py:25 with nz_pitch padded to 16 bytes = 2 doubles).
//
// Layout: C-order (nx, ny, nzp) float64, z contiguous, nzp even, pad lanes zero (the CPU
// programs pad nz to 2 doubles, pad 0 or 1; a 64-byte GPU pitch, K7/gpu/fields.py:45-50,
// gives pads up to 7 and the same masks by index).  One thread owns one
// aligned z-pair (16 bytes per array, the reference's SSE2 vector); consecutive lanes own
// consecutive pairs, so a wave streams 1 KiB per array.  Arithmetic order and masks are
// those of the float kernels (value*0 on masked lanes keeps the reference's signed zeros);
// compiled with -ffp-contract=off the results equal the reference's SSE2 code bit for bit.

#include "common.hpp"

namespace mifdtd {

namespace {

constexpr int kBlockD = 256;

struct D2 {
    double v[2];
};

__device__ __forceinline__ D2 ld2(const double *__restrict__ p)
{
    const double2 t = *reinterpret_cast<const double2 *>(p);
    return D2{{t.x, t.y}};
}

__device__ __forceinline__ void st2(double *__restrict__ p, const D2 &q)
{
    *reinterpret_cast<double2 *>(p) = make_double2(q.v[0], q.v[1]);
}

__device__ __forceinline__ D2 ldc(const double *__restrict__ c, size_t id)
{
    return c ? ld2(c + id) : D2{{0.5, 0.5}};
}

__global__ __launch_bounds__(kBlockD) void k_update_e_f64(
    int nx, int ny, int nzp, int nz, unsigned p0, unsigned p1, unsigned P,
    double *__restrict__ ex, double *__restrict__ ey, double *__restrict__ ez,
    const double *__restrict__ hx, const double *__restrict__ hy, const double *__restrict__ hz,
    const double *__restrict__ cex, const double *__restrict__ cey, const double *__restrict__ cez)
{
    const unsigned f = p0 + blockIdx.x * kBlockD + threadIdx.x;  // pairs [p0, p1) of the grid
    if (f >= p1) return;
    const unsigned row = f / P;
    const int k0 = (int)(f - row * P) * 2;
    const int i = (int)(row / (unsigned)ny);
    const int j = (int)row - i * ny;
    const size_t id = (size_t)row * nzp + k0;
    const size_t sx = (size_t)ny * nzp;

    const D2 hx0 = ld2(hx + id), hy0 = ld2(hy + id), hz0 = ld2(hz + id);
    if (j < ny - 1) {
        const D2 e = ld2(ex + id), c = ldc(cex, id), h1 = ld2(hz + id + nzp);
        const double hyn = hy[id + 2];  // next row's first element on the last pair (row j+1 exists)
        D2 r;
#pragma unroll
        for (int l = 0; l < 2; l++) {
            const double h2 = l < 1 ? hy0.v[l + 1] : hyn;
            const double v = e.v[l] + c.v[l] * ((h1.v[l] - hz0.v[l]) - (h2 - hy0.v[l]));
            r.v[l] = (k0 + l < nz - 1) ? v : v * 0.0;
        }
        st2(ex + id, r);
    }
    if (i < nx - 1) {
        const D2 e = ld2(ey + id), c = ldc(cey, id), h2 = ld2(hz + id + sx);
        const double hxn = hx[id + 2];
        D2 r;
#pragma unroll
        for (int l = 0; l < 2; l++) {
            const double h1 = l < 1 ? hx0.v[l + 1] : hxn;
            const double v = e.v[l] + c.v[l] * ((h1 - hx0.v[l]) - (h2.v[l] - hz0.v[l]));
            r.v[l] = (k0 + l < nz - 1) ? v : v * 0.0;
        }
        st2(ey + id, r);
        if (j < ny - 1) {
            const D2 e3 = ld2(ez + id), c3 = ldc(cez, id), g1 = ld2(hy + id + sx), g2 = ld2(hx + id + nzp);
            D2 r3;
#pragma unroll
            for (int l = 0; l < 2; l++) {
                const double v = e3.v[l] + c3.v[l] * ((g1.v[l] - hy0.v[l]) - (g2.v[l] - hx0.v[l]));
                r3.v[l] = (k0 + l < nz) ? v : v * 0.0;
            }
            st2(ez + id, r3);
        }
    }
}

__global__ __launch_bounds__(kBlockD) void k_update_h_f64(
    int ny, int nzp, int nz, unsigned p0, unsigned p1, unsigned P,
    const double *__restrict__ ex, const double *__restrict__ ey, const double *__restrict__ ez,
    double *__restrict__ hx, double *__restrict__ hy, double *__restrict__ hz,
    const double *__restrict__ chx, const double *__restrict__ chy, const double *__restrict__ chz)
{
    const unsigned f = p0 + blockIdx.x * kBlockD + threadIdx.x;  // pairs [p0, p1) of the grid
    if (f >= p1) return;
    const unsigned row = f / P;
    const int k0 = (int)(f - row * P) * 2;
    const int i = (int)(row / (unsigned)ny);
    const int j = (int)row - i * ny;
    const size_t id = (size_t)row * nzp + k0;
    const size_t sx = (size_t)ny * nzp;

    const D2 ex0 = ld2(ex + id), ey0 = ld2(ey + id), ez0 = ld2(ez + id);
    if (j > 0) {
        const D2 h = ld2(hx + id), c = ldc(chx, id), e1 = ld2(ez + id - nzp);
        const double eyp = ey[id - 1];  // previous row's last element when k0 == 0 (row j-1 exists)
        D2 r;
#pragma unroll
        for (int l = 0; l < 2; l++) {
            const double e2 = l > 0 ? ey0.v[l - 1] : eyp;
            const double v = h.v[l] - c.v[l] * ((ez0.v[l] - e1.v[l]) - (ey0.v[l] - e2));
            const int k = k0 + l;
            r.v[l] = (k > 0 && k < nz) ? v : v * 0.0;
        }
        st2(hx + id, r);
    }
    if (i > 0) {
        const D2 h = ld2(hy + id), c = ldc(chy, id), e2 = ld2(ez + id - sx);
        const double exp_ = ex[id - 1];
        D2 r;
#pragma unroll
        for (int l = 0; l < 2; l++) {
            const double e1 = l > 0 ? ex0.v[l - 1] : exp_;
            const double v = h.v[l] - c.v[l] * ((ex0.v[l] - e1) - (ez0.v[l] - e2.v[l]));
            const int k = k0 + l;
            r.v[l] = (k > 0 && k < nz) ? v : v * 0.0;
        }
        st2(hy + id, r);
        if (j > 0) {
            const D2 h3 = ld2(hz + id), c3 = ldc(chz, id), g1 = ld2(ey + id - sx), g2 = ld2(ex + id - nzp);
            D2 r3;
#pragma unroll
            for (int l = 0; l < 2; l++) {
                const double v = h3.v[l] - c3.v[l] * ((ey0.v[l] - g1.v[l]) - (ex0.v[l] - g2.v[l]));
                r3.v[l] = (k0 + l < nz) ? v : v * 0.0;
            }
            st2(hz + id, r3);
        }
    }
}

int check_grid_f64(int nx, int ny, int nzp, int pad)
{
    if (nx < 2 || ny < 2 || nzp < 2) return fail(MIFDTD_EINVAL, "grid (%d,%d,%d) too small", nx, ny, nzp);
    if (nzp % 2) return fail(MIFDTD_EALIGN, "nz_pitch %d is not a multiple of 2 (16-byte rows of doubles)", nzp);
    if (pad < 0 || pad > nzp - 2) return fail(MIFDTD_EINVAL, "pad %d invalid for nz_pitch %d", pad, nzp);
    const unsigned long long pairs = (unsigned long long)nx * ny * (nzp / 2);
    if (pairs >= (1ull << 31)) return fail(MIFDTD_ERANGE, "grid (%d,%d,%d) exceeds 2^31 pairs", nx, ny, nzp);
    return 0;
}

int launch_core_f64(bool e_half, int nx, int ny, int nzp, int pad, double *ex, double *ey, double *ez, double *hx,
                    double *hy, double *hz, const double *c0, const double *c1, const double *c2, hipStream_t st,
                    long long idx0 = 0, long long nmax = -1)
{
    if (int rc = check_grid_f64(nx, ny, nzp, pad)) return rc;
    const double *ptrs[6] = {ex, ey, ez, hx, hy, hz};
    for (const double *p : ptrs) {
        if (!p) return fail(MIFDTD_ENULL, "field pointer is NULL");
        if (!aligned16(p)) return fail(MIFDTD_EALIGN, "field pointer %p is not 16-byte aligned", (const void *)p);
    }
    if (c0 || c1 || c2) {
        if (!(c0 && c1 && c2)) return fail(MIFDTD_ENULL, "coefficient arrays must be all set or all NULL");
        if (!aligned16(c0) || !aligned16(c1) || !aligned16(c2))
            return fail(MIFDTD_EALIGN, "coefficient pointer is not 16-byte aligned");
    }
    const unsigned P = (unsigned)(nzp / 2);
    const unsigned npairs = (unsigned)nx * ny * P;
    unsigned p0 = 0, p1 = npairs;
    if (nmax >= 0) {  // the split form's element range, a pair per 2-element vector
        const long long n = 2ll * npairs;
        if (idx0 < 0 || idx0 > nmax || nmax > n)
            return fail(MIFDTD_EINVAL, "element range [%lld, %lld) outside the grid's %lld", idx0, nmax, n);
        if (idx0 % 2) return fail(MIFDTD_EALIGN, "idx0 %lld does not start a 2-element vector", idx0);
        p0 = (unsigned)(idx0 / 2);
        p1 = (unsigned)((nmax + 1) / 2);
    }
    if (p1 == p0) return 0;
    const unsigned grid = (p1 - p0 + kBlockD - 1) / kBlockD;
    const int nz = nzp - pad;
    if (e_half) {
        k_update_e_f64<<<grid, kBlockD, 0, st>>>(nx, ny, nzp, nz, p0, p1, P, ex, ey, ez, hx, hy, hz, c0, c1, c2);
        MIFDTD_CHECK_LAUNCH("update_e (double) launch");
    } else {
        k_update_h_f64<<<grid, kBlockD, 0, st>>>(ny, nzp, nz, p0, p1, P, ex, ey, ez, hx, hy, hz, c0, c1, c2);
        MIFDTD_CHECK_LAUNCH("update_h (double) launch");
    }
    return 0;
}

}  // namespace

}  // namespace mifdtd

using namespace mifdtd;

extern "C" {

int mifdtd_update_e_f64(int nx, int ny, int nz_pitch, int pad, double *ex, double *ey, double *ez,
                        const double *hx, const double *hy, const double *hz,
                        const double *cex, const double *cey, const double *cez, mifdtd_stream_t stream)
{
    return launch_core_f64(true, nx, ny, nz_pitch, pad, ex, ey, ez, const_cast<double *>(hx), const_cast<double *>(hy),
                           const_cast<double *>(hz), cex, cey, cez, to_stream(stream));
}

int mifdtd_update_h_f64(int nx, int ny, int nz_pitch, int pad, const double *ex, const double *ey, const double *ez,
                        double *hx, double *hy, double *hz,
                        const double *chx, const double *chy, const double *chz, mifdtd_stream_t stream)
{
    return launch_core_f64(false, nx, ny, nz_pitch, pad, const_cast<double *>(ex), const_cast<double *>(ey),
                           const_cast<double *>(ez), hx, hy, hz, chx, chy, chz, to_stream(stream));
}

int mifdtd_update_e_f64_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax, double *ex,
                              double *ey, double *ez, const double *hx, const double *hy, const double *hz,
                              const double *cex, const double *cey, const double *cez, mifdtd_stream_t stream)
{
    if (nmax < 0) return fail(MIFDTD_EINVAL, "nmax %lld < 0", nmax);
    return launch_core_f64(true, nx, ny, nz_pitch, pad, ex, ey, ez, const_cast<double *>(hx), const_cast<double *>(hy),
                           const_cast<double *>(hz), cex, cey, cez, to_stream(stream), idx0, nmax);
}

int mifdtd_update_h_f64_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax,
                              const double *ex, const double *ey, const double *ez, double *hx, double *hy,
                              double *hz, const double *chx, const double *chy, const double *chz,
                              mifdtd_stream_t stream)
{
    if (nmax < 0) return fail(MIFDTD_EINVAL, "nmax %lld < 0", nmax);
    return launch_core_f64(false, nx, ny, nz_pitch, pad, const_cast<double *>(ex), const_cast<double *>(ey),
                           const_cast<double *>(ez), hx, hy, hz, chx, chy, chz, to_stream(stream), idx0, nmax);
}

}  // extern "C"
