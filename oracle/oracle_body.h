/*
 * ORACLE BODY — test infrastructure only.  Included twice by oracle.c, once with
 * REAL=float (suffix _f32) and once with REAL=double (suffix _f64).
 *
 * Scalar C restatement of the reference FDTD hot path on the KEMP field layout:
 * C-order (nx, ny, nzp) arrays, z contiguous, nzp = nz + pad (pad lanes are zero).
 * Every arithmetic expression keeps the reference's evaluation order so that the
 * float build is bit-comparable with the reference's SSE code (no FMA contraction:
 * the Makefile builds with -ffp-contract=off).
 */

#define IX(i, j, k) (((size_t)(i) * (size_t)ny + (size_t)(j)) * (size_t)nzp + (size_t)(k))

/* K7/cpu/src/core.c:11-55 (update_e) with the masks of K7/cpu/core.py:33-49:
 *   ex += cex*((hz[j+1]-hz) - (hy[k+1]-hy))   if j<ny-1   (k >= nz-1 lanes -> 0)
 *   ey += cey*((hx[k+1]-hx) - (hz[i+1]-hz))   if i<nx-1   (k >= nz-1 lanes -> 0)
 *   ez += cez*((hy[i+1]-hy) - (hx[j+1]-hx))   if i<nx-1 && j<ny-1 (pad lanes -> 0)
 * A NULL coefficient pointer means the constant 0.5 (K7/cpu/core.py:63-65).
 * Masked lanes are the computed value multiplied by 0, as the reference's SSE mask
 * multiply does, so even the sign of the resulting zero matches. */
void SFX(ora_update_e)(int nx, int ny, int nzp, int pad,
                       REAL *ex, REAL *ey, REAL *ez,
                       const REAL *hx, const REAL *hy, const REAL *hz,
                       const REAL *cex, const REAL *cey, const REAL *cez)
{
    const int nz = nzp - pad;
    const REAL half = (REAL)0.5;
    const REAL zero = (REAL)0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nzp; k++) {
                const size_t id = IX(i, j, k);
                /* hy[id+1] / hx[id+1] at k = nzp-1 is the first element of the next row,
                 * as in the reference's flat loop; it exists whenever the branch runs. */
                if (j < ny - 1) {
                    const REAL c = cex ? cex[id] : half;
                    const REAL v = ex[id] + c * ((hz[id + nzp] - hz[id]) - (hy[id + 1] - hy[id]));
                    ex[id] = k < nz - 1 ? v : v * zero;
                }
                if (i < nx - 1) {
                    const REAL c = cey ? cey[id] : half;
                    const REAL v = ey[id] + c * ((hx[id + 1] - hx[id]) - (hz[id + (size_t)ny * nzp] - hz[id]));
                    ey[id] = k < nz - 1 ? v : v * zero;
                }
                if (i < nx - 1 && j < ny - 1) {
                    const REAL c = cez ? cez[id] : half;
                    const REAL v = ez[id] + c * ((hy[id + (size_t)ny * nzp] - hy[id]) - (hx[id + nzp] - hx[id]));
                    ez[id] = k < nz ? v : v * zero;
                }
            }
}

/* K7/cpu/src/core.c:59-105 (update_h):
 *   hx -= chx*((ez-ez[j-1]) - (ey-ey[k-1]))   if j>0   (k==0 and pad lanes -> 0)
 *   hy -= chy*((ex-ex[k-1]) - (ez-ez[i-1]))   if i>0   (k==0 and pad lanes -> 0)
 *   hz -= chz*((ey-ey[i-1]) - (ex-ex[j-1]))   if i>0 && j>0 (pad lanes -> 0) */
void SFX(ora_update_h)(int nx, int ny, int nzp, int pad,
                       const REAL *ex, const REAL *ey, const REAL *ez,
                       REAL *hx, REAL *hy, REAL *hz,
                       const REAL *chx, const REAL *chy, const REAL *chz)
{
    const int nz = nzp - pad;
    const REAL half = (REAL)0.5;
    const REAL zero = (REAL)0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nzp; k++) {
                const size_t id = IX(i, j, k);
                /* ey[id-1] / ex[id-1] at k = 0 is the last element of the previous row. */
                if (j > 0) {
                    const REAL c = chx ? chx[id] : half;
                    const REAL v = hx[id] - c * ((ez[id] - ez[id - nzp]) - (ey[id] - ey[id - 1]));
                    hx[id] = (k > 0 && k < nz) ? v : v * zero;
                }
                if (i > 0) {
                    const REAL c = chy ? chy[id] : half;
                    const REAL v = hy[id] - c * ((ex[id] - ex[id - 1]) - (ez[id] - ez[id - (size_t)ny * nzp]));
                    hy[id] = (k > 0 && k < nz) ? v : v * zero;
                }
                if (i > 0 && j > 0) {
                    const REAL c = chz ? chz[id] : half;
                    const REAL v = hz[id] - c * ((ey[id] - ey[id - (size_t)ny * nzp]) - (ex[id] - ex[id - nzp]));
                    hz[id] = k < nz ? v : v * zero;
                }
            }
}

/* Cells that the core update writes (same masks as above).  The CPML correction is
 * applied, and its psi advanced, only on these cells, so the walls of the KEMP box
 * (tangential E on the high faces, tangential H on the low faces, pad lanes) stay 0. */
static inline int SFX(valid_e)(int c, int i, int j, int k, int nx, int ny, int nz)
{
    if (c == 0) return j < ny - 1 && k < nz - 1;
    if (c == 1) return i < nx - 1 && k < nz - 1;
    return i < nx - 1 && j < ny - 1 && k < nz;
}
static inline int SFX(valid_h)(int c, int i, int j, int k, int nx, int ny, int nz)
{
    (void)nx; (void)ny;
    if (c == 0) return j > 0 && k > 0 && k < nz;
    if (c == 1) return i > 0 && k > 0 && k < nz;
    return i > 0 && j > 0 && k < nz;
}

/* CPML E-slab update, one face (axis 0/1/2 = x/y/z, side 0 = front (low), 1 = back).
 * Restates KU/cpu/core/cpml_non_kappa.c:14-88 (x), :156-224 (y), :294-362 (z) and the
 * uniform x/y/z indexing of Y2/CUDA/fdtd3d_018_cpu_250x250x320cell_15npml.c:156-280:
 *   psi = b[d]*psi + a[d]*(H(+1) - H);  E_t1 -= CE*psi1;  E_t2 += CE*psi2
 *   x: psi1 -> ey from hz,  psi2 -> ez from hy
 *   y: psi1 -> ez from hx,  psi2 -> ex from hz
 *   z: psi1 -> ex from hy,  psi2 -> ey from hx
 * Slab planes along the axis (n = extent of the axis, n_e = n for x/y, nz for z):
 *   front: p in [0,npml), index p,              table entry p+1
 *   back : p in [0,npml), index n_e-1-npml+p,   table entry npml+1+p
 * psi layout: x (npml, ny, nzp), y (nx, npml, nzp), z (nx, ny, npml). */
void SFX(ora_cpml_e)(int axis, int side, int nx, int ny, int nzp, int pad, int npml,
                     REAL *ex, REAL *ey, REAL *ez,
                     const REAL *hx, const REAL *hy, const REAL *hz,
                     const REAL *cex, const REAL *cey, const REAL *cez,
                     REAL *psi1, REAL *psi2, const REAL *b, const REAL *a)
{
    const int nz = nzp - pad;
    const REAL half = (REAL)0.5;
    const int n_e = axis == 0 ? nx : axis == 1 ? ny : nz;
    const int base = side ? n_e - 1 - npml : 0;
    const int toff = side ? npml + 1 : 1;
    const size_t sx = (size_t)ny * nzp, sy = (size_t)nzp, sz = 1;
    /* iterate the slab only: the slab axis runs over [base, base+npml) */
    const int i0 = axis == 0 ? base : 0, i1 = axis == 0 ? base + npml : nx;
    const int j0 = axis == 1 ? base : 0, j1 = axis == 1 ? base + npml : ny;
    const int k0 = axis == 2 ? base : 0, k1 = axis == 2 ? base + npml : nzp;
#pragma omp parallel for schedule(static)
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++)
            for (int k = k0; k < k1; k++) {
                const int p = (axis == 0 ? i : axis == 1 ? j : k) - base;
                const size_t id = IX(i, j, k);
                size_t pid;
                if (axis == 0) pid = ((size_t)p * ny + j) * nzp + k;
                else if (axis == 1) pid = ((size_t)i * npml + p) * nzp + k;
                else pid = ((size_t)i * ny + j) * npml + p;
                const REAL bb = b[toff + p], aa = a[toff + p];
                if (axis == 0) {
                    if (SFX(valid_e)(1, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi1[pid] + aa * (hz[id + sx] - hz[id]);
                        psi1[pid] = t;
                        ey[id] = ey[id] - (cey ? cey[id] : half) * t;
                    }
                    if (SFX(valid_e)(2, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi2[pid] + aa * (hy[id + sx] - hy[id]);
                        psi2[pid] = t;
                        ez[id] = ez[id] + (cez ? cez[id] : half) * t;
                    }
                } else if (axis == 1) {
                    if (SFX(valid_e)(2, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi1[pid] + aa * (hx[id + sy] - hx[id]);
                        psi1[pid] = t;
                        ez[id] = ez[id] - (cez ? cez[id] : half) * t;
                    }
                    if (SFX(valid_e)(0, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi2[pid] + aa * (hz[id + sy] - hz[id]);
                        psi2[pid] = t;
                        ex[id] = ex[id] + (cex ? cex[id] : half) * t;
                    }
                } else {
                    if (SFX(valid_e)(0, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi1[pid] + aa * (hy[id + sz] - hy[id]);
                        psi1[pid] = t;
                        ex[id] = ex[id] - (cex ? cex[id] : half) * t;
                    }
                    if (SFX(valid_e)(1, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi2[pid] + aa * (hx[id + sz] - hx[id]);
                        psi2[pid] = t;
                        ey[id] = ey[id] + (cey ? cey[id] : half) * t;
                    }
                }
            }
}

/* CPML H-slab update, one face.  KU/cpu/core/cpml_non_kappa.c:91-153 (x), :227-291 (y),
 * :365-429 (z):  psi = b*psi + a*(E - E(-1));  H_t1 += CH*psi1;  H_t2 -= CH*psi2
 *   x: psi1 -> hy from ez,  psi2 -> hz from ey
 *   y: psi1 -> hz from ex,  psi2 -> hx from ez
 *   z: psi1 -> hx from ey,  psi2 -> hy from ex
 *   front: p in [0,npml), index p (p = 0 is the wall and is skipped by the mask),
 *          table entry p+1;  back: index n_e-npml+p, table entry npml+2+p. */
void SFX(ora_cpml_h)(int axis, int side, int nx, int ny, int nzp, int pad, int npml,
                     const REAL *ex, const REAL *ey, const REAL *ez,
                     REAL *hx, REAL *hy, REAL *hz,
                     const REAL *chx, const REAL *chy, const REAL *chz,
                     REAL *psi1, REAL *psi2, const REAL *b, const REAL *a)
{
    const int nz = nzp - pad;
    const REAL half = (REAL)0.5;
    const int n_e = axis == 0 ? nx : axis == 1 ? ny : nz;
    const int base = side ? n_e - npml : 0;
    const int toff = side ? npml + 2 : 1;
    const size_t sx = (size_t)ny * nzp, sy = (size_t)nzp, sz = 1;
    /* iterate the slab only: the slab axis runs over [base, base+npml) */
    const int i0 = axis == 0 ? base : 0, i1 = axis == 0 ? base + npml : nx;
    const int j0 = axis == 1 ? base : 0, j1 = axis == 1 ? base + npml : ny;
    const int k0 = axis == 2 ? base : 0, k1 = axis == 2 ? base + npml : nzp;
#pragma omp parallel for schedule(static)
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++)
            for (int k = k0; k < k1; k++) {
                const int p = (axis == 0 ? i : axis == 1 ? j : k) - base;
                const size_t id = IX(i, j, k);
                size_t pid;
                if (axis == 0) pid = ((size_t)p * ny + j) * nzp + k;
                else if (axis == 1) pid = ((size_t)i * npml + p) * nzp + k;
                else pid = ((size_t)i * ny + j) * npml + p;
                const REAL bb = b[toff + p], aa = a[toff + p];
                if (axis == 0) {
                    if (SFX(valid_h)(1, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi1[pid] + aa * (ez[id] - ez[id - sx]);
                        psi1[pid] = t;
                        hy[id] = hy[id] + (chy ? chy[id] : half) * t;
                    }
                    if (SFX(valid_h)(2, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi2[pid] + aa * (ey[id] - ey[id - sx]);
                        psi2[pid] = t;
                        hz[id] = hz[id] - (chz ? chz[id] : half) * t;
                    }
                } else if (axis == 1) {
                    if (SFX(valid_h)(2, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi1[pid] + aa * (ex[id] - ex[id - sy]);
                        psi1[pid] = t;
                        hz[id] = hz[id] + (chz ? chz[id] : half) * t;
                    }
                    if (SFX(valid_h)(0, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi2[pid] + aa * (ez[id] - ez[id - sy]);
                        psi2[pid] = t;
                        hx[id] = hx[id] - (chx ? chx[id] : half) * t;
                    }
                } else {
                    if (SFX(valid_h)(0, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi1[pid] + aa * (ey[id] - ey[id - sz]);
                        psi1[pid] = t;
                        hx[id] = hx[id] + (chx ? chx[id] : half) * t;
                    }
                    if (SFX(valid_h)(1, i, j, k, nx, ny, nz)) {
                        REAL t = bb * psi2[pid] + aa * (ex[id] - ex[id - sz]);
                        psi2[pid] = t;
                        hy[id] = hy[id] - (chy ? chy[id] : half) * t;
                    }
                }
            }
}

/* Source injection, K7/naive/incident_direct.py:52-70 and K7/cpu/get_set_fields.py:110-123:
 *   f[region] += REAL(tfunc(t)) * spatial_value      (overwrite: f[region] = ...)
 * The region is the inclusive box [i0,i1]x[j0,j1]x[k0,k1]; sv (if non-NULL) is the
 * C-order spatial array of the box, else the scalar svs. */
void SFX(ora_src)(REAL *f, int nx, int ny, int nzp,
                  int i0, int i1, int j0, int j1, int k0, int k1,
                  REAL tval, const REAL *sv, REAL svs, int overwrite)
{
    (void)nx;
    const int lj = j1 - j0 + 1, lk = k1 - k0 + 1;
    for (int i = i0; i <= i1; i++)
        for (int j = j0; j <= j1; j++)
            for (int k = k0; k <= k1; k++) {
                const REAL s = sv ? sv[((size_t)(i - i0) * lj + (j - j0)) * lk + (k - k0)] : svs;
                const REAL v = tval * s;
                const size_t id = IX(i, j, k);
                f[id] = overwrite ? v : f[id] + v;
            }
}

#undef IX

/* ---------------- 4th-order symplectic path (x contiguous) ----------------
 * S4/Python-C/dielectric.c:5-49 (update_h) and :52-96 (update_e).  Fortran order:
 * idx = i + j*nx + k*nx*ny.  Loops cover i<nx-1, j<ny-1, k<nz-1 as the C loops do;
 * guards keep the 2-cell stencil inside the grid. */
#define FX(i, j, k) ((size_t)(i) + (size_t)(j) * (size_t)nx + (size_t)(k) * (size_t)nx * (size_t)ny)

void SFX(ora_s4_update_h)(int nx, int ny, int nz, REAL cl,
                          const REAL *ex, const REAL *ey, const REAL *ez,
                          REAL *hx, REAL *hy, REAL *hz,
                          const REAL *cx, const REAL *cy, const REAL *cz)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const REAL c27 = (REAL)27;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz - 1; k++)
        for (int j = 0; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) {
                const size_t id = FX(i, j, k);
                if (j > 1 && j < ny - 1 && k > 1 && k < nz - 1)
                    hx[id] -= cl * cx[id] * (c27 * (ez[id] - ez[id - sy] - ey[id] + ey[id - sz])
                                             - (ez[id + sy] - ez[id - 2 * sy] - ey[id + sz] + ey[id - 2 * sz]));
                if (i > 1 && i < nx - 1 && k > 1 && k < nz - 1)
                    hy[id] -= cl * cy[id] * (c27 * (ex[id] - ex[id - sz] - ez[id] + ez[id - 1])
                                             - (ex[id + sz] - ex[id - 2 * sz] - ez[id + 1] + ez[id - 2]));
                if (i > 1 && i < nx - 1 && j > 1 && j < ny - 1)
                    hz[id] -= cl * cz[id] * (c27 * (ey[id] - ey[id - 1] - ex[id] + ex[id - sy])
                                             - (ey[id + 1] - ey[id - 2] - ex[id + sy] + ex[id - 2 * sy]));
            }
}

void SFX(ora_s4_update_e)(int nx, int ny, int nz, REAL dl,
                          REAL *ex, REAL *ey, REAL *ez,
                          const REAL *hx, const REAL *hy, const REAL *hz,
                          const REAL *cx, const REAL *cy, const REAL *cz)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const REAL c27 = (REAL)27;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz - 1; k++)
        for (int j = 0; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) {
                const size_t id = FX(i, j, k);
                if (j > 0 && j < ny - 2 && k > 0 && k < nz - 2)
                    ex[id] += dl * cx[id] * (c27 * (hz[id + sy] - hz[id] - hy[id + sz] + hy[id])
                                             - (hz[id + 2 * sy] - hz[id - sy] - hy[id + 2 * sz] + hy[id - sz]));
                if (i > 0 && i < nx - 2 && k > 0 && k < nz - 2)
                    ey[id] += dl * cy[id] * (c27 * (hx[id + sz] - hx[id] - hz[id + 1] + hz[id])
                                             - (hx[id + 2 * sz] - hx[id - sz] - hz[id + 2] + hz[id - 1]));
                if (i > 0 && i < nx - 2 && j > 0 && j < ny - 2)
                    ez[id] += dl * cz[id] * (c27 * (hy[id + 1] - hy[id] - hx[id + sy] + hx[id])
                                             - (hy[id + 2] - hy[id - 1] - hx[id + 2 * sy] + hx[id - sy]));
            }
}
/* ---- 4th-order CPML (new: the reference's 4th-order path has no absorbing boundary).
 * Same face/table/sign conventions as the Yee CPML of KU/cpu/core/cpml_non_kappa.c
 * (x: psi1 -> ey from hz, psi2 -> ez from hy; y: psi1 -> ez from hx, psi2 -> ex from
 * hz; z: psi1 -> ex from hy, psi2 -> ey from hx; E front plane p uses table entry p+1,
 * back plane n-1-npml+p entry npml+1+p; H front p -> p+1, back n-npml+p -> npml+2+p),
 * with the 4th-order difference of S4/Python-C/dielectric.c:82-89 / :35-42 as driver:
 *   E half: D = 27*(f[+1]-f[0]) - (f[+2]-f[-1]),  H half: D = 27*(f[0]-f[-1]) - (f[+1]-f[-2])
 *   psi = b*psi + a*D;   E -/+= w*CE*psi;   H +/-= w*CH*psi
 * w is the stage weight (d_s for E, c_s for H) and b/a are the tables of that stage's
 * time step (b = exp(-sigma*w*dt), a = b-1).  Applied only where the 4th-order core
 * writes (guards of dielectric.c:34-41,81-88 and its i<nx-1, j<ny-1, k<nz-1 loops).
 * psi layouts (x fastest): x face (npml, ny, nz), y face (nx, npml, nz), z face (nx, ny, npml). */
static inline int SFX(s4_valid_e)(int c, int i, int j, int k, int nx, int ny, int nz)
{
    if (i > nx - 2 || j > ny - 2 || k > nz - 2) return 0;
    if (c == 0) return j > 0 && j < ny - 2 && k > 0 && k < nz - 2;
    if (c == 1) return i > 0 && i < nx - 2 && k > 0 && k < nz - 2;
    return i > 0 && i < nx - 2 && j > 0 && j < ny - 2;
}
static inline int SFX(s4_valid_h)(int c, int i, int j, int k, int nx, int ny, int nz)
{
    if (i > nx - 2 || j > ny - 2 || k > nz - 2) return 0;
    if (c == 0) return j > 1 && k > 1;
    if (c == 1) return i > 1 && k > 1;
    return i > 1 && j > 1;
}

void SFX(ora_s4_cpml_e)(int axis, int side, int nx, int ny, int nz, int npml, REAL w,
                        REAL *ex, REAL *ey, REAL *ez,
                        const REAL *hx, const REAL *hy, const REAL *hz,
                        const REAL *cx, const REAL *cy, const REAL *cz,
                        REAL *psi1, REAL *psi2, const REAL *b, const REAL *a)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const REAL c27 = (REAL)27;
    const int n_e = axis == 0 ? nx : axis == 1 ? ny : nz;
    const size_t s = axis == 0 ? 1 : axis == 1 ? sy : sz;
    /* tangential pair of this face: (E1, H driving E1, C1, comp1), (E2, ...) */
    REAL *e1 = axis == 0 ? ey : axis == 1 ? ez : ex;
    REAL *e2 = axis == 0 ? ez : axis == 1 ? ex : ey;
    const REAL *g1 = axis == 0 ? hz : axis == 1 ? hx : hy;
    const REAL *g2 = axis == 0 ? hy : axis == 1 ? hz : hx;
    const REAL *c1 = axis == 0 ? cy : axis == 1 ? cz : cx;
    const REAL *c2 = axis == 0 ? cz : axis == 1 ? cx : cy;
    const int comp1 = axis == 0 ? 1 : axis == 1 ? 2 : 0;
    const int comp2 = axis == 0 ? 2 : axis == 1 ? 0 : 1;
    for (int p = 0; p < npml; p++) {
        const int q = side ? n_e - 1 - npml + p : p;
        const int t = side ? npml + 1 + p : p + 1;
        /* (u, v) run over the two tangential axes of the face */
        const int nu = axis == 0 ? ny : nx, nv = axis == 2 ? ny : nz;
        for (int v = 0; v < nv; v++)
            for (int u = 0; u < nu; u++) {
                    const int ii = axis == 0 ? q : u;
                    const int jj = axis == 0 ? u : axis == 1 ? q : v;
                    const int kk = axis == 2 ? q : v;
                    const size_t id = FX(ii, jj, kk);
                    const size_t pid = axis == 0 ? (size_t)p + (size_t)npml * ((size_t)jj + (size_t)ny * kk)
                                     : axis == 1 ? (size_t)ii + (size_t)nx * ((size_t)p + (size_t)npml * kk)
                                                 : (size_t)ii + (size_t)nx * ((size_t)jj + (size_t)ny * p);
                    if (SFX(s4_valid_e)(comp1, ii, jj, kk, nx, ny, nz)) {
                        const REAL d = c27 * (g1[id + s] - g1[id]) - (g1[id + 2 * s] - g1[id - s]);
                        const REAL tt = b[t] * psi1[pid] + a[t] * d;
                        psi1[pid] = tt;
                        e1[id] -= w * c1[id] * tt;
                    }
                    if (SFX(s4_valid_e)(comp2, ii, jj, kk, nx, ny, nz)) {
                        const REAL d = c27 * (g2[id + s] - g2[id]) - (g2[id + 2 * s] - g2[id - s]);
                        const REAL tt = b[t] * psi2[pid] + a[t] * d;
                        psi2[pid] = tt;
                        e2[id] += w * c2[id] * tt;
                    }
                }
    }
}

void SFX(ora_s4_cpml_h)(int axis, int side, int nx, int ny, int nz, int npml, REAL w,
                        const REAL *ex, const REAL *ey, const REAL *ez,
                        REAL *hx, REAL *hy, REAL *hz,
                        const REAL *cx, const REAL *cy, const REAL *cz,
                        REAL *psi1, REAL *psi2, const REAL *b, const REAL *a)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const REAL c27 = (REAL)27;
    const int n_e = axis == 0 ? nx : axis == 1 ? ny : nz;
    const size_t s = axis == 0 ? 1 : axis == 1 ? sy : sz;
    /* x: psi1 -> hy from ez (+), psi2 -> hz from ey (-); y: psi1 -> hz from ex (+),
     * psi2 -> hx from ez (-); z: psi1 -> hx from ey (+), psi2 -> hy from ex (-) */
    REAL *h1 = axis == 0 ? hy : axis == 1 ? hz : hx;
    REAL *h2 = axis == 0 ? hz : axis == 1 ? hx : hy;
    const REAL *g1 = axis == 0 ? ez : axis == 1 ? ex : ey;
    const REAL *g2 = axis == 0 ? ey : axis == 1 ? ez : ex;
    const REAL *c1 = axis == 0 ? cy : axis == 1 ? cz : cx;
    const REAL *c2 = axis == 0 ? cz : axis == 1 ? cx : cy;
    const int comp1 = axis == 0 ? 1 : axis == 1 ? 2 : 0;
    const int comp2 = axis == 0 ? 2 : axis == 1 ? 0 : 1;
    for (int p = 0; p < npml; p++) {
        const int q = side ? n_e - npml + p : p;
        const int t = side ? npml + 2 + p : p + 1;
        /* (u, v) run over the two tangential axes of the face */
        const int nu = axis == 0 ? ny : nx, nv = axis == 2 ? ny : nz;
        for (int v = 0; v < nv; v++)
            for (int u = 0; u < nu; u++) {
                    const int ii = axis == 0 ? q : u;
                    const int jj = axis == 0 ? u : axis == 1 ? q : v;
                    const int kk = axis == 2 ? q : v;
                    const size_t id = FX(ii, jj, kk);
                    const size_t pid = axis == 0 ? (size_t)p + (size_t)npml * ((size_t)jj + (size_t)ny * kk)
                                     : axis == 1 ? (size_t)ii + (size_t)nx * ((size_t)p + (size_t)npml * kk)
                                                 : (size_t)ii + (size_t)nx * ((size_t)jj + (size_t)ny * p);
                    if (SFX(s4_valid_h)(comp1, ii, jj, kk, nx, ny, nz)) {
                        const REAL d = c27 * (g1[id] - g1[id - s]) - (g1[id + s] - g1[id - 2 * s]);
                        const REAL tt = b[t] * psi1[pid] + a[t] * d;
                        psi1[pid] = tt;
                        h1[id] += w * c1[id] * tt;
                    }
                    if (SFX(s4_valid_h)(comp2, ii, jj, kk, nx, ny, nz)) {
                        const REAL d = c27 * (g2[id] - g2[id - s]) - (g2[id + s] - g2[id - 2 * s]);
                        const REAL tt = b[t] * psi2[pid] + a[t] * d;
                        psi2[pid] = tt;
                        h2[id] -= w * c2[id] * tt;
                    }
                }
    }
}
#undef FX

