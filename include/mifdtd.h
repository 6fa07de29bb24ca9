/*
 * mifdtd — MI355X (gfx950) FDTD time-step engine: C ABI.
 *
 * Every pointer argument is a DEVICE pointer (e.g. torch.Tensor.data_ptr() of a
 * CUDA/HIP tensor) unless stated otherwise.  Arrays use the KEMP layout of the
 * reference: C-order (nx, ny, nz_pitch) float32, z contiguous, nz_pitch = nz + pad
 * with pad lanes held at zero (K7/cpu/fields.py:27-32, K7/gpu/fields.py:45-50).
 * nz_pitch must be a multiple of 4 and every array 16-byte aligned.
 *
 * Functions returning int return 0 on success, a positive hipError_t from the
 * runtime, or a negative MIFDTD_E* code for rejected arguments; the message of the
 * last failure on the calling thread is available from mifdtd_last_error().
 * The void functions (the reference's own names and argument orders: s4_update_h/e
 * here, update_e/update_h in the KEMP program libraries of mifdtd_kemp.h) report
 * failures only through mifdtd_last_error(), as the reference's void kernels have no
 * error channel (K7/cpu/src/core.c:11,59).
 *
 * All work is enqueued on the given HIP stream (NULL = the null stream) and is
 * asynchronous with respect to the host.
 */
#ifndef MIFDTD_H
#define MIFDTD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void *mifdtd_stream_t; /* hipStream_t */

enum {
    MIFDTD_OK = 0,
    MIFDTD_EINVAL = -1,   /* bad size / pad / npml / region */
    MIFDTD_EALIGN = -2,   /* misaligned pointer or nz_pitch % 4 != 0 */
    MIFDTD_ENULL = -3,    /* required pointer is NULL */
    MIFDTD_ERANGE = -4    /* size beyond the supported index range */
};

const char *mifdtd_last_error(void);
int mifdtd_version(void);

/* ------------------------------------------------------------------ core ----
 * The KEMP core programs.  The reference generates one shared library per
 * (precision_float, coeff_use, pad) from K7/cpu/src/core.c:11 (update_e) and :59
 * (update_h) (K7/cpu/core.py:52-85) and binds update_e/update_h with 3 ints and 6 or 9
 * arrays (K7/cpu/core.py:89-105), the pad baked into the program's masks
 * (K7/cpu/core.py:33-49).  The same programs exist here as one small library each
 * (libkemp_<single|double>_<none|e|h|eh>_pad<p>.so, declared in mifdtd_kemp.h), each
 * forwarding to the status-returning calls below with its own pad.  libmifdtd.so
 * itself exports no update_e/update_h symbol.
 * Semantics (masks as K7/cpu/src/core.c:31,41,51,79-80,90-91,101), nz = nz_pitch - pad:
 *   ex += cex*((hz[j+1]-hz)-(hy[k+1]-hy))  j<ny-1,          lanes k>=nz-1 -> 0
 *   ey += cey*((hx[k+1]-hx)-(hz[i+1]-hz))  i<nx-1,          lanes k>=nz-1 -> 0
 *   ez += cez*((hy[i+1]-hy)-(hx[j+1]-hx))  i<nx-1 && j<ny-1, pad lanes -> 0
 *   hx -= chx*((ez-ez[j-1])-(ey-ey[k-1]))  j>0,             lane 0 and pad -> 0
 *   hy -= chy*((ex-ex[k-1])-(ez-ez[i-1]))  i>0,             lane 0 and pad -> 0
 *   hz -= chz*((ey-ey[i-1])-(ex-ex[j-1]))  i>0 && j>0,      pad lanes -> 0
 * A masked lane stores value*0 (the reference's SSE mask multiply), so signed zeros
 * match.  NULL coefficient arrays mean the constant 0.5 (K7/cpu/core.py:63-65). */
/* update_e / update_h of the 'single' programs, with the pad explicit, a status return
 * and a stream. */
int mifdtd_update_e(int nx, int ny, int nz_pitch, int pad,
                    float *ex, float *ey, float *ez,
                    const float *hx, const float *hy, const float *hz,
                    const float *cex, const float *cey, const float *cez,
                    mifdtd_stream_t stream);
int mifdtd_update_h(int nx, int ny, int nz_pitch, int pad,
                    const float *ex, const float *ey, const float *ez,
                    float *hx, float *hy, float *hz,
                    const float *chx, const float *chy, const float *chz,
                    mifdtd_stream_t stream);

/* The same for the 'double' programs (K7/cpu/core.py:52-54, DTYPE double, 2 lanes;
 * K7/cpu/fields.py:25): float64 arrays, nz_pitch even (16-byte rows), 16-byte aligned
 * pointers, pad in [0, nz_pitch-2]. */
int mifdtd_update_e_f64(int nx, int ny, int nz_pitch, int pad,
                        double *ex, double *ey, double *ez,
                        const double *hx, const double *hy, const double *hz,
                        const double *cex, const double *cey, const double *cez,
                        mifdtd_stream_t stream);
int mifdtd_update_h_f64(int nx, int ny, int nz_pitch, int pad,
                        const double *ex, const double *ey, const double *ez,
                        double *hx, double *hy, double *hz,
                        const double *chx, const double *chy, const double *chz,
                        mifdtd_stream_t stream);

/* The split form (K7/cpu/src/core_split.c:11,59, bound by K7/cpu/core_split.py:89-123 for
 * computation/communication overlap): the same update restricted to the flat element
 * range [idx0, nmax) of the (nx, ny, nz_pitch) arrays, vector by vector as the
 * reference's SSE loop steps (a 4-element vector -- 2 for double -- is updated when its
 * first element lies in the range; idx0 must start a vector).  Cells outside the range
 * are not written. */
int mifdtd_update_e_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax,
                          float *ex, float *ey, float *ez,
                          const float *hx, const float *hy, const float *hz,
                          const float *cex, const float *cey, const float *cez,
                          mifdtd_stream_t stream);
int mifdtd_update_h_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax,
                          const float *ex, const float *ey, const float *ez,
                          float *hx, float *hy, float *hz,
                          const float *chx, const float *chy, const float *chz,
                          mifdtd_stream_t stream);
int mifdtd_update_e_f64_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax,
                              double *ex, double *ey, double *ez,
                              const double *hx, const double *hy, const double *hz,
                              const double *cex, const double *cey, const double *cez,
                              mifdtd_stream_t stream);
int mifdtd_update_h_f64_range(int nx, int ny, int nz_pitch, int pad, long long idx0, long long nmax,
                              const double *ex, const double *ey, const double *ez,
                              double *hx, double *hy, double *hz,
                              const double *chx, const double *chy, const double *chz,
                              mifdtd_stream_t stream);

/* update_e with the CE arrays given as per-component uint8 material indices
 * (mat_c, same shape as the fields, 4-byte aligned) into lut (DEVICE, 3 x 256 float32:
 * row c holds the distinct CE values of component c).  Same results as mifdtd_update_e
 * with the expanded arrays, 39 instead of 48 bytes per cell.  No reference
 * counterpart (a storage form of the CE arrays of K7/cpu/fields.py:112-122 for media
 * with at most 256 distinct values per component). */
int mifdtd_update_e_lut(int nx, int ny, int nz_pitch, int pad,
                        float *ex, float *ey, float *ez,
                        const float *hx, const float *hy, const float *hz,
                        const unsigned char *mat_x, const unsigned char *mat_y,
                        const unsigned char *mat_z, const float *lut, mifdtd_stream_t stream);

/* ------------------------------------------------------------------ CPML ----
 * CPML face updates on the KEMP field layout.  The per-face arithmetic, the signs and
 * the argument order follow kuFDTD's non-kappa CPML (KU/cpu/core/cpml_non_kappa.c:14
 * update_x_e, :91 update_x_h, :156 update_y_e, :227 update_y_h, :294 update_z_e, :365
 * update_z_h; KU/gpu/core/cpml_non_kapa.cu:17-130), with the CPU thread count dropped
 * and the pad, a status return and a stream added.  They are NOT binary drop-ins for
 * kuFDTD's buffers, hence the mifdtd_ prefix: kuFDTD's fields carry a ghost plane and
 * row at index 0 and its psi arrays are (Npml, Ny-1, Nz), (Nx-1, Npml, Nz) and
 * (Nx-1, Ny-1, Npml+1) with the front slab from plane 1 (KU/cpu/cpml.py:35-85,
 * cpml_non_kappa.c:60); here the fields are KEMP's (no ghost layer, nz = nz_pitch) and
 * the psi shapes and slab planes are those below.  The Python Cpml checks psi shapes
 * before any call.  rcmb/rcma are DEVICE tables of 2*(npml+1) floats in kuFDTD's
 * layout (KU/cpu/cpml.py:107-126): E front entry p+1, E back npml+1+p; H front p+1,
 * H back npml+2+p.
 *   x: psi1->ey from hz, psi2->ez from hy   psi shape (npml, ny, nz_pitch)
 *   y: psi1->ez from hx, psi2->ex from hz   psi shape (nx, npml, nz_pitch)
 *   z: psi1->ex from hy, psi2->ey from hx   psi shape (nx, ny, npml)
 *   E:  psi = b*psi + a*(H(+1)-H);  E_1 -= CE*psi1;  E_2 += CE*psi2
 *   H:  psi = b*psi + a*(E-E(-1));  H_1 += 0.5*psi1; H_2 -= 0.5*psi2
 * Slab planes: E front p, E back n_e-1-npml+p, H front p, H back n_e-npml+p
 * (p in [0,npml), n_e = nx, ny or nz - pad); updates only cells the core writes.
 * Requires n_e >= 2*npml+2 on the axis and 1 <= npml <= MIFDTD_MAX_NPML.
 * cex..cez may be NULL (the constant 0.5). */
#define MIFDTD_MAX_NPML 64
int mifdtd_cpml_face_x_e(int nx, int ny, int nz_pitch, int pad, int npml, int backward,
                         float *ey, float *ez, const float *hy, const float *hz,
                         const float *cey, const float *cez, float *psi1, float *psi2,
                         const float *rcmb, const float *rcma, mifdtd_stream_t stream);
int mifdtd_cpml_face_x_h(int nx, int ny, int nz_pitch, int pad, int npml, int backward,
                         const float *ey, const float *ez, float *hy, float *hz,
                         float *psi1, float *psi2, const float *rcmb, const float *rcma,
                         mifdtd_stream_t stream);
int mifdtd_cpml_face_y_e(int nx, int ny, int nz_pitch, int pad, int npml, int backward,
                         float *ex, float *ez, const float *hx, const float *hz,
                         const float *cex, const float *cez, float *psi1, float *psi2,
                         const float *rcmb, const float *rcma, mifdtd_stream_t stream);
int mifdtd_cpml_face_y_h(int nx, int ny, int nz_pitch, int pad, int npml, int backward,
                         const float *ex, const float *ez, float *hx, float *hz,
                         float *psi1, float *psi2, const float *rcmb, const float *rcma,
                         mifdtd_stream_t stream);
int mifdtd_cpml_face_z_e(int nx, int ny, int nz_pitch, int pad, int npml, int backward,
                         float *ex, float *ey, const float *hx, const float *hy,
                         const float *cex, const float *cey, float *psi1, float *psi2,
                         const float *rcmb, const float *rcma, mifdtd_stream_t stream);
int mifdtd_cpml_face_z_h(int nx, int ny, int nz_pitch, int pad, int npml, int backward,
                         const float *ex, const float *ey, float *hx, float *hy,
                         float *psi1, float *psi2, const float *rcmb, const float *rcma,
                         mifdtd_stream_t stream);

/* All CPML faces of one half step, in the reference order (x f, x b, y f, y b, z f,
 * z b): one launch per axis.  psi1[2*axis+side], psi2[2*axis+side] may be NULL for
 * a face that is off.  ch arrays (H half) may be NULL = 0.5. */
typedef struct {
    int npml;
    const float *b, *a;          /* device tables, 2*(npml+1) floats */
    float *psi1[6], *psi2[6];    /* index 2*axis + side */
} mifdtd_cpml_half;

int mifdtd_cpml_e(int nx, int ny, int nz_pitch, int pad, const mifdtd_cpml_half *cp,
                  float *ex, float *ey, float *ez,
                  const float *hx, const float *hy, const float *hz,
                  const float *cex, const float *cey, const float *cez,
                  mifdtd_stream_t stream);
int mifdtd_cpml_h(int nx, int ny, int nz_pitch, int pad, const mifdtd_cpml_half *cp,
                  const float *ex, const float *ey, const float *ez,
                  float *hx, float *hy, float *hz,
                  const float *chx, const float *chy, const float *chz,
                  mifdtd_stream_t stream);

/* ---------------------------------------------------------------- sources ----
 * Replaces the IncidentDirect -> SetFields update of K7/cpu/incident_direct.py:75-77
 * and K7/cpu/get_set_fields.py:110-123 (and KU/gpu/core/source.cu:11-20):
 *   f[i0..i1, j0..j1, k0..k1] (+)= tval * (sv ? sv[box index] : svs)
 * tval is the host's float32 cast of tfunc(tstep); sv is a C-order device array of
 * the box (inclusive bounds) or NULL. */
int mifdtd_source(float *f, int nx, int ny, int nz_pitch,
                  int i0, int i1, int j0, int j1, int k0, int k1,
                  float tval, const float *sv, float svs, int overwrite,
                  mifdtd_stream_t stream);

/* Periodic boundary: replaces Pbc.update of K7/cpu/pbc.py:25-31 (the slice table of
 * K7/util/common_exchange.py:1-29; K7/gpu/pbc.py is the same copy on device).  For axis
 * 0, 1, 2 (x, y, z), f1 and f2 are the two components tangential to it (x: ey, ez or
 * hy, hz; y: ex, ez or hx, hz; z: ex, ey or hx, hy).  e_half != 0 (after the E half
 * step): plane n-1 takes plane 0; e_half == 0 (after the H half step): plane 0 takes
 * plane n-1.  n is nx, ny or the real nz (nz_pitch - pad); x and y planes include the
 * pad.  Axes are applied one call each, in the caller's order. */
int mifdtd_pbc(float *f1, float *f2, int nx, int ny, int nz_pitch, int pad, int axis, int e_half,
               mifdtd_stream_t stream);

/* Drude dispersion (auxiliary current): replaces drude_ade.update_e of
 * kuFDTD/v0.2.0/dim3/cpu/core/drude_ade.c:14-99 (called after the core E update,
 * KU/dim3/cpu/matter.py:210-221).  On the region [i0, i0+li) x [j0, j0+lj) x
 * [k0, k0+lk), per component c: e' = e - ce J, J' = cja J + cjb e'.  j, cja, cjb are
 * region-sized C-order arrays; ce may be NULL (or hold NULLs) for the constant 0.5. */
int mifdtd_drude_e(float *const e[3], const float *const ce[3], float *const j[3], const float *const cja[3],
                   const float *const cjb[3], int nx, int ny, int nz_pitch, int i0, int j0, int k0, int li, int lj,
                   int lk, mifdtd_stream_t stream);

/* TFSF plane wave, normal incidence: replaces Tfsf.updateE / updateH of
 * kuFDTD/v0.2.0/dim3/tfsf.py:229-242 (Fdtd1d :14-40, faces and incident coefficients
 * :138-226).  mifdtd_tfsf_line advances the 1-D incident line of n points: e_half != 0:
 * E[1:n-1] -= 0.5 (H[2:] - H[1:n-1]), two-step-delay ends (abc[0..1] front, abc[2..3]
 * back, device floats), then E[3] = src; e_half == 0: H[1:] -= 0.5 (E[1:] - E[:-1]).
 * mifdtd_tfsf_faces adds, for every patch in table order, 0.5 * (coef * line[base +
 * offset along the propagation axis]) to the patch box of one component (f: ex, ey, ez,
 * hx, hy, hz); a cell on a box edge takes the corrections of both faces, in that order. */
#define MIFDTD_TFSF_MAX_PATCH 12
typedef struct {
    int n;                                 /* patches */
    int axis;                              /* propagation axis 0, 1, 2 */
    int comp[MIFDTD_TFSF_MAX_PATCH];       /* 0..5 = ex..hz */
    int box[MIFDTD_TFSF_MAX_PATCH][6];     /* i0, i1, j0, j1, k0, k1 (inclusive) */
    int base[MIFDTD_TFSF_MAX_PATCH];       /* line index at the box start along the axis */
    float coef[MIFDTD_TFSF_MAX_PATCH];
    unsigned start[MIFDTD_TFSF_MAX_PATCH + 1];  /* filled by the call */
} mifdtd_tfsf_patches;
int mifdtd_tfsf_line(float *e, float *h, float *abc, int n, int e_half, float src, mifdtd_stream_t stream);
int mifdtd_tfsf_faces(const mifdtd_tfsf_patches *p, float *const f[6], const float *line, int line_n, int nx, int ny,
                      int nz_pitch, mifdtd_stream_t stream);

/* Probe / GetFields gather (K7/cpu/get_set_fields.py:47-50, K7/gpu/src/copy.cl:3-14):
 *   out[n] = f[offsets[n]]  for n < count   (offsets: device int64 element offsets) */
int mifdtd_gather(const float *f, const long long *offsets, int count, float *out,
                  mifdtd_stream_t stream);


/* ------------------------------------------------------------- fused step ----
 * The whole time step (core E, CPML-E on all faces, E sources, core H, CPML-H, H
 * sources, probes) as one single-pass sweep over the interior box of cells that carry
 * no CPML term (each array streamed once; a work-group of rows_per_thread+1 waves owns
 * that many y rows of an x chunk and recomputes E of the row below;
 * rows_per_thread * planes_per_chunk bounds its LDS), between an E pass and an H pass
 * over the CPML shell around it (three launches per step with CPML, one without):
 * state of step n is read from e[] / psiE*[], step n+1 is written to e_alt[] /
 * psiE*_alt[] (every cell, walls copied), and the two sets swap every step.  psiH is
 * updated in place.  Same per-cell arithmetic as update_e/update_h + the CPML and
 * source calls above (values bit-identical; masked zero lanes may differ in sign).
 * No reference counterpart: it fuses the reference step loop of
 * KU/cpu/test/020-dielectric-cpml.py:58-73 (E, CPML-E, source, H, CPML-H).
 * Limits: CPML on all six faces or none, CH = 0.5 (no CH arrays),
 * at most MIFDTD_FUSED_MAX_SRC sources and MIFDTD_FUSED_MAX_PROBE probes.
 * Components: 0..5 = ex, ey, ez, hx, hy, hz.  Face index f = 2*axis + side.
 * MIFDTD_PSI_PACKED layout of psiE / psiE_alt / psiH (floats): for f = 0..5, psi1 then
 * psi2 of face f, each face_size(f) long: x faces npml*ny*nz_pitch, y faces
 * nx*npml*nz_pitch, z faces nx*ny*npml (z last so x/y faces stay 16-byte aligned);
 * 4*(sX+sY+sZ) floats per half.  The per-face psi arrays of mifdtd_cpml_e/h may be
 * views into such a buffer. */
#define MIFDTD_FUSED_MAX_SRC 4
#define MIFDTD_FUSED_MAX_PROBE 8
typedef struct {
    int nx, ny, nz_pitch, pad;
    float *e[6];          /* ex, ey, ez, hx, hy, hz at step n */
    float *e_alt[6];      /* the other buffer set */
    const float *ce[3];   /* cex, cey, cez or all NULL (0.5) */
    int npml;             /* 0: no CPML; else all six faces */
    const float *bE, *aE, *bH, *aH;
    float *psiE, *psiE_alt, *psiH;  /* packed psi buffers (MIFDTD_PSI_PACKED layout) */
    int nsrc;
    int src_comp[MIFDTD_FUSED_MAX_SRC];
    int src_box[MIFDTD_FUSED_MAX_SRC][6];     /* i0, i1, j0, j1, k0, k1 (inclusive) */
    const float *src_sv[MIFDTD_FUSED_MAX_SRC]; /* C-order box array or NULL */
    float src_svs[MIFDTD_FUSED_MAX_SRC];
    int src_overwrite[MIFDTD_FUSED_MAX_SRC];
    int nprobe;
    int probe_comp[MIFDTD_FUSED_MAX_PROBE];
    int probe_ijk[MIFDTD_FUSED_MAX_PROBE][3];
    int rows_per_thread;  /* fused: y rows owned by one work-group, 1..15 (T+1 waves) */
    int planes_per_chunk; /* fused: x planes marched by one work-group */
    /* CE as uint8 material indices (nx, ny, nz_pitch) per component into
     * ce_lut[3][256] (the float CE values), read by the fused sweep, the shell E pass
     * and the slab E pass; when ce_mat[0] is set, ce[] is ignored.  Index arrays must
     * be 4-byte aligned. */
    const unsigned char *ce_mat[3];
    const float *ce_lut;
    /* fused only, with CPML: a third buffer set (ex..hz).  When set, steps advance in
     * pairs: the interior sweep forms two steps per pass (the first step's interior
     * state stays on chip) over a box two cells further from the CPML slabs, and the
     * shell passes write the first step's shell values here; a trailing odd step is an
     * ordinary one.  rows_per_thread must then be 1..9 (rows+3 waves per work-group).
     * NULL: one step per sweep. */
    float *e_mid[6];
    /* fused only: one Drude box (auxiliary current, kuFDTD matter.py:104-221 with
     * drude_ade.c: after the core E update, before CPML-E and the sources, e' = e - CE J,
     * J' = CJA J + CJB e'), or drude_j[0] == NULL.  drude_box: i0, i1, j0, j1, k0, k1
     * (inclusive, inside the grid); drude_a / drude_b: CJA / CJB per component as C-order
     * box arrays; drude_j: J of step n, drude_j_alt: receives step n+1 (the two swap with
     * the field sets: *parity 1 means the current J is in drude_j_alt).  Arrays 16-byte
     * aligned.  Not with e_mid. */
    int drude_box[6];
    const float *drude_a[3], *drude_b[3];
    float *drude_j[3], *drude_j_alt[3];
    /* fused only: a TFSF plane wave (kuFDTD tfsf.py:90-242, normal incidence), or
     * tfsf_line_n == 0.  Each step: the 1-D line's E half (mifdtd_tfsf_line with the hard
     * source tfsf_src[s], a HOST array of nsteps values), the step with the E patches
     * (tfsf_e, reading the line's H) and the H patches (tfsf_h, reading the line's E)
     * applied after the sources, then the line's H half.  Patch tables as for
     * mifdtd_tfsf_faces (start[] unused; E patches of components 0..2, H of 3..5).  No
     * source may share a cell and component with a patch.  Not with e_mid. */
    mifdtd_tfsf_patches tfsf_e, tfsf_h;
    float *tfsf_line_e, *tfsf_line_h, *tfsf_abc;
    int tfsf_line_n;
    const float *tfsf_src;
    /* fused only: CPML on a subset of axes -- bit a set: both faces of axis a (x = 1,
     * y = 2, z = 4); 0 with npml > 0 means all three -- and periodic boundaries
     * (K7/cpu/pbc.py:5-39): pbc_axes[0..npbc) in the order the copies run, after each
     * step's E (plane n-1 takes plane 0) and H (plane 0 takes plane n-1) as
     * mifdtd_pbc does.  An axis is CPML or periodic, not both; probes are recorded
     * before the copies, so none may sit on a copied cell. */
    int cpml_axes;
    int npbc;
    int pbc_axes[3];
    /* fused only, optional with ce_mat: CE as one uint16 index per cell (an array like a
     * field, 8-byte aligned) into a DEVICE table of ce_tri_n <= MIFDTD_FUSED_MAX_CE_TRI
     * (cex, cey, cez) float triples -- the same values as ce_mat/ce_lut, 2 B per cell
     * instead of 3.  Read by the interior sweep of single steps; the shell passes keep
     * ce_mat. */
    const unsigned short *ce_tri;
    const float *ce_tri_lut;
    int ce_tri_n;
} mifdtd_fused_desc;
#define MIFDTD_FUSED_MAX_CE_TRI 4096

/* Run nsteps fused steps.  tvals: HOST array (nsteps, nsrc) of the float32 time values
 * of each source; probe_series: DEVICE array (nsteps, nprobe).  On return *parity is 1
 * when the current state is in e_alt[] / psiE_alt (nsteps & 1 with one step per sweep;
 * with e_mid[] set, step pairs and a trailing odd step each swap the two sets once). */
int mifdtd_fused_run(const mifdtd_fused_desc *d, int nsteps, const float *tvals,
                     float *probe_series, int *parity, mifdtd_stream_t stream);
/* The same, with HIP events around each interior-sweep launch: *sweep_ms receives the
 * sweep time per step (the sum of the sweep launches' durations / nsteps; waits for the
 * stream to reach the last sweep).  No reference counterpart (measurement). */
int mifdtd_fused_run_timed(const mifdtd_fused_desc *d, int nsteps, const float *tvals,
                           float *probe_series, int *parity, mifdtd_stream_t stream,
                           float *sweep_ms);

/* ------------------------------------------------------------- slab step ----
 * The same whole step, in place, as two launches per slab of slab_planes consecutive
 * x planes: E pass (core E + CPML-E + E sources + E probes) then H pass (core H +
 * CPML-H + H sources + H probes), slabs in increasing x.  A slab small enough for its
 * E, H and CE planes to stay in the last-level cache between its two passes makes
 * HBM see each array about once per step.  Results equal update_e/update_h + the CPML
 * and source calls above bit for bit (masked lanes included).  Uses e[], ce[], npml,
 * the CPML tables, psiE and psiH (both updated in place), sources and probes of the
 * descriptor; e_alt, psiE_alt, rows_per_thread and planes_per_chunk are ignored.
 * No reference counterpart (it reorders the step loop of
 * KU/cpu/test/020-dielectric-cpml.py:58-73 by slab).  tvals: HOST (nsteps, nsrc);
 * probe_series: DEVICE (nsteps, nprobe). */
int mifdtd_slab_run(const mifdtd_fused_desc *d, int slab_planes, int nsteps,
                    const float *tvals, float *probe_series, mifdtd_stream_t stream);

/* ---------------------------------------------------------------- batched ----
 * Many independent Yee (+CPML) simulations of uneven sizes advanced together: every
 * step is two launches for the whole batch (E pass of all instances, then H pass), each
 * block of a launch belonging to one instance.  Per instance the per-cell arithmetic is
 * that of update_e/update_h + CPML + sources (as mifdtd_slab_run).  No reference
 * counterpart: the reference runs one simulation per process (K7/main.py:296-326).
 * Each instance: KEMP-layout fields (own nx, ny, nz_pitch, pad), CE arrays on every
 * instance or on none, npml 0 or the batch's npml (tables shared), packed psi buffers
 * (MIFDTD_PSI_PACKED, updated in place), up to MIFDTD_FUSED_MAX_SRC sources and
 * MIFDTD_FUSED_MAX_PROBE probes.  Source n of instance m reads column
 * (sources of instances < m) + n of the tvals row of the step; probes likewise fill the
 * columns of a probe_series row. */
typedef struct {
    int nx, ny, nz_pitch, pad;
    float *e[6];
    const float *ce[3];
    int npml;
    float *psiE, *psiH;
    int nsrc;
    int src_comp[MIFDTD_FUSED_MAX_SRC];
    int src_box[MIFDTD_FUSED_MAX_SRC][6];
    const float *src_sv[MIFDTD_FUSED_MAX_SRC];
    float src_svs[MIFDTD_FUSED_MAX_SRC];
    int src_overwrite[MIFDTD_FUSED_MAX_SRC];
    int nprobe;
    int probe_comp[MIFDTD_FUSED_MAX_PROBE];
    int probe_ijk[MIFDTD_FUSED_MAX_PROBE][3];
    /* optional: CE as uint8 material indices (each array like a field, 4-byte aligned)
     * into a DEVICE table of 3 x 256 floats (rows cex, cey, cez) of the instance, read
     * instead of ce[]; set on every instance or on none */
    const unsigned char *ce_mat[3];
    const float *ce_lut;
} mifdtd_instance;

typedef struct mifdtd_batch mifdtd_batch; /* opaque; holds device copies of the descriptors */

/* Validate the instances and upload their descriptors (allocates device memory). */
int mifdtd_batch_create(const mifdtd_instance *inst, int ninst,
                        const float *bE, const float *aE, const float *bH, const float *aH,
                        mifdtd_batch **out);
/* Columns of a tvals row and of a probe row; blocks of one launch. */
int mifdtd_batch_sizes(const mifdtd_batch *b, int *total_src, int *total_probe,
                       unsigned *total_blocks);
/* Run nsteps steps.  tvals: DEVICE (nsteps, total_src) float32 source values;
 * probe_series: DEVICE (nsteps, total_probe).  use_graph != 0 captures the run's
 * launches into one hipGraph and launches it (needs a non-null stream; the call then
 * returns after the graph completed).  remap_rows > 0 deals each instance's blocks to
 * the 8 XCDs in runs of remap_rows consecutive blocks. */
int mifdtd_batch_run(mifdtd_batch *b, int nsteps, const float *tvals, float *probe_series,
                     int use_graph, int remap_rows, mifdtd_stream_t stream);
int mifdtd_batch_destroy(mifdtd_batch *b);

/* The batched run without any device allocation: the same validation, kernels and
 * results as mifdtd_batch_create + mifdtd_batch_run (use_graph 0) + mifdtd_batch_destroy,
 * with the instance descriptors held in a static device array of the library instead of
 * memory from hipMalloc.  No reference counterpart: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  At most MIFDTD_BATCH_STATIC_MAX_INST instances per call (rejected with
 * MIFDTD_ERANGE beyond).  Each GPU holds one such array until the process ends; every
 * call on a GPU rewrites it (nothing has to start at zero), so calls that overlap on one
 * GPU (other streams, other host threads) race.  tvals: DEVICE (nsteps, sum of the
 * instances' nsrc); probe_series: DEVICE (nsteps, sum of nprobe); columns as for
 * mifdtd_batch_run. */
#define MIFDTD_BATCH_STATIC_MAX_INST 128
int mifdtd_batch_run_static(const mifdtd_instance *inst, int ninst,
                            const float *bE, const float *aE, const float *bH, const float *aH,
                            int nsteps, const float *tvals, float *probe_series, int remap_rows,
                            mifdtd_stream_t stream);

/* ---------------------------------------------------------------- adjoint ----
 * Discrete adjoint of one in-place Yee + CPML step (update_e, mifdtd_cpml_e, sources,
 * update_h, mifdtd_cpml_h with CH = 0.5 and the packed psi layout), for gradients with
 * respect to the CE coefficients of a design box.  No reference counterpart (the
 * reference has no gradient path).  lam[0..5]: adjoint fields (same layout as the
 * fields; on entry the adjoint of the state after step n, on return that of the state
 * before it); lpsiE / lpsiH: adjoint psi (packed layout).  eplane (or NULL): the
 * forward E^n on the monitor plane x = iplane, (3, ny, nz_pitch), whose |E|^2 the
 * objective sums at this step (adds 2*E^n to the adjoint).  curl: (3, di, dj, dk) curl
 * of H^{n-1} on the box as update_e forms it (mifdtd_design_curl, taken before step n);
 * grad (or NULL): (3, di, dj, dk) double accumulator of dJ/dCE_c.
 * lpsiE is carried one step behind: a call leaves the b_E scaling of its psi-E adjoint
 * to the next call (two launches per step), so after the last step lpsiE is not yet
 * the adjoint of the initial psi-E (the gradient and lE, lH, lpsiH are complete). */
typedef struct {
    int i0, j0, k0, di, dj, dk;
} mifdtd_design_box;

int mifdtd_adjoint_step(int nx, int ny, int nz_pitch, int pad,
                        float *const lam[6], const float *const ce[3], int npml,
                        float *lpsiE, float *lpsiH,
                        const float *bE, const float *aE, const float *bH, const float *aH,
                        const float *eplane, int iplane, const mifdtd_design_box *box,
                        const float *curl, double *grad, mifdtd_stream_t stream);
/* out (3, di, dj, dk): (hz[j+1]-hz)-(hy[k+1]-hy), (hx[k+1]-hx)-(hz[i+1]-hz),
 * (hy[i+1]-hy)-(hx[j+1]-hx) on the box (0 where update_e does not write). */
int mifdtd_design_curl(int nx, int ny, int nz_pitch, int pad,
                       const float *hx, const float *hy, const float *hz,
                       const mifdtd_design_box *box, float *out, mifdtd_stream_t stream);

/* ------------------------------------------------------- 4th-order symplectic ----
 * Layout of the reference's 4th-order path: (nx, ny, nz) Fortran order (x contiguous,
 * no padding), S4/PyCUDA/100-base.py:74-89; nx must be a multiple of 4.
 * s4_update_h / s4_update_e replace the Python-C extension functions
 * S4/Python-C/dielectric.c:5 (update_h, "iiifOOOOOOOOO": nx, ny, nz, cl, Ex..Hz, Cx..Cz)
 * and :52 (update_e, same with dl), and the CUDA kernels S4/PyCUDA/100-base.py:10,28:
 *   hx -= cl*chx*(27*(ez-ez[j-1]-ey+ey[k-1]) - (ez[j+1]-ez[j-2]-ey[k+1]+ey[k-2]))   2<=j<=ny-2, 2<=k<=nz-2
 *   hy -= cl*chy*(27*(ex-ex[k-1]-ez+ez[i-1]) - (ex[k+1]-ex[k-2]-ez[i+1]+ez[i-2]))   2<=i<=nx-2, 2<=k<=nz-2
 *   hz -= cl*chz*(27*(ey-ey[i-1]-ex+ex[j-1]) - (ey[i+1]-ey[i-2]-ex[j+1]+ex[j-2]))   2<=i<=nx-2, 2<=j<=ny-2
 *   ex += dl*cex*(27*(hz[j+1]-hz-hy[k+1]+hy) - (hz[j+2]-hz[j-1]-hy[k+2]+hy[k-1]))   1<=j<=ny-3, 1<=k<=nz-3
 *   ey += dl*cey*(27*(hx[k+1]-hx-hz[i+1]+hz) - (hx[k+2]-hx[k-1]-hz[i+2]+hz[i-1]))   1<=i<=nx-3, 1<=k<=nz-3
 *   ez += dl*cez*(27*(hy[i+1]-hy-hx[j+1]+hx) - (hy[i+2]-hy[i-1]-hx[j+2]+hx[j-1]))   1<=i<=nx-3, 1<=j<=ny-3
 * restricted to i < nx-1, j < ny-1, k < nz-1 (the loop bounds of dielectric.c:30-32).
 * The coefficient arrays hold S/(24*eps_r) (S/24 in vacuum, 100-base.py:75). */
void s4_update_h(int nx, int ny, int nz, float cl, float *ex, float *ey, float *ez,
                 float *hx, float *hy, float *hz, float *cx, float *cy, float *cz);
void s4_update_e(int nx, int ny, int nz, float dl, float *ex, float *ey, float *ez,
                 float *hx, float *hy, float *hz, float *cx, float *cy, float *cz);

/* Status-returning, stream-taking forms; NULL coefficient arrays mean the constant
 * ch0 / ce0 (e.g. S/24 for mu_r = 1). */
int mifdtd_s4_update_h(int nx, int ny, int nz, float cl,
                       const float *ex, const float *ey, const float *ez,
                       float *hx, float *hy, float *hz,
                       const float *chx, const float *chy, const float *chz, float ch0,
                       mifdtd_stream_t stream);
int mifdtd_s4_update_e(int nx, int ny, int nz, float dl,
                       float *ex, float *ey, float *ez,
                       const float *hx, const float *hy, const float *hz,
                       const float *cex, const float *cey, const float *cez, float ce0,
                       mifdtd_stream_t stream);

/* mifdtd_s4_update_e with CE given as per-component uint8 material indices (Fortran
 * (nx, ny, nz), 4-byte aligned) into lut (DEVICE, 3 x 256 float32: row c holds the
 * distinct CE values of component c): the same results, 39 instead of 48 bytes per cell.
 * No reference counterpart (a storage form of the CE arrays of S4/PyCUDA/100-base.py:84-86
 * for media with at most 256 distinct values per component). */
int mifdtd_s4_update_e_lut(int nx, int ny, int nz, float dl,
                           float *ex, float *ey, float *ez,
                           const float *hx, const float *hy, const float *hz,
                           const unsigned char *mat_x, const unsigned char *mat_y,
                           const unsigned char *mat_z, const float *lut, mifdtd_stream_t stream);

/* CPML for the 4th-order path (no reference counterpart: the reference's 4th-order
 * path has no absorbing boundary).  Face, table and sign conventions of the Yee CPML
 * above (kuFDTD), driven by the 4th-order difference of the stage update:
 *   E stage: psi = b*psi + a*(27*(H[+1]-H) - (H[+2]-H[-1]));   E_1 -= dl*CE*psi1, E_2 += dl*CE*psi2
 *   H stage: psi = b*psi + a*(27*(E-E[-1]) - (E[+1]-E[-2]));   H_1 += cl*CH*psi1, H_2 -= cl*CH*psi2
 * b/a are the tables of that stage's time step (b = exp(-sigma*w*dt), a = b-1 for the
 * stage weight w = dl or cl; negative weights give b > 1 and the product over a step
 * is exp(-sigma*dt)).  psi layouts, x fastest: x face (npml, ny, nz), y face
 * (nx, npml, nz), z face (nx, ny, npml).  Applied only where the stage update writes. */
int mifdtd_s4_cpml_e(int nx, int ny, int nz, const mifdtd_cpml_half *cp, float dl,
                     float *ex, float *ey, float *ez,
                     const float *hx, const float *hy, const float *hz,
                     const float *cex, const float *cey, const float *cez, float ce0,
                     mifdtd_stream_t stream);
int mifdtd_s4_cpml_h(int nx, int ny, int nz, const mifdtd_cpml_half *cp, float cl,
                     const float *ex, const float *ey, const float *ez,
                     float *hx, float *hy, float *hz,
                     const float *chx, const float *chy, const float *chz, float ch0,
                     mifdtd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MIFDTD_H */
