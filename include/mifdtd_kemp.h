/*
 * mifdtd KEMP core programs: the two symbols of each program library
 * libkemp_<precision>_<coeff>_pad<p>.so (fdtd_accelerate_amd/kemp/).
 *
 * The reference builds one core program per (precision_float, coeff_use, pad) by text
 * substitution of K7/cpu/src/core.c (K7/cpu/core.py:33-85) and binds, per program,
 *   update_e(int nx, int ny, int nz, T *ex, T *ey, T *ez, T *hx, T *hy, T *hz [, T *cex, T *cey, T *cez])
 *   update_h(int nx, int ny, int nz, T *ex, T *ey, T *ez, T *hx, T *hy, T *hz [, T *chx, T *chy, T *chz])
 * (K7/cpu/src/core.c:11,59; argtypes K7/cpu/core.py:89-105: 3 c_int + 6 or 9 arrays,
 * restype None).  Each library here is the same program over the HIP kernels:
 *   precision  single (T = float, pad 0..3) or double (T = double, pad 0..1)
 *   coeff      none, e, h or eh: update_e takes the three CE arrays iff coeff has 'e',
 *              update_h the three CH arrays iff coeff has 'h' (else the constant 0.5)
 *   pad        nz_pitch - nz, fixed per program like the reference's masks
 * nz is nz_pitch; the arrays are DEVICE pointers (16-byte aligned, C order
 * (nx, ny, nz_pitch)); work goes to the null stream.  A rejected call does nothing
 * and leaves its reason in mifdtd_last_error() of libmifdtd.so.
 * The declarations below are those of the 'single', coeff 'eh' program; the others
 * drop the coefficient arrays or use double as stated.
 *
 * Split form, libkemp_<precision>_<coeff>_pad<p>_split.so: the program that
 * K7/cpu/core_split.py:8-98 builds from K7/cpu/src/core_split.c:11,59, whose two
 * symbols take idx0, nmax (c_int) after nz and update only the flat element range
 * [idx0, nmax) of the (nx, ny, nz_pitch) grid: a 4-float (2-double) vector is updated
 * when its first element lies in the range, so idx0 must start a vector and a range
 * outside the grid is rejected.  core_split.py uses the whole grid and its 'pre' /
 * 'post' plane ranges (E: [ny*nz, 3*ny*nz) / [0, ny*nz); H: [0, 2*ny*nz) /
 * [2*ny*nz, 3*ny*nz)).  Declared below under MIFDTD_KEMP_SPLIT, since the two forms
 * share their symbol names.
 */
#ifndef MIFDTD_KEMP_H
#define MIFDTD_KEMP_H

#ifdef __cplusplus
extern "C" {
#endif

#ifdef MIFDTD_KEMP_SPLIT
void update_e(int nx, int ny, int nz, int idx0, int nmax, float *ex, float *ey, float *ez,
              float *hx, float *hy, float *hz, float *cex, float *cey, float *cez);
void update_h(int nx, int ny, int nz, int idx0, int nmax, float *ex, float *ey, float *ez,
              float *hx, float *hy, float *hz, float *chx, float *chy, float *chz);
#else
void update_e(int nx, int ny, int nz, float *ex, float *ey, float *ez,
              float *hx, float *hy, float *hz, float *cex, float *cey, float *cez);
void update_h(int nx, int ny, int nz, float *ex, float *ey, float *ez,
              float *hx, float *hy, float *hz, float *chx, float *chy, float *chz);
#endif

/* (split << 16) | (double << 12) | (ce << 8) | (ch << 4) | pad of the loaded program.  No reference
 * counterpart (lets a loader check which program it bound). */
int mifdtd_kemp_program_id(void);

#ifdef __cplusplus
}
#endif
#endif /* MIFDTD_KEMP_H */
