/*
 * One KEMP core program: the shared library that K7/cpu/core.py:52-85 generates per
 * (precision_float, coeff_use, pad) from K7/cpu/src/core.c and loads with
 * np.ctypeslib.load_library, here as a thin C shim over libmifdtd.so's HIP kernels.
 *
 * Compiled once per variant (fdtd_accelerate_amd/build.py) with
 *   -DKEMP_DOUBLE=0|1   'single' (float, 4-lane SSE, pad 0..3) or 'double' (2 lanes, pad 0..1)
 *   -DKEMP_CE=0|1       coeff_use has 'e': update_e takes cex, cey, cez (9 arrays, else 6)
 *   -DKEMP_CH=0|1       coeff_use has 'h': update_h takes chx, chy, chz (9 arrays, else 6)
 *   -DKEMP_PAD=p        the pad baked into the program's masks (K7/cpu/core.py:33-49)
 *   -DKEMP_SPLIT=0|1    the split form of K7/cpu/src/core_split.c:11,59 (core_split.py):
 *                       update_e / update_h take idx0, nmax after nz and update the
 *                       flat element range [idx0, nmax) only
 * Exports exactly the reference's two symbols with the reference's arity and meaning
 * (K7/cpu/src/core.c:11,59; argtypes K7/cpu/core.py:89-105): nz is nz_pitch, arrays
 * are DEVICE pointers, the coefficient-free forms use the constant 0.5.  Work goes to
 * the null stream; failures are recorded for mifdtd_last_error() (the reference's
 * calls have no error channel).
 */
#include "mifdtd.h"

#ifndef KEMP_SPLIT
#define KEMP_SPLIT 0
#endif

#if KEMP_DOUBLE
#define KEMP_T double
#define KEMP_E mifdtd_update_e_f64
#define KEMP_H mifdtd_update_h_f64
#define KEMP_E_RANGE mifdtd_update_e_f64_range
#define KEMP_H_RANGE mifdtd_update_h_f64_range
#else
#define KEMP_T float
#define KEMP_E mifdtd_update_e
#define KEMP_H mifdtd_update_h
#define KEMP_E_RANGE mifdtd_update_e_range
#define KEMP_H_RANGE mifdtd_update_h_range
#endif

#if KEMP_CE
#define KEMP_CE_ARGS , KEMP_T *cex, KEMP_T *cey, KEMP_T *cez
#define KEMP_CE_PASS cex, cey, cez
#else
#define KEMP_CE_ARGS
#define KEMP_CE_PASS 0, 0, 0
#endif

#if KEMP_CH
#define KEMP_CH_ARGS , KEMP_T *chx, KEMP_T *chy, KEMP_T *chz
#define KEMP_CH_PASS chx, chy, chz
#else
#define KEMP_CH_ARGS
#define KEMP_CH_PASS 0, 0, 0
#endif

#if KEMP_SPLIT
__attribute__((visibility("default"))) void update_e(int nx, int ny, int nz, int idx0, int nmax, KEMP_T *ex, KEMP_T *ey,
                                                     KEMP_T *ez, KEMP_T *hx, KEMP_T *hy, KEMP_T *hz KEMP_CE_ARGS)
{
    (void)KEMP_E_RANGE(nx, ny, nz, KEMP_PAD, idx0, nmax, ex, ey, ez, hx, hy, hz, KEMP_CE_PASS, 0);
}

__attribute__((visibility("default"))) void update_h(int nx, int ny, int nz, int idx0, int nmax, KEMP_T *ex, KEMP_T *ey,
                                                     KEMP_T *ez, KEMP_T *hx, KEMP_T *hy, KEMP_T *hz KEMP_CH_ARGS)
{
    (void)KEMP_H_RANGE(nx, ny, nz, KEMP_PAD, idx0, nmax, ex, ey, ez, hx, hy, hz, KEMP_CH_PASS, 0);
}
#else
__attribute__((visibility("default"))) void update_e(int nx, int ny, int nz, KEMP_T *ex, KEMP_T *ey, KEMP_T *ez,
                                                     KEMP_T *hx, KEMP_T *hy, KEMP_T *hz KEMP_CE_ARGS)
{
    (void)KEMP_E(nx, ny, nz, KEMP_PAD, ex, ey, ez, hx, hy, hz, KEMP_CE_PASS, 0);
}

__attribute__((visibility("default"))) void update_h(int nx, int ny, int nz, KEMP_T *ex, KEMP_T *ey, KEMP_T *ez,
                                                     KEMP_T *hx, KEMP_T *hy, KEMP_T *hz KEMP_CH_ARGS)
{
    (void)KEMP_H(nx, ny, nz, KEMP_PAD, ex, ey, ez, hx, hy, hz, KEMP_CH_PASS, 0);
}
#endif

/* Which program this is: (split, double, ce, ch, pad), for loaders that check what they bound. */
__attribute__((visibility("default"))) int mifdtd_kemp_program_id(void)
{
    return (KEMP_SPLIT << 16) | (KEMP_DOUBLE << 12) | (KEMP_CE << 8) | (KEMP_CH << 4) | KEMP_PAD;
}
