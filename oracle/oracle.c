/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.  The
 * product (fdtd_accelerate_amd) never links, imports or calls it.
 *
 * CPU restatement of the reference FDTD hot path; the function bodies live in
 * oracle_body.h and are instantiated for float (_f32) and double (_f64).
 */
#include <stddef.h>
#include <xmmintrin.h>

#define REAL float
#define SFX(name) name##_f32
#include "oracle_body.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX(name) name##_f64
#include "oracle_body.h"
#undef REAL
#undef SFX

/* Flush-to-zero and denormals-are-zero on the calling thread's MXCSR (threads created
 * afterwards inherit it): the FTZ variant of the CPU baseline timing. */
void ora_set_ftz(int on)
{
    unsigned m = _mm_getcsr();
    m = on ? (m | 0x8040u) : (m & ~0x8040u);
    _mm_setcsr(m);
}
