/* Internal host-side helpers shared by the C-ABI translation units. */
#ifndef INFOLI_HOST_H_
#define INFOLI_HOST_H_

#include <stdint.h>
#include "infoli_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Record a printf-style message for infoli_last_error(); returns `code`. */
int infoli_fail(int code, const char *fmt, ...)
#ifdef __GNUC__
    __attribute__((format(printf, 2, 3)))
#endif
    ;
void infoli_clear_error(void);

/* Counter-based random numbers: the same (seed, stream, a, b) gives the same
 * 64 bits on every host. */
uint64_t infoli_rng_bits(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b);
/* Uniform double in (0, 1] from 64 random bits. */
double infoli_rng_u01(uint64_t bits);

/* Stream identifiers of the synthetic inputs. */
enum {
    INFOLI_STREAM_GRAPH = 0, INFOLI_STREAM_W = 1, INFOLI_STREAM_GCAL = 2,
    INFOLI_STREAM_IAPP = 3, INFOLI_STREAM_MASK = 4, INFOLI_STREAM_SIZES = 5,
    INFOLI_STREAM_DIRECTED = 6
};

#ifdef __cplusplus
}
#endif
#endif
