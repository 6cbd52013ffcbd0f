/* Shared index helpers for the oracle sources (test infrastructure only). */
#ifndef CGLI_ORACLE_COMMON_H
#define CGLI_ORACLE_COMMON_H

#include <math.h>
#include "cgli_oracle.h"

/* Reference index names (Resolution.h:15-29): NY = ny_ref = Y_out+1 etc. */
#define DIMS(c)                                  \
    const long Ym = (c)->ny + 4, Zm = (c)->nz + 4; \
    const int gend = (c)->nx + 2;                \
    const int NY = (c)->ny + 1, NZ = (c)->nz + 1; \
    const int NYY = NY + 1, NZZ = NZ + 1;        \
    const int NYYY = NYY + 1, NZZZ = NZZ + 1;    \
    (void)gend; (void)NY; (void)NZ; (void)NYY; (void)NZZ; (void)NYYY; (void)NZZZ

#define ID(i, j, k) ((((long)(i)) * Ym + (long)(j)) * Zm + (long)(k))

/* max() of the reference (Flux_X.cpp:14): evaluates the winner twice,
 * which is the same value. */
#define OMAX(a, b) (((a) > (b)) ? (a) : (b))

#endif
