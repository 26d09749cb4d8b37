/*
 * sfx_profile.h -- kernel timing of the propagator sweeps (diagnostic
 * interface). It has no counterpart in the reference, which times nothing on
 * the GPU (only clock() ticks in sf_vectoriterations.cpp:7-9): it is for
 * callers that report achieved bandwidth, such as benchmark drivers.
 *
 * Between sfx_profile_start and sfx_profile_stop every propagator kernel of
 * the system is bracketed by HIP events on the stream it runs on. A launch
 * advances one rank, or several (the blocked 3-D simple-cubic kernels). The
 * report sums the event durations and ranks per kernel kind and precision and
 * gives the algorithmic bytes of an average launch: ranks per launch x sites
 * (interior) x the words of one site-rank:
 *                                                   first order  second order
 *   SFX_K_FWD   G_s = g A[G_{s-1}]                      3            13
 *   SFX_K_BWD   G_inv,s = g A[G_inv,s+1], phi += G G_inv 6            21
 *   SFX_K_TFWD  base + tangent forward step              6            26
 *   SFX_K_TBWD  base + tangent backward step + compose   10           46
 * (a word is 8 bytes for FP64 and 4 bytes for FP32; second order moves the
 * six direction components of every propagator, one word for g and v and
 * two for the accumulator).
 */
#ifndef SFX_PROFILE_H
#define SFX_PROFILE_H

#include "sfx.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SFX_K_FWD = 0, SFX_K_BWD = 1, SFX_K_TFWD = 2, SFX_K_TBWD = 3, SFX_K_COUNT = 4 };

typedef struct sfx_profile_entry {
  int kind;                 /* SFX_K_* */
  int precision;            /* SFX_FP64 / SFX_FP32 */
  long launches;
  long ranks;               /* ranks those launches advanced */
  double total_ms;          /* summed event durations */
  double bytes_per_launch;  /* algorithmic bytes of an average launch */
} sfx_profile_entry;

typedef struct sfx_profile_report {
  int count;
  sfx_profile_entry entries[2 * SFX_K_COUNT];
} sfx_profile_report;

int sfx_profile_start(sfx_system* s);
/* As sfx_profile_start, timing only the (kind, precision) pairs whose bit
 * 2*kind + precision is set in `select`; the other launches run without
 * events (every event pair adds a few microseconds of queue time). */
int sfx_profile_start_select(sfx_system* s, unsigned select);
int sfx_profile_stop(sfx_system* s, sfx_profile_report* report);

#ifdef __cplusplus
}
#endif
#endif /* SFX_PROFILE_H */
