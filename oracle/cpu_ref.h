/*
 * cpu_ref.h -- TEST INFRASTRUCTURE ONLY. C-ABI of the CPU restatement of
 * SoftGLRender's RendererSoft triangle path (oracle/cpu_ref.cpp).
 */
#ifndef CPU_REF_H
#define CPU_REF_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replays an rhip_stream.h stream (setup once, frame part `repeats` times, the
 * RHS_READ commands once at the end). Returns output bytes, -1 on a malformed
 * stream or unsupported command, -2 when out_cap is too small. */
long long cpu_ref_replay(const void *stream, size_t bytes, void *out, size_t out_cap, int repeats, double *frame_ms);

/* 0: powf / log2f / exp2f from the host libm, as the reference calls them.
 * 1: the same functions correctly rounded (double precision, one rounding). */
void cpu_ref_set_math(int correctly_rounded);

/* Worker threads for the fragment stage (1 = serial). Results do not depend on it. */
void cpu_ref_set_threads(int n);

/* Counters of the last replayed frame: [0] fragment-shader invocations,
 * [1] samples written, [2] triangles rasterised, [3] triangles that needed clipping,
 * [4] triangles culled, [5] 2x2 quads visited, [6] texel fetches, [7] draws. */
void cpu_ref_stats(long long *out8);

#ifdef __cplusplus
}
#endif
#endif
