/* Chunk and split-point arithmetic shared by the host plan (smoc_api.cpp)
 * and the device plan (smoc_kernels.hip).  Track t holds stream segments
 * [b, e).  Split points are the stream indices inside a track that are
 * multiples of SMOC_RUN_SPLIT: a run of a region-sorted stream ends before
 * each.  A chunked sweep cuts a track at base + k C (k >= 1, below e), where
 * base is b, or, for C a multiple of SMOC_RUN_SPLIT, b rounded up to a
 * multiple of it, so that every cut is a split point; the first chunk is
 * then up to SMOC_RUN_SPLIT - 1 segments longer than C.  A track of at most C
 * segments is one chunk. */
#ifndef SMOC_PLAN_H
#define SMOC_PLAN_H

#include <stdint.h>

#include "smoc_internal.h"

__host__ __device__ inline int64_t smoc_chunk_base(int64_t b, int64_t C)
{
    return C % SMOC_RUN_SPLIT == 0 ? (b + SMOC_RUN_SPLIT - 1) / SMOC_RUN_SPLIT * SMOC_RUN_SPLIT
                                   : b;
}

/* chunks of track [b, e) */
__host__ __device__ inline int64_t smoc_chunk_count(int64_t b, int64_t e, int64_t C)
{
    const int64_t B = smoc_chunk_base(b, C);
    return e - B > C ? (e - B + C - 1) / C : 1;
}

/* first segment of chunk k of track [b, e) */
__host__ __device__ inline int64_t smoc_chunk_begin(int64_t b, int64_t k, int64_t C)
{
    return k == 0 ? b : smoc_chunk_base(b, C) + k * C;
}

/* end (exclusive) of chunk k of track [b, e) */
__host__ __device__ inline int64_t smoc_chunk_end(int64_t b, int64_t e, int64_t k, int64_t C)
{
    const int64_t x = smoc_chunk_base(b, C) + (k + 1) * C;
    return x < e ? x : e;
}

/* split points of track [b, e): the multiples of SMOC_RUN_SPLIT in (b, e) */
__host__ __device__ inline int64_t smoc_split_count(int64_t b, int64_t e)
{
    return e > b ? (e - 1) / SMOC_RUN_SPLIT - b / SMOC_RUN_SPLIT : 0;
}

/* 1-based index, among the split points of the track that begins at b, of
 * split point x (a multiple of SMOC_RUN_SPLIT in (b, e)) */
__host__ __device__ inline int64_t smoc_split_index(int64_t b, int64_t x)
{
    return x / SMOC_RUN_SPLIT - b / SMOC_RUN_SPLIT;
}

#endif
