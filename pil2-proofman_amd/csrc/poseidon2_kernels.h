// Launch interface of the Poseidon2 Merkle kernels (internal to libpil2_commit_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define P2_THREADS 256
// tree levels with at most this many parents use 16 lanes per permutation (default)
#ifndef P2_COOP_MAX_PARENTS
#define P2_COOP_MAX_PARENTS 16384
#endif

// leaf hashing of at most this many rows uses 16 lanes per row
#ifndef P2_COOP_LEAF_MAX_ROWS
#define P2_COOP_LEAF_MAX_ROWS 16384
#endif

// Transcript sponge state (device memory), see k_transcript.
struct P2TranscriptState {
    uint64_t state[16], pending[12], out[16];
    uint32_t pending_cursor, out_cursor;
};

extern "C" int p2_launch_leaves(uint32_t width, const uint64_t *src, uint64_t rs, uint64_t cs, uint64_t ncols,
                                uint64_t nrows, uint64_t *out, hipStream_t st);
// carry != 0: continue the row sponges from the capacity words already in out
extern "C" int p2_launch_leaves_carry(uint32_t width, const uint64_t *src, uint64_t rs, uint64_t cs, uint64_t ncols,
                                      uint64_t nrows, uint64_t *out, uint32_t carry, hipStream_t st);
extern "C" int p2_launch_tree_levels(uint32_t width, uint64_t *tree, uint64_t nleaves, uint64_t coop_max,
                                     hipStream_t st);
extern "C" int p2_launch_permute(uint32_t width, uint64_t *states, uint64_t count, hipStream_t st);
extern "C" int p2_launch_mul_peak(const uint64_t *in, uint64_t *out, uint64_t lanes, uint32_t iters, hipStream_t st);
extern "C" int p2_launch_copy_root(const uint64_t *src, uint64_t *dst, hipStream_t st);
extern "C" int p2_launch_transcript(uint32_t width, P2TranscriptState *t, uint64_t *io, uint64_t n, int op,
                                    hipStream_t st);
extern "C" int p2_launch_grind(const uint64_t *ch, uint64_t level, uint64_t base, uint64_t count,
                               uint64_t *found, hipStream_t st);
extern "C" int p2_launch_group_proofs(const uint64_t *src, uint64_t rs, uint64_t cs, uint64_t ncols, const uint64_t *tree,
                                      uint64_t nrows, uint64_t arity, uint64_t cut, const uint64_t *idx,
                                      uint64_t nq, uint64_t pw, uint64_t *out, hipStream_t st);
