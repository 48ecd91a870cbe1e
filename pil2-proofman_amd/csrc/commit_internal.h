// Interface between the host sources of libpil2_commit_hip (not exported in the
// public header).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

// LDE + Poseidon2 leaves + tree levels enqueued on st (no phase timing).
int64_t commit_enqueue(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *d_trace, uint64_t *d_scratch,
                       uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, hipStream_t st);

// Records msg as this thread's last error and returns code.
int64_t api_fail(int64_t code, const std::string &msg);
