// Launch interface of the FRI kernels (internal to libpil2_commit_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct NttTablesHost;
// tw (optional): the NTT's split root tables, from which a lane takes its coset
// point's root w_{2^prev}^-g with one product; without them it is a power.
extern "C" int fri_launch_fold(uint64_t *out, const uint64_t *pol, const uint64_t *challenge, uint64_t n_bits_ext,
                               uint64_t prev_bits, uint64_t cur_bits, const NttTablesHost *tw, hipStream_t st);
// As fri_launch_fold with the challenge read from device memory (d_challenge, 3
// words) by the kernel when d_challenge is not null (then `challenge` is unused).
extern "C" int fri_launch_fold2(uint64_t *out, const uint64_t *pol, const uint64_t *challenge,
                                const uint64_t *d_challenge, uint64_t n_bits_ext, uint64_t prev_bits,
                                uint64_t cur_bits, const NttTablesHost *tw, hipStream_t st);
extern "C" int fri_launch_transpose(uint64_t *aux, const uint64_t *pol, uint64_t degree, uint32_t t_bits,
                                    hipStream_t st);
extern "C" int fri_launch_gl_binop(uint64_t *out, const uint64_t *a, const uint64_t *b, uint64_t n, int op, hipStream_t st);
