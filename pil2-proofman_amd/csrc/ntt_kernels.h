// Launch interface of the LDE / transpose kernels (internal to libpil2_commit_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NTT_THREADS 1024                   // max threads per block (16 values each)
#define NTT_LOG_TILE 14                    // 2^14 values per block tile
#define NTT_LDS_WORDS (16384 + 1024)       // tile + one pad word per 16 (139,264 bytes)
#define NTT_MAX_LOG 28                     // largest transform handled (2^28 rows)
#define NTT_MAX_STRIDED 10                 // most index bits one strided launch covers

// Split tables of one domain 2^(2*bits): w^x = lo[x & (2^bits - 1)] * hi[x >> bits]
// (x < 2^(2*bits), w = the domain's root, forward and inverse) and 7^e likewise.
struct NttSplitTables {
    uint64_t *twl_f, *twh_f, *twl_i, *twh_i;  // 2^bits entries each
    uint64_t *shl, *shh;                   // 2^bits entries each
    uint32_t bits;
};

struct NttTablesHost {
    uint64_t *tkf, *tki;                   // 2^15 entries each: [2^k + y] = w_{2^k}^(+-y), k <= 14
    NttSplitTables split[2];               // [0]: the 2^24 domain (bits 12), [1]: 2^28 (bits 14)
};

// LDE launch plan: the largest block tile (log2 values, <= 14; a shape that needs a
// larger one gets it) and the most index bits per in-tile round (radix 2^radix_log).
struct NttPlan {
    uint32_t tile_log, radix_log;
};

extern "C" int ntt_init_device(void);
// Nonzero signature of the launches ntt_launch_lde makes for this shape and plan
// (plans with equal signatures launch the same kernels); 0 for a shape the plan
// machinery does not cover (log_n < 4 uses a direct kernel).
extern "C" uint64_t ntt_lde_plan_signature(uint32_t log_n, uint32_t log_ext, const NttPlan *plan);
extern "C" int ntt_launch_lde(uint64_t *lde, const uint64_t *trace, uint64_t *scratch, uint32_t log_n, uint32_t log_ext,
                              uint64_t ncols, const NttTablesHost *th, const NttPlan *plan, hipStream_t st,
                              hipEvent_t *ev);
// computeQ (INTT over NExt, x shiftIn^p, regroup, zero-extend, NTT over NExt): q element
// (k, d) at q[k*q_rs + d*q_cs] (q_rs == 1: column-major; else q_cs == 1), cmq column-major
// (q_deg*q_dim columns of NExt), coef scratch >= q_deg*q_dim*N words.
extern "C" int ntt_launch_compute_q(uint64_t *cmq, const uint64_t *q, uint64_t q_rs, uint64_t q_cs, uint64_t *coef,
                                    uint32_t log_n, uint32_t log_ext, uint32_t q_deg, uint32_t q_dim,
                                    const NttTablesHost *th, hipStream_t st);
// NTT / INTT (natural order; the inverse includes 1/M): column-major, out of place, M >= 16
extern "C" int ntt_launch_ntt(uint64_t *dst, const uint64_t *src, uint32_t lm, uint64_t ncols, int inverse,
                              const NttTablesHost *th, hipStream_t st);
// M < 16, element (r, c) at base[r*rs + c*cs]; dst may equal src
extern "C" int ntt_launch_ntt_tiny(uint64_t *dst, const uint64_t *src, uint64_t rs, uint64_t cs, uint32_t lm,
                                   uint64_t ncols, int inverse, hipStream_t st);
extern "C" int ntt_launch_cm2rm(uint64_t *out, const uint64_t *in, uint64_t rows, uint64_t cols, uint64_t in_cs,
                                uint64_t out_rs, hipStream_t st);
extern "C" int ntt_launch_rm2cm(uint64_t *out, const uint64_t *in, uint64_t rows, uint64_t cols, uint64_t in_rs,
                                uint64_t out_cs, hipStream_t st);
extern "C" int ntt_launch_fill_uniform(uint64_t *out, uint64_t n, uint64_t seed, uint64_t start, hipStream_t st);
// columns [0, w) of `rows` host rows (stride ld words, page-locked memory) -> dst, rows x w row-major
extern "C" int ntt_launch_gather_rows(uint64_t *dst, const uint64_t *src, uint64_t rows, uint32_t w, uint64_t ld,
                                      hipStream_t st);
