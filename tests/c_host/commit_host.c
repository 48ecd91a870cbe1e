/* A plain C host of libpil2_commit_hip.so: the caller proofman's Rust FFI would be,
 * through the same header and the same entry points (set_hash_family,
 * write_custom_commit, pil2hip_last_error; starks_api.hpp:28,96).
 *
 * usage: commit_host TRACE_FILE N_BITS N_BITS_EXT N_COLS ARITY
 * TRACE_FILE holds 2^N_BITS x N_COLS little-endian u64 words, row-major. Prints
 * the four root words in hex on one line; exits non-zero on any error. */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/pil2_commit_hip.h"

int main(int argc, char **argv)
{
    if (argc != 6) {
        fprintf(stderr, "usage: %s TRACE_FILE N_BITS N_BITS_EXT N_COLS ARITY\n", argv[0]);
        return 2;
    }
    uint64_t nbits = strtoull(argv[2], NULL, 10), nbits_ext = strtoull(argv[3], NULL, 10);
    uint64_t ncols = strtoull(argv[4], NULL, 10), arity = strtoull(argv[5], NULL, 10);
    if (nbits > 24 || ncols > 4096) {
        fprintf(stderr, "shape out of range\n");
        return 2;
    }
    size_t n = ((size_t)1 << nbits) * ncols;
    uint64_t *trace = malloc(n * sizeof(uint64_t) + 8);
    FILE *f = fopen(argv[1], "rb");
    if (!trace || !f || fread(trace, sizeof(uint64_t), n, f) != n) {
        fprintf(stderr, "cannot read %zu words from %s\n", n, argv[1]);
        return 2;
    }
    fclose(f);
    pil2hip_set_abort_on_error(false);
    set_hash_family(2);
    uint64_t root[4] = {0, 0, 0, 0};
    write_custom_commit(root, arity, nbits, nbits_ext, ncols, NULL, trace, "");
    const char *err = pil2hip_last_error();
    if (err && err[0]) {
        fprintf(stderr, "write_custom_commit: %s\n", err);
        return 1;
    }
    printf("%016" PRIx64 " %016" PRIx64 " %016" PRIx64 " %016" PRIx64 "\n", root[0], root[1], root[2], root[3]);
    free(trace);
    return 0;
}
