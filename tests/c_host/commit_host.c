/* A plain C host of libpil2_commit_hip.so: the caller proofman's Rust FFI would be,
 * through the same header and the same entry points (set_hash_family,
 * write_custom_commit, pil2hip_commit_witness, pil2hip_last_error;
 * starks_api.hpp:28,96,98).
 *
 * usage: commit_host TRACE_FILE N_BITS N_BITS_EXT N_COLS ARITY
 *        commit_host --witness TRACE_FILE N_BITS N_BITS_EXT N_COLS ARITY OUT_FILE
 * TRACE_FILE holds 2^N_BITS x N_COLS little-endian u64 words, row-major. Prints
 * the four root words in hex on one line; exits non-zero on any error.
 *
 * --witness lays out an aux_trace the way the reference's StarkInfo does for
 * stage 1 (the trace section, then the extended "cm1" section, then the "mt1"
 * nodes, with gaps between them), fills it with a sentinel word, calls
 * pil2hip_commit_witness with the two offsets, checks that every word outside
 * the two output sections still holds the sentinel, and writes the LDE section
 * followed by the nodes section to OUT_FILE. */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/pil2_commit_hip.h"

#define SENTINEL 0x5A5A5A5A5A5A5A5AULL

static uint64_t *read_trace(const char *path, size_t n)
{
    uint64_t *trace = malloc(n * sizeof(uint64_t) + 8);
    FILE *f = fopen(path, "rb");
    if (!trace || !f || fread(trace, sizeof(uint64_t), n, f) != n) {
        fprintf(stderr, "cannot read %zu words from %s\n", n, path);
        exit(2);
    }
    fclose(f);
    return trace;
}

static int witness(int argc, char **argv)
{
    if (argc != 8) {
        fprintf(stderr, "usage: %s --witness TRACE_FILE N_BITS N_BITS_EXT N_COLS ARITY OUT_FILE\n", argv[0]);
        return 2;
    }
    uint64_t nbits = strtoull(argv[3], NULL, 10), nbits_ext = strtoull(argv[4], NULL, 10);
    uint64_t ncols = strtoull(argv[5], NULL, 10), arity = strtoull(argv[6], NULL, 10);
    if (nbits > 22 || nbits_ext > 22 || nbits_ext < nbits || ncols > 4096 || arity < 2 || arity > 4) {
        fprintf(stderr, "shape out of range\n");
        return 2;
    }
    size_t n = ((size_t)1 << nbits) * ncols, lde = ((size_t)1 << nbits_ext) * ncols;
    size_t nodes = pil2hip_tree_num_elements((uint64_t)1 << nbits_ext, arity);
    uint64_t *trace = read_trace(argv[2], n);
    /* [trace | gap | cm1 extended | gap | mt1 | tail] */
    size_t gap = 37, off_dst = n + gap, off_mt = off_dst + lde + gap, total = off_mt + nodes + gap;
    uint64_t *aux = malloc(total * sizeof(uint64_t));
    if (!aux) return 2;
    for (size_t i = 0; i < total; i++) aux[i] = SENTINEL;
    memcpy(aux, trace, n * sizeof(uint64_t));
    pil2hip_set_abort_on_error(false);
    set_hash_family(2);
    void *dev = gen_device_buffers(0, 1, NULL, (uint32_t)arity, (uint32_t)nbits_ext);
    uint64_t root[4] = {0, 0, 0, 0};
    int64_t rc = pil2hip_commit_witness(root, arity, nbits, nbits_ext, ncols, aux, aux, off_dst, off_mt, dev);
    free_device_buffers(dev);
    if (rc) {
        fprintf(stderr, "pil2hip_commit_witness: %" PRId64 " %s\n", rc, pil2hip_last_error());
        return 1;
    }
    if (memcmp(aux, trace, n * sizeof(uint64_t))) {
        fprintf(stderr, "trace section modified\n");
        return 1;
    }
    for (size_t i = n; i < total; i++) {
        int in_out = (i >= off_dst && i < off_dst + lde) || (i >= off_mt && i < off_mt + nodes);
        if (!in_out && aux[i] != SENTINEL) {
            fprintf(stderr, "word %zu outside the output sections was written\n", i);
            return 1;
        }
    }
    FILE *f = fopen(argv[7], "wb");
    if (!f || fwrite(aux + off_dst, 8, lde, f) != lde || fwrite(aux + off_mt, 8, nodes, f) != nodes || fclose(f)) {
        fprintf(stderr, "cannot write %s\n", argv[7]);
        return 2;
    }
    printf("%016" PRIx64 " %016" PRIx64 " %016" PRIx64 " %016" PRIx64 "\n", root[0], root[1], root[2], root[3]);
    free(aux);
    free(trace);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && !strcmp(argv[1], "--witness")) return witness(argc, argv);
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
    uint64_t *trace = read_trace(argv[1], n);
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
