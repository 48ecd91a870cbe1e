/*
 * pil2_commit_hip.h — C-ABI of libpil2_commit_hip.so, the MI355X implementation of
 * pil2-stark's polynomial-commitment path (Goldilocks LDE, Poseidon2 Merkle commit,
 * FRI fold).
 *
 * Two groups of entry points:
 *
 *  1. Drop-ins for the reference ABI (pil2-stark/src/api/starks_api.hpp). Same
 *     names, argument order and meaning; proofman's Rust FFI
 *     (provers/starks-lib-c/src/ffi_starks.rs) links them unchanged. Buffers are
 *     host memory, row-major, exactly as the CPU backend takes them.
 *
 *  2. Device-resident entry points (prefix pil2hip_) with no single reference
 *     counterpart: the same operations on buffers already in HBM, launched on a
 *     caller-supplied hipStream_t passed as void* (NULL = the null stream). They
 *     return 0 on success and a negative code on error; pil2hip_last_error()
 *     then gives a message. Device layouts:
 *       trace  row-major  N x nCols   (element (r, c) at r*nCols + c)
 *       LDE    col-major  nCols x NExt (element (r, c) at c*NExt + r)
 *       nodes  the reference Merkle layout (goldilocks_tooling.hpp:9-21):
 *              leaf digests (4 words per row), zero pad digests, next level,
 *              ..., root = last 4 words; pil2hip_tree_num_elements() words.
 *
 * All field values written by this library are canonical (< p). Inputs may hold
 * any 64-bit word; they are reduced mod p on load, matching the reference, whose
 * results are compared through Goldilocks::toU64.
 */
#ifndef PIL2_COMMIT_HIP_H
#define PIL2_COMMIT_HIP_H
#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------ reference ABI drop-ins -- */

/* starks_api.hpp:28 (impl starks_api.cpp:43). 1 = Poseidon1, 2 = Poseidon2.
 * Only Poseidon2 is implemented; commits under family 1 (or unset) fail. */
void set_hash_family(uint8_t fam);

/* starks_api.hpp:96 (CPU impl starks_api.cpp:633-681). buffer: host row-major
 * trace of 2^nBits x nCols u64; root: 4 u64 out; bufferFile: "" for none, else
 * the file [root 32 B | trace | LDE (row-major) | tree nodes] is written, as the
 * CPU backend writes it. arity 2/3/4 selects Poseidon2 width 8/12/16. d_buffers
 * may be NULL or a handle from gen_device_buffers. Errors print a message and
 * abort, as the reference does (exitProcess), unless
 * pil2hip_set_abort_on_error(false) was called. */
void write_custom_commit(void *root, uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols,
                         void *d_buffers, void *buffer, char *bufferFile);

/* commit_witness (starks_api.hpp:98; CPU impl commit_witness_cpu,
 * starks_api.cpp:683-723) for an unpacked trace whose witness columns are already
 * computed (no packed cm1, no witness hints: the SetupCtx-dependent steps before
 * the LDE stay with the caller). The host trace (row-major 2^nBits x nCols, the
 * "cm1" stage-1 section) is extended and merkelized on the GPU, and the outputs
 * land where commit_witness_cpu leaves them for the later stages and openings:
 *   aux_trace[offset_dst ..]  the row-major LDE, 2^nBitsExt x nCols ("cm1" extended)
 *   aux_trace[offset_mt ..]   the Merkle nodes, pil2hip_tree_num_elements words ("mt1")
 *   root                      4 words (MerkleTreeGL::getRoot)
 * The numbers are the ones commit_witness_cpu reads from its SetupCtx:
 * starkStruct.nBits, nBitsExt and merkleTreeArity, mapSectionsN["cm1"],
 * mapOffsets[("cm1", true)] and mapOffsets[("mt1", true)]. Nothing else in
 * aux_trace is written. d_buffers: NULL or a gen_device_buffers handle (its
 * stream is used). Synchronous. Returns 0, or a negative code (-35: the two
 * sections overlap). */
int64_t pil2hip_commit_witness(uint64_t *root, uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols,
                               const uint64_t *trace, uint64_t *aux_trace, uint64_t offset_dst, uint64_t offset_mt,
                               void *d_buffers);

/* starks_api.hpp:190 / :193. Returns an opaque handle that owns a HIP stream on
 * the current device (default priority). node_rank/node_size/numa_nodes are
 * accepted for signature compatibility; this build uses one GPU per process. */
void *gen_device_buffers(uint32_t node_rank, uint32_t node_size, const int32_t *numa_nodes, uint32_t arity,
                         uint32_t max_n_bits_ext);
void free_device_buffers(void *d_buffers);

/* ConstTree (pil2-stark/src/starkpil/const_pols.hpp:25-103), GL trees; the
 * reference entry points (starks_api.hpp:56-63) take a StarkInfo, these take its
 * four numbers. The const tree buffer is the LDE of the constant polynomials,
 * row-major 2^nBitsExt x nConstants, followed by the Merkle nodes (the CPU
 * layout; the same bytes are the file).
 * getConstTreeSizeGL (const_pols.hpp:25-30): words in the buffer. */
uint64_t pil2hip_const_tree_size(uint64_t nBitsExt, uint64_t nConstants, uint64_t arity);

/* calculateConstTreeGL (const_pols.hpp:33-44) on device: d_constPols row-major
 * 2^nBits x nConstants; d_tree receives pil2hip_const_tree_size words;
 * d_scratch >= (2^nBits + 2^nBitsExt) * nConstants words. Asynchronous. */
int64_t pil2hip_const_tree(uint64_t *d_tree, const uint64_t *d_constPols, uint64_t *d_scratch, uint64_t nBits,
                           uint64_t nBitsExt, uint64_t nConstants, uint64_t arity, void *stream);

/* calculate_const_tree (starks_api.hpp:61) with host buffers: pConstPols row-major
 * 2^nBits x nConstants in, pConstTree (pil2hip_const_tree_size words) out.
 * d_buffers (gen_device_buffers) selects the stream, or NULL. Synchronous. */
int64_t pil2hip_calculate_const_tree(void *pConstPols, void *pConstTree, uint64_t nBits, uint64_t nBitsExt,
                                     uint64_t nConstants, uint64_t arity, void *d_buffers);

/* write_const_tree (starks_api.hpp:63) / MerkleTreeGL::writeFile
 * (merkleTreeGL.cpp:338-345): writes the buffer as the file. Host only. */
int64_t pil2hip_write_const_tree(const void *pConstTree, uint64_t nBitsExt, uint64_t nConstants, uint64_t arity,
                                 const char *treeFilename);

/* load_const_tree (starks_api.hpp:56) / ConstTree::loadConstTree GL branch
 * (const_pols.hpp:66-103): reads constTreeSize words into pConstTree and, when
 * verkeyFilename is given, checks the root (last 4 words) against that JSON
 * array of 4 integers. Returns 1 if loaded (and verified), 0 if not, a negative
 * code if the key file is not 4 integers. Host only. */
int64_t pil2hip_load_const_tree(void *pConstTree, const char *treeFilename, uint64_t constTreeSize,
                                const char *verkeyFilename);

/* starks_api.hpp:203 */
uint64_t get_num_gpus(void);

/* starks_api.hpp:230 (starks_backend.cpp:233-243). This library is the GPU
 * backend only: set_gpu_mode(true) returns true, set_gpu_mode(false) returns
 * false (there is no CPU fallback inside this library). */
bool set_gpu_mode(bool use_gpu);

/* ---------------------------------------------------- device-resident path -- */

/* getTreeNumElements (goldilocks_tooling.hpp:9-21): u64 words in the tree buffer. */
uint64_t pil2hip_tree_num_elements(uint64_t nRows, uint64_t arity);

/* NTT_Goldilocks::LDE (ntt_goldilocks.cpp:369-404) on device:
 * d_lde[c*NExt + j] = p_c(7 * w_NExt^j), p_c interpolating column c of d_trace.
 * d_scratch: >= 2^nBits * nCols words. nBitsExt - nBits >= 0, nBitsExt <= 28. */
int64_t pil2hip_lde(uint64_t *d_lde, const uint64_t *d_trace, uint64_t *d_scratch, uint64_t nBits, uint64_t nBitsExt,
                    uint64_t nCols, void *stream);

/* Poseidon2Goldilocks<4*arity>::merkletree (poseidon2_goldilocks.cpp:130-175) as
 * called by MerkleTreeGL::merkelize (merkleTreeGL.cpp:316-336). Row r, column c
 * of the source is d_src[r*rowStride + c*colStride]. */
int64_t pil2hip_merkelize(uint64_t *d_nodes, const uint64_t *d_src, uint64_t nRows, uint64_t nCols, uint64_t rowStride,
                          uint64_t colStride, uint64_t arity, void *stream);

/* LDE followed by merkelize over the LDE rows: commit_witness_cpu's core
 * (starks_api.cpp:683-723) without SetupCtx. */
int64_t pil2hip_commit(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *d_trace, uint64_t *d_scratch, uint64_t arity,
                       uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, void *stream);

/* No counterpart in the reference. pil2hip_commit without d_scratch, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): nothing is allocated, by this call or
 * any call before it. The scratch is a device array inside the library of
 * PIL2HIP_NOALLOC_SCRATCH_WORDS words (128 MiB), one per GPU, held from the
 * library's load until the process ends; 2^nBits * nCols above that returns -36.
 * All calls on one GPU share that array: calls that overlap on one GPU (on
 * different streams, or from different threads without ordering) race and give
 * wrong LDEs and roots. Calls ordered on one stream are safe. Outputs as
 * pil2hip_commit's (d_lde column-major, d_nodes), on `stream`. */
#define PIL2HIP_NOALLOC_SCRATCH_WORDS (1ull << 24)
int64_t pil2hip_commit_noalloc(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *d_trace, uint64_t arity,
                               uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, void *stream);

/* pil2hip_commit for a trace in host memory, with the transfer overlapped (no
 * reference counterpart with this signature: it is the column-chunk absorb of the
 * reference's streaming commit, stream_commit.cuh:9-20 / commit_witness_streaming,
 * starks_api.hpp:216, for an unpacked trace and Poseidon2). The row-major host
 * trace (2^nBits x nCols) is copied in groups of groupCols columns (a multiple of
 * the sponge rate 4*arity-4; 0 picks one group under 256 MiB of trace, else a
 * first group of about an eighth of the columns, at least 16, and later groups
 * twice as wide) on a copy stream of the library's own (the first group is read
 * from the page-locked rows by a kernel on `stream`, which the device has free
 * then), double-buffered, while `stream` extends the previous group into
 * its columns of d_lde and absorbs it into the row sponges, whose capacity words
 * carry over in the leaf digests of d_nodes. d_lde and d_nodes end up equal to
 * pil2hip_commit's. d_stage: pil2hip_commit_host_stage_words() words. h_trace
 * must be page-locked (hipHostMalloc): from pageable memory the runtime splits
 * the 2D copies into small pieces. Returns once every host-to-device copy has
 * completed, so h_trace is free on return. d_stage is NOT free on return: the
 * LDE of the last column group still reads it on `stream`, so it may be reused
 * or freed only by work ordered after the commit on `stream` (or after
 * synchronising it). The LDE and tree complete on `stream`. -33: groupCols not a
 * multiple of the rate; -34: h_trace not page-locked. */
int64_t pil2hip_commit_host(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *h_trace, uint64_t *d_stage,
                            uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, uint64_t groupCols,
                            void *stream);
/* Staging words pil2hip_commit_host needs: 3 * 2^nBits * (group width); 0 for a bad arity. */
uint64_t pil2hip_commit_host_stage_words(uint64_t nBits, uint64_t nCols, uint64_t arity, uint64_t groupCols);

/* NTT_Goldilocks::NTT / ::INTT (ntt_goldilocks.hpp:182-191, ntt_goldilocks.cpp:211-260):
 * per column, out_k = sum_j x_j w^(j*k) with w = W[nBits] (inverse: w^-1 and a
 * factor 2^-nBits), natural order in and out. rowMajor: element (r, c) at
 * r*nCols + c (the CPU layout); otherwise column c at c*2^nBits (the GPU layout).
 * d_dst == d_src or NULL transforms in place. d_scratch: 2 * 2^nBits * nCols words
 * for row-major data; 2^nBits * nCols for in-place column-major data; unused
 * otherwise and for 2^nBits < 16. nBits <= 28; nCols == 0 is a no-op. */
int64_t pil2hip_ntt(uint64_t *d_dst, const uint64_t *d_src, uint64_t *d_scratch, uint64_t nBits, uint64_t nCols,
                    bool inverse, bool rowMajor, void *stream);

/* Starks::computeQ (starks.hpp:187-239; GPU twin NTTGoldilocksGPU::computeQ,
 * ntt_goldilocks.cu:1583-1604): the quotient q, qDim columns over the extended
 * coset (element (k, d) at d_q[k*qRowStride + d*qColStride]; column-major with
 * qRowStride 1 and qColStride >= NExt, or row-major with qColStride 1), is
 * interpolated over NExt points; coefficient p*N + i of column d times
 * shiftIn^p (shiftIn = 7^-N) becomes coefficient i of column p*qDim + d, and
 * each of the qDeg*qDim columns is zero-extended and transformed over NExt
 * points into d_cmQ (column-major, column c at d_cmQ + c*NExt). d_q is not
 * modified and must not overlap d_cmQ. d_scratch: >= qDeg*qDim*2^nBits words.
 * With d_nodes != NULL the rows of d_cmQ are then merkelized (the step's
 * merkelize, starks.hpp:228-235); the root is read with pil2hip_get_root.
 * 1 <= qDeg <= 2^(nBitsExt - nBits), nBitsExt <= 28. */
int64_t pil2hip_compute_q(uint64_t *d_cmQ, uint64_t *d_nodes, const uint64_t *d_q, uint64_t qRowStride,
                          uint64_t qColStride, uint64_t *d_scratch, uint64_t nBits, uint64_t nBitsExt, uint64_t qDeg,
                          uint64_t qDim, uint64_t arity, void *stream);

/* One commit of a batch: the arguments of pil2hip_commit. Buffers are device
 * memory owned by the caller and must not overlap between jobs. */
typedef struct Pil2CommitJob {
    const uint64_t *d_trace; /* row-major 2^n_bits x n_cols */
    uint64_t *d_lde;         /* column-major 2^n_bits_ext x n_cols */
    uint64_t *d_nodes;       /* pil2hip_tree_num_elements(2^n_bits_ext, arity) words */
    uint64_t *d_scratch;     /* 2^n_bits x n_cols words */
    uint64_t n_bits, n_bits_ext, n_cols, arity;
} Pil2CommitJob;

/* Batched commit of many AIR instances of uneven shapes: the per-instance
 * commits that proofman issues one by one (commit_witness, starks_api.hpp:98,
 * once per instance from proofman.rs) run as ONE hipGraph launch on `stream`.
 * Jobs are spread over n_streams concurrent branches of the graph (longest
 * first, each to the least-loaded branch); each job's root is copied to
 * d_roots[4*i .. 4*i+4) (d_roots may be NULL). The instantiated graph is cached
 * per job list (pointers and shapes) and replayed on later calls.
 * n_streams = 0 autotunes the whole launch plan on the first call for a job
 * list: the NTT block tile and in-tile radix per LDE shape and the cooperative
 * tree-level threshold per tree shape (each shape measured once per process, on
 * the first batch that has it; see pil2hip_batch_shape_plan), then the branch
 * count (1, 2, 4, 8, 16, each timed over one replay). These runs recompute the
 * same outputs. n_streams > 0 uses that many branches and the default plan
 * (Pil2CommitPlan below; pil2hip_commit_batch_plan takes an explicit one).
 * Returns the branch count used (> 0), 0 for an empty batch, or a negative error
 * code. Asynchronous like pil2hip_commit. */
int64_t pil2hip_commit_batch(const Pil2CommitJob *jobs, uint64_t n_jobs, uint32_t n_streams, uint64_t *d_roots,
                             void *stream);

/* Launch plan of a commit. Every plan computes the same values; they differ in
 * speed only. n_streams = 0 asks the batch autotuner for the branch count; the
 * other fields left 0 take their defaults, except that a plan of all 0 asks the
 * autotuner for everything.
 *   n_streams         graph branches, 1..16
 *   ntt_tile_log      largest NTT block tile, log2 values, 12..14 (a shape that
 *                     needs a larger tile gets it); smaller tiles fit more blocks per CU
 *   ntt_radix_log     most index bits per in-tile round of the NTT block pass: 2, 3, 4
 *                     (radix 4, 8, 16); the strided passes always take the fewest
 *                     rounds of at most 4 bits
 *   coop_max_parents  tree levels with at most this many parents hash with 16 lanes
 *                     per permutation, larger levels with one lane per parent
 * The plain entry points and fields left 0 use {-, 13, 4, 16384}. */
typedef struct Pil2CommitPlan {
    uint32_t n_streams, ntt_tile_log, ntt_radix_log, coop_max_parents;
} Pil2CommitPlan;

/* pil2hip_commit_batch with an explicit plan. A plan with every field 0 (or
 * plan == NULL) is tuned completely, as pil2hip_commit_batch with n_streams = 0
 * does; *chosen then holds the branch count and 0 in the other fields (they are
 * per shape). Otherwise n_streams = 0 is tuned on the first call for this job
 * list (each branch count timed over one replay) and the other fields left 0
 * take their defaults. The plan used is written to *chosen (may be NULL).
 * Returns the branch count used, 0 for an empty batch, or a negative code. */
int64_t pil2hip_commit_batch_plan(const Pil2CommitJob *jobs, uint64_t n_jobs, const Pil2CommitPlan *plan,
                                  Pil2CommitPlan *chosen, uint64_t *d_roots, void *stream);

/* The plan a fully tuned batch uses for a job of this shape: the tuner's NTT tile
 * and radix for (nBits, nBitsExt, ceil(log2 nCols)) and its level threshold for
 * (nBitsExt, arity), with defaults where a shape has not been measured yet
 * (out->n_streams = 0). Returns 1 when both were measured, 0 otherwise, or a
 * negative code. No counterpart in the reference: it reports the autotuner's
 * per-shape choices. */
int64_t pil2hip_batch_shape_plan(uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, uint64_t arity,
                                 Pil2CommitPlan *out);

/* Destroys every cached batch graph. */
void pil2hip_batch_release(void);

/* MerkleTreeGL::getGroupProof (merkleTreeGL.cpp:145-175) for nQueries row
 * indices d_idx (device memory) at once: proof q occupies
 * d_out[q*pw .. (q+1)*pw), pw = pil2hip_proof_words(nRows, nCols, arity, llv):
 * the nCols values of row idx (element (r, c) at d_src[r*rowStride + c*colStride]:
 * rowStride 1, colStride NExt for the column-major LDE; rowStride = width,
 * colStride 1 for a row-major source such as a FRI layer), then per tree level
 * above the last-level cut (genMerkleProof, merkleTreeGL.cpp:157-175) the
 * arity-1 sibling digests in ascending order. An index >= nRows yields an
 * all-zero proof. FRI::proveFRIQueries (fri.hpp:131-139) is this call on a
 * layer tree with indices reduced mod 2^currentBits. */
int64_t pil2hip_group_proofs(uint64_t *d_out, const uint64_t *d_src, uint64_t rowStride, uint64_t colStride,
                             uint64_t nCols, const uint64_t *d_nodes, uint64_t nRows, uint64_t arity, uint64_t llv,
                             const uint64_t *d_idx, uint64_t nQueries, void *stream);

/* Merkle proof levels (MerkleTreeGL::getMerkleProofLength, merkleTreeGL.cpp:50-55)
 * and words per group proof (nCols + (arity-1)*4*levels). */
uint64_t pil2hip_proof_levels(uint64_t nRows, uint64_t arity, uint64_t llv);
uint64_t pil2hip_proof_words(uint64_t nRows, uint64_t nCols, uint64_t arity, uint64_t llv);

/* Proof-of-work grinding: Poseidon2Goldilocks<8>::grinding
 * (poseidon2_goldilocks.cpp:177-232; gen_proof's runGrinding through
 * starks_api_internal.hpp:21). *nonce = the smallest i with
 * permute([challenge[0..3), i, 0, 0, 0, 0])[0] < 2^(64 - nBits), comparing the
 * canonical value as the reference's tests do; challenge is 3 words in host
 * memory. Synchronous on `stream`. -37 if no nonce exists below the reference's
 * bound ceil(-128 ln 2 / ln(1 - 2^-nBits)). */
int64_t pil2hip_grinding(uint64_t *nonce, const uint64_t *challenge, uint32_t nBits, void *stream);

/* Fiat-Shamir transcript (TranscriptGL, transcriptGL.hpp:11-58 /
 * transcriptGL.cpp:4-119) with its state in device memory: Poseidon2 of width
 * 4*arity, rate 4*(arity-1), capacity 4. put absorbs words from device memory
 * (e.g. a Merkle root where the commit left it) or from host memory;
 * get_field squeezes 3 words (a cubic challenge); get_state returns the first
 * n state words; get_permutations draws n query indices of nBits each from
 * ceil(n*nBits/63) squeezed fields, 63 low bits per field. Words returned are
 * canonical (the reference returns its internal word, equal mod p). */
void *pil2hip_transcript_create(uint64_t arity);
int64_t pil2hip_transcript_put(void *t, const uint64_t *d_in, uint64_t n, void *stream);
int64_t pil2hip_transcript_put_host(void *t, const uint64_t *h_in, uint64_t n, void *stream);
int64_t pil2hip_transcript_get_field(void *t, uint64_t *h_out, void *stream);
int64_t pil2hip_transcript_get_state(void *t, uint64_t *h_out, uint64_t n, void *stream);
int64_t pil2hip_transcript_get_permutations(void *t, uint64_t *h_out, uint64_t n, uint64_t nBits, void *stream);
void pil2hip_transcript_destroy(void *t);

/* ---------------------------------------------------------------- serving --
 * A commit server: jobs arrive one by one (pil2hip_serve_submit), are admitted
 * first-come first-served as in-flight slots and HBM budget free up, run
 * concurrently on several device streams (LDE, Merkle tree, the job's group
 * proofs), and report completions (pil2hip_serve_poll) with their timestamps.
 * A job's LDE and tree can stay resident for later openings
 * (pil2hip_serve_open) until pil2hip_serve_release. The reference's analogue is
 * proofman's per-instance commit_witness calls and its stream/slot API
 * (configure_stream_commit_slots / commit_witness_streaming,
 * starks_api.hpp:214-216); this engine admits by HBM budget instead. */
typedef struct Pil2ServeConfig {
    uint64_t hbm_budget_bytes; /* device arena holding every admitted job's buffers */
    uint32_t n_streams;        /* concurrent device streams, 1..16 */
    uint32_t max_inflight;     /* admitted, not yet completed jobs, 1..256 */
    uint64_t arity;            /* Merkle arity, 2..4 (Poseidon2 width 4*arity) */
    uint64_t llv;              /* last level verification of the group proofs */
} Pil2ServeConfig;

typedef struct Pil2ServeRequest {
    uint64_t n_bits, n_bits_ext, n_cols;
    const uint64_t *h_trace;   /* host row-major trace, read before completion, or NULL */
    const uint64_t *d_trace;   /* device row-major trace owned by the caller, or NULL */
    uint64_t seed;             /* both NULL: synthetic trace pil2hip_fill_uniform(seed) */
    uint64_t n_queries;        /* group proofs computed with the commit */
    const uint64_t *h_queries; /* row indices < 2^n_bits_ext (copied at submit) */
    uint64_t *h_root;          /* 4 words, written before the completion is reported */
    uint64_t *h_proofs;        /* n_queries * pil2hip_proof_words(2^n_bits_ext, n_cols, arity, llv), or NULL */
    uint64_t tag;              /* returned in the completion */
    uint32_t keep_resident;    /* 1: LDE and tree stay in HBM until pil2hip_serve_release */
} Pil2ServeRequest;

typedef struct Pil2ServeCompletion {
    uint64_t job_id, tag;
    int64_t status; /* 0, or a negative error code (-31: cancelled) */
    double t_submit_us, t_admit_us, t_done_us; /* engine clock, pil2hip_serve_now_us */
    uint32_t stream;
} Pil2ServeCompletion;

typedef struct Pil2ServeStats {
    uint64_t submitted, admitted, completed, failed, cancelled, queued, running, resident_jobs;
    uint64_t bytes_in_use, peak_bytes, budget_bytes;
    uint32_t peak_running;
} Pil2ServeStats;

/* Creates a server on the current device (allocates the arena, streams and a
 * scheduler thread). NULL on error (pil2hip_last_error gives the reason). */
void *pil2hip_serve_create(const Pil2ServeConfig *cfg);
/* Queues a job; returns its id (>= 0) or a negative code: -30 when the job
 * can never fit the HBM budget, -10/-11/-12 for a bad shape, -28 for a query
 * index out of range. Never blocks on the device. */
int64_t pil2hip_serve_submit(void *srv, const Pil2ServeRequest *req);
/* Copies up to max completions into out; waits up to timeout_ms for the first
 * one. Returns the number copied. */
int64_t pil2hip_serve_poll(void *srv, Pil2ServeCompletion *out, uint64_t max, int32_t timeout_ms);
/* Cancels a queued job (completion status -31); -32 if it is already admitted. */
int64_t pil2hip_serve_cancel(void *srv, uint64_t job_id);
/* Group proofs on a completed, resident job (synchronous). */
int64_t pil2hip_serve_open(void *srv, uint64_t job_id, const uint64_t *h_queries, uint64_t n_queries,
                           uint64_t *h_proofs);
/* Frees a completed job's resident LDE and tree. */
int64_t pil2hip_serve_release(void *srv, uint64_t job_id);
int64_t pil2hip_serve_stats(void *srv, Pil2ServeStats *out);
double pil2hip_serve_now_us(void *srv);
/* HBM bytes a job holds while admitted (LDE, tree, trace, scratch, queries, proofs). */
uint64_t pil2hip_serve_job_bytes(uint64_t n_bits, uint64_t n_bits_ext, uint64_t n_cols, uint64_t arity,
                                 uint64_t n_queries, uint64_t llv, bool device_trace);
/* Drains admitted jobs, cancels queued ones, frees everything. */
void pil2hip_serve_destroy(void *srv);


/* MerkleTreeGL::getRoot (merkleTreeGL.cpp:103-107): synchronous copy of the root
 * (4 words) into host memory. */
int64_t pil2hip_get_root(uint64_t *root, const uint64_t *d_nodes, uint64_t nRows, uint64_t arity, void *stream);

/* MerkleTreeGL::getLevel (merkleTreeGL.cpp:84-101), the last_levels that
 * FRI::merkelize stores (fri.hpp:108): the nodes of the first tree level with at
 * most arity^llv entries, copied into h_level (host memory) and zero-padded to
 * arity^llv digests, i.e. pil2hip_level_words(arity, llv) words. llv = 0 leaves
 * h_level untouched, as the reference does. Synchronous on `stream`. */
uint64_t pil2hip_level_words(uint64_t arity, uint64_t llv);
int64_t pil2hip_get_level(uint64_t *h_level, const uint64_t *d_nodes, uint64_t nRows, uint64_t arity, uint64_t llv,
                          void *stream);

/* FRI::fold (fri.hpp:32-99) for step > 0: d_out[g] (g < 2^currentBits, cubic,
 * 3 words each) from d_pol (2^prevBits cubic values). challenge: 3 words in host
 * memory. d_out may equal d_pol (in place, as the reference). step == 0 is a no-op.
 * Fold ratios 2^1 .. 2^10 (prevBits - currentBits <= 10; -17 above that). */
int64_t pil2hip_fri_fold(uint64_t *d_out, const uint64_t *d_pol, const uint64_t *challenge, uint64_t step,
                         uint64_t nBitsExt, uint64_t prevBits, uint64_t currentBits, void *stream);

/* FRI getTransposed (fri.hpp:209-226): d_aux[(i*h + j)] = d_pol[(j*w + i)],
 * w = 2^tBits, h = degree / w, 3-word elements. */
int64_t pil2hip_fri_transpose(uint64_t *d_aux, const uint64_t *d_pol, uint64_t degree, uint64_t tBits, void *stream);

/* The FRI layer loop of gen_proof: for s = 0 .. nSteps-1, step s > 0 folds
 * (FRI::fold, fri.hpp:32-99) d_pol (s = 1) or d_work (later steps, in place) into
 * d_work with the challenge at d_challenges[3s .. 3s+3) (device memory, read when
 * the fold runs), and step s < nSteps-1 commits its layer (FRI::merkelize,
 * fri.hpp:102-109): getTransposed into d_sources[s] (2^steps[s] cubic values)
 * and merkelize into d_nodes[s] (pil2hip_tree_num_elements(2^steps[s+1], arity)
 * words), i.e. a tree of 2^steps[s+1] rows of 3*2^(steps[s]-steps[s+1]) words.
 * steps[0] must equal nBitsExt and each step drops 1 .. 10 bits
 * (stark_struct.rs:162-166 gives the schedule). graph = true captures the whole
 * sequence once per argument set (all pointers, steps, arity) and replays the
 * cached graph on later calls, so new challenges written into d_challenges are
 * used without a new capture; in the graph every layer's tree is a branch that
 * depends only on that layer's transpose, so the trees of different layers run
 * side by side and the roots are in d_nodes when the launch completes on
 * `stream`. graph = false enqueues the kernels directly on `stream`, one after
 * another (for a caller's own capture). d_work needs 3*2^steps[1] words. Asynchronous on
 * `stream`. No counterpart in the reference API: gen_proof (starks_api.hpp:122)
 * runs this loop internally. */
int64_t pil2hip_fri_layers(const uint64_t *d_pol, uint64_t *d_work, const uint64_t *d_challenges, uint64_t nBitsExt,
                           const uint64_t *steps, uint64_t nSteps, uint64_t arity, uint64_t *const *d_sources,
                           uint64_t *const *d_nodes, bool graph, void *stream);

/* Destroys every cached FRI layer graph. */
void pil2hip_fri_release(void);

/* Column-major -> row-major copy (e.g. to hand an LDE back in the CPU layout). */
int64_t pil2hip_transpose_cm2rm(uint64_t *d_out, const uint64_t *d_in, uint64_t rows, uint64_t cols,
                                uint64_t inColStride, uint64_t outRowStride, void *stream);

/* Poseidon2Goldilocks<width>::permute (poseidon2_goldilocks.cpp:10-41) on
 * `count` consecutive states of `width` words (8, 12 or 16), in place. */
int64_t pil2hip_p2_permute(uint64_t *d_states, uint64_t count, uint32_t width, void *stream);

/* Synthetic inputs: d_out[i] = element (start + i) of the seeded uniform stream
 * over [0, p) (counter-based SplitMix64, re-mixed while >= p). Not part of the
 * reference ABI; used to build benchmark inputs directly in HBM. */
int64_t pil2hip_fill_uniform(uint64_t *d_out, uint64_t n, uint64_t seed, uint64_t start, void *stream);

/* Element-wise device field arithmetic, for checking the GPU's Goldilocks
 * operations against exact integers (no reference counterpart; the reference's
 * goldilocks_*_ffi entry points, starks_api.hpp:172-185, are the scalar CPU
 * analogue). op 0: add, 1: sub, 2: mul (canonical results of a[i] op b[i]);
 * op 3: canonical residue of the 128-bit value b[i]*2^64 + a[i]. */
int64_t pil2hip_gl_binop(uint64_t *d_out, const uint64_t *d_a, const uint64_t *d_b, uint64_t n, int32_t op,
                         void *stream);

/* Throughput calibration of the Poseidon2 kernels' reduced 64x64 product (no
 * reference counterpart; it gives the measured product peak that the leaf
 * kernel's field multiplications per second are compared with). Lane i runs 18
 * independent chains a[k] <- a[k]*b[k] for `iters` steps, with a[k] = d_in[36i+k],
 * b[k] = d_in[36i+18+k], and writes d_out[18i+k] = a[k]*b[k]^iters mod p
 * (canonical): lanes*iters*18 products. */
int64_t pil2hip_mul_peak(uint64_t *d_out, const uint64_t *d_in, uint64_t lanes, uint32_t iters, void *stream);

/* Phase timing of pil2hip_commit with HIP events on the commit's own stream (no
 * reference counterpart; the reference's TimerGPU, utils/gpu_timer.cuh, plays
 * this role). While enabled, every pil2hip_commit synchronises its stream and
 * adds the elapsed milliseconds of its phases to accumulators:
 * [0] row->column transpose, [1] iNTT strided pass, [2] fused iNTT/coset/NTT
 * block, [3] NTT strided pass, [4] Poseidon2 leaf hashing, [5] tree levels.
 * pil2hip_timing_read copies the 6 sums and returns the number of timed commits. */
void pil2hip_timing_enable(bool on);
uint64_t pil2hip_timing_read(double *phase_ms, bool reset);

/* Message of the last failing call on this thread ("" if none). */
const char *pil2hip_last_error(void);

/* When true (default, as the reference), errors inside the void reference
 * entry points print and abort; when false they only record the error. */
void pil2hip_set_abort_on_error(bool abort_on_error);

#ifdef __cplusplus
}
#endif
#endif
