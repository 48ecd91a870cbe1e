// Goldilocks LDE kernels for gfx950 (MI355X).
//
// Implements the semantics of NTT_Goldilocks::LDE
// (pil2-stark/src/goldilocks/src/ntt_goldilocks.cpp:369-404): for every column,
// out[j] = p(7 * w_NExt^j) where p interpolates the column over <w_N>.
//
// Device layout: a trace arrives row-major (N x C, the CPU path's layout,
// starks.hpp:151-152) and is transposed once into column-major scratch; the LDE
// is produced column-major (column c at lde + c*NExt), which keeps every NTT pass
// and the row hashing coalesced.
//
// Transform plan (four-step, no explicit bit-reversal pass):
//   iNTT (size N = n1*n2)   pass 1: strided DIF over m (n1 points, stride n2),
//                                   then x w_N^-(i*bitrev(m))
//                           pass 2: contiguous DIF over i (n2 points) -> the
//                                   coefficients in bit-reversed order
//   coset+extend            scale coefficient e by 7^e/N; in bit-reversed order
//                           the zero-padded array of size NExt = N*B holds each
//                           coefficient followed by B-1 zeros, and the first
//                           log2(B) DIT stages turn that into B copies
//   NTT (size NExt = n1*n2B) pass 3: contiguous DIT over blocks of n2*B
//                                   (fused with pass 2 in one kernel), then
//                                   x w_NExt^(bitrev(m)*k)
//                           pass 4: strided DIT over m (n1 points, stride n2*B)
//                                   -> natural order
#include "gl_field.h"
#include "ntt_kernels.h"

// --------------------------------------------------------------- tables ----
// All twiddles derive from the reference's root chain W[k] (goldilocks_base_field.cpp:11-44).
//   roots_f[h + j] = w_{2h}^j        (h = 1, 2, ..., LDS_ELEMS/2; j < h)   sub-FFT stages
//   roots_i[h + j] = w_{2h}^-j
//   tw_f[0][x & 4095] * tw_f[1][x >> 12] = w_{2^24}^x                     inter-pass
//   sh[0][e & 4095] * sh[1][e >> 12] = 7^e                                  coset shift
struct NttTables {
    const uint64_t *roots_f;
    const uint64_t *roots_i;
    const uint64_t *twl_f, *twh_f, *twl_i, *twh_i;
    const uint64_t *shl, *shh;
};

__device__ __forceinline__ uint64_t tw_pow(const uint64_t *lo, const uint64_t *hi, uint32_t x)
{
    uint64_t a = lo[x & 4095u];
    uint32_t h = x >> 12;
    return h ? gl_mul(a, hi[h]) : a;
}

// ------------------------------------------------------------ transpose ----
// in: rows x cols row-major (row stride in_rs), out: column c at out + c*out_cs.
#define TR_T 32
__global__ __launch_bounds__(256) void k_transpose_rm2cm(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                         uint64_t rows, uint64_t cols, uint64_t in_rs, uint64_t out_cs)
{
    __shared__ uint64_t tile[TR_T][TR_T + 1];
    uint64_t r0 = (uint64_t)blockIdx.x * TR_T, c0 = (uint64_t)blockIdx.y * TR_T;
    uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (uint32_t k = ty; k < TR_T; k += 8) {
        uint64_t r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[r * in_rs + c];
    }
    __syncthreads();
    for (uint32_t k = ty; k < TR_T; k += 8) {
        uint64_t c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) out[c * out_cs + r] = tile[tx][k];
    }
}

// column-major (col c at in + c*in_cs) -> row-major rows x cols (row stride out_rs)
__global__ __launch_bounds__(256) void k_transpose_cm2rm(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                         uint64_t rows, uint64_t cols, uint64_t in_cs, uint64_t out_rs)
{
    __shared__ uint64_t tile[TR_T][TR_T + 1];
    uint64_t r0 = (uint64_t)blockIdx.x * TR_T, c0 = (uint64_t)blockIdx.y * TR_T;
    uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (uint32_t k = ty; k < TR_T; k += 8) {
        uint64_t c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[c * in_cs + r];
    }
    __syncthreads();
    for (uint32_t k = ty; k < TR_T; k += 8) {
        uint64_t r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) out[r * out_rs + c] = tile[tx][k];
    }
}

// ---------------------------------------------------------- LDS stages ----
// Sub-FFT over "m" inside an LDS tile laid out [m][ti] (ti fastest), n points,
// tcount independent lines. DIF: natural in -> bit-reversed out. DIT: reverse.
template <bool INV>
__device__ __forceinline__ void lds_dif(uint64_t *s, uint32_t logn, uint32_t tcount, uint32_t log_t,
                                        const uint64_t *roots)
{
    uint32_t n = 1u << logn;
    uint32_t nbf = (n >> 1) << log_t;
    for (int t = (int)logn - 1; t >= 0; t--) {
        uint32_t h = 1u << t;
        for (uint32_t b = threadIdx.x; b < nbf; b += blockDim.x) {
            uint32_t ii = b & (tcount - 1), j = b >> log_t;
            uint32_t jj = j & (h - 1);
            uint32_t p0 = ((j >> t) << (t + 1)) + jj;
            uint32_t a0 = (p0 << log_t) + ii, a1 = ((p0 + h) << log_t) + ii;
            uint64_t u = s[a0], v = s[a1];
            s[a0] = gl_add(u, v);
            uint64_t d = gl_sub(u, v);
            s[a1] = jj ? gl_mul(d, roots[h + jj]) : d;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void lds_dit(uint64_t *s, uint32_t first_t, uint32_t logn, uint32_t tcount, uint32_t log_t,
                                        const uint64_t *roots)
{
    uint32_t n = 1u << logn;
    uint32_t nbf = (n >> 1) << log_t;
    for (uint32_t t = first_t; t < logn; t++) {
        uint32_t h = 1u << t;
        for (uint32_t b = threadIdx.x; b < nbf; b += blockDim.x) {
            uint32_t ii = b & (tcount - 1), j = b >> log_t;
            uint32_t jj = j & (h - 1);
            uint32_t p0 = ((j >> t) << (t + 1)) + jj;
            uint32_t a0 = (p0 << log_t) + ii, a1 = ((p0 + h) << log_t) + ii;
            uint64_t u = s[a0], v = s[a1];
            if (jj) v = gl_mul(v, roots[h + jj]);
            s[a0] = gl_add(u, v);
            s[a1] = gl_sub(u, v);
        }
        __syncthreads();
    }
}

// -------------------------------------------------------- strided passes ----
// Column c = blockIdx.y at data + c*cs; length n = n1 * n2 (log_n = log_n1 + log_n2).
// A block owns the n1 points {m*n2 + i : m} for ti consecutive i (i0 .. i0+ti).
// INV = true : DIF with inverse roots, then x w_n^-(i * bitrev(m))  (iNTT pass 1)
// INV = false: DIT with forward roots on bit-reversed m, no twiddle (NTT pass 4)
template <bool INV>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_strided(uint64_t *__restrict__ data, uint64_t cs, uint32_t log_n1,
                                                             uint32_t log_n2, uint32_t log_t, NttTables tb)
{
    extern __shared__ uint64_t s[];
    uint32_t ti = 1u << log_t, n1 = 1u << log_n1;
    uint64_t n2 = 1ull << log_n2;
    uint64_t i0 = (uint64_t)blockIdx.x << log_t;
    uint64_t *col = data + (uint64_t)blockIdx.y * cs;
    uint32_t tile = n1 << log_t;
    for (uint32_t x = threadIdx.x; x < tile; x += blockDim.x) {
        uint32_t m = x >> log_t, ii = x & (ti - 1);
        s[x] = col[(uint64_t)m * n2 + i0 + ii];
    }
    __syncthreads();
    if (INV) {
        lds_dif<true>(s, log_n1, ti, log_t, tb.roots_i);
        uint32_t shift = 24 - (log_n1 + log_n2);  // w_n = w_{2^24}^(2^shift)
        for (uint32_t x = threadIdx.x; x < tile; x += blockDim.x) {
            uint32_t m = x >> log_t, ii = x & (ti - 1);
            uint32_t e = (uint32_t)((i0 + ii) * bitrev32(m, log_n1));
            uint64_t v = s[x];
            if (e) v = gl_mul(v, tw_pow(tb.twl_i, tb.twh_i, e << shift));
            col[(uint64_t)m * n2 + i0 + ii] = v;
        }
    } else {
        lds_dit(s, 0, log_n1, ti, log_t, tb.roots_f);
        for (uint32_t x = threadIdx.x; x < tile; x += blockDim.x) {
            uint32_t m = x >> log_t, ii = x & (ti - 1);
            col[(uint64_t)m * n2 + i0 + ii] = s[x];
        }
    }
}

// ------------------------------------------------------ fused LDE block ----
// Block (m = blockIdx.x, column c = blockIdx.y):
//   load n2 values of the half-transformed column (scratch, column stride cs_in)
//   DIF over i with inverse roots -> coefficients in bit-reversed order,
//   scale coefficient e = bitrev_L(m*n2 + q) by 7^e / N,
//   write each into B consecutive slots, DIT stages log2(B) .. log2(n2*B)-1,
//   multiply by w_NExt^(bitrev_{log n1}(m) * k) and store block m of the
//   column-major LDE (column stride cs_out).
__global__ __launch_bounds__(NTT_THREADS) void k_lde_block(const uint64_t *__restrict__ in, uint64_t cs_in,
                                                           uint64_t *__restrict__ out, uint64_t cs_out,
                                                           uint32_t log_n1, uint32_t log_n2, uint32_t log_b,
                                                           uint64_t n_inv, NttTables tb)
{
    extern __shared__ uint64_t s[];
    uint32_t n2 = 1u << log_n2, nb = n2 << log_b, m = blockIdx.x;
    const uint64_t *src = in + (uint64_t)blockIdx.y * cs_in + ((uint64_t)m << log_n2);
    for (uint32_t q = threadIdx.x; q < n2; q += blockDim.x) s[q] = src[q];
    __syncthreads();
    lds_dif<true>(s, log_n2, 1, 0, tb.roots_i);
    uint32_t log_n = log_n1 + log_n2;
    // scale and spread: thread-private registers hold the values between syncs
    uint64_t v[NTT_LDS_ELEMS / NTT_THREADS];
    uint32_t k = 0;
    for (uint32_t q = threadIdx.x; q < n2; q += blockDim.x, k++) {
        uint32_t e = bitrev32((m << log_n2) + q, log_n);
        v[k] = gl_mul(gl_mul(s[q], n_inv), tw_pow(tb.shl, tb.shh, e));
    }
    __syncthreads();
    k = 0;
    for (uint32_t q = threadIdx.x; q < n2; q += blockDim.x, k++)
        for (uint32_t r = 0; r < (1u << log_b); r++) s[(q << log_b) + r] = v[k];
    __syncthreads();
    lds_dit(s, log_b, log_n2 + log_b, 1, 0, tb.roots_f);
    uint64_t *dst = out + (uint64_t)blockIdx.y * cs_out + ((uint64_t)m << (log_n2 + log_b));
    if (log_n1 == 0) {
        for (uint32_t x = threadIdx.x; x < nb; x += blockDim.x) dst[x] = s[x];
    } else {
        uint32_t e1 = bitrev32(m, log_n1);
        uint32_t shift = 24 - (log_n + log_b);
        for (uint32_t x = threadIdx.x; x < nb; x += blockDim.x) {
            uint32_t e = e1 * x;
            uint64_t val = s[x];
            if (e) val = gl_mul(val, tw_pow(tb.twl_f, tb.twh_f, e << shift));
            dst[x] = val;
        }
    }
}

// ------------------------------------------------------------ launchers ----
extern "C" int ntt_launch_lde(uint64_t *lde, const uint64_t *trace, uint64_t *scratch, uint32_t log_n, uint32_t log_ext,
                              uint64_t ncols, const NttTablesHost *th, hipStream_t st, hipEvent_t *ev)
{
    // ev (optional): events recorded after passes 1..4 (ev[0..3]) for phase timing
    NttTables tb = {th->roots_f, th->roots_i, th->twl_f, th->twh_f, th->twl_i, th->twh_i, th->shl, th->shh};
    uint64_t N = 1ull << log_n;
    uint32_t log_b = log_ext - log_n;
    // plan: n2 * B <= LDS_ELEMS, n1 = N / n2
    uint32_t log_lds = NTT_LOG_LDS;
    if (log_b > log_lds) return -2;
    uint32_t log_n2 = log_n;
    if (log_n2 + log_b > log_lds) log_n2 = log_lds > log_b ? log_lds - log_b : 0;
    uint32_t log_n1 = log_n - log_n2;
    if (log_n1 > log_lds) return -2;
    if (log_ext > 24) return -3;

    // 1. row-major trace -> column-major scratch
    dim3 tg((unsigned)((N + TR_T - 1) / TR_T), (unsigned)((ncols + TR_T - 1) / TR_T));
    hipLaunchKernelGGL(k_transpose_rm2cm, tg, dim3(256), 0, st, trace, scratch, N, ncols, ncols, N);
    if (ev) hipEventRecord(ev[0], st);
    // 2. strided inverse DIF
    if (log_n1 > 0) {
        uint32_t log_t = log_lds - log_n1;
        if (log_t > log_n2) log_t = log_n2;
        dim3 g((unsigned)(1ull << (log_n2 - log_t)), (unsigned)ncols);
        size_t sh = (size_t)8 << (log_n1 + log_t);
        hipLaunchKernelGGL(k_ntt_strided<true>, g, dim3(NTT_THREADS), sh, st, scratch, N, log_n1, log_n2, log_t, tb);
    }
    if (ev) hipEventRecord(ev[1], st);
    // 3. fused block: DIF, coset scale, extend, DIT, twiddle
    uint64_t n_inv = gl_inv(N % GL_P);
    {
        dim3 g((unsigned)(1u << log_n1), (unsigned)ncols);
        size_t sh = (size_t)8 << (log_n2 + log_b);
        hipLaunchKernelGGL(k_lde_block, g, dim3(NTT_THREADS), sh, st, scratch, N, lde, N << log_b, log_n1, log_n2,
                           log_b, n_inv, tb);
    }
    if (ev) hipEventRecord(ev[2], st);
    // 4. strided forward DIT
    if (log_n1 > 0) {
        uint32_t log_n2b = log_n2 + log_b;
        uint32_t log_t = log_lds - log_n1;
        if (log_t > log_n2b) log_t = log_n2b;
        dim3 g((unsigned)(1ull << (log_n2b - log_t)), (unsigned)ncols);
        size_t sh = (size_t)8 << (log_n1 + log_t);
        hipLaunchKernelGGL(k_ntt_strided<false>, g, dim3(NTT_THREADS), sh, st, lde, N << log_b, log_n1, log_n2b,
                           log_t, tb);
    }
    if (ev) hipEventRecord(ev[3], st);
    return (int)hipGetLastError();
}

extern "C" int ntt_launch_cm2rm(uint64_t *out, const uint64_t *in, uint64_t rows, uint64_t cols, uint64_t in_cs,
                                uint64_t out_rs, hipStream_t st)
{
    dim3 tg((unsigned)((rows + TR_T - 1) / TR_T), (unsigned)((cols + TR_T - 1) / TR_T));
    hipLaunchKernelGGL(k_transpose_cm2rm, tg, dim3(256), 0, st, in, out, rows, cols, in_cs, out_rs);
    return (int)hipGetLastError();
}

extern "C" int ntt_launch_rm2cm(uint64_t *out, const uint64_t *in, uint64_t rows, uint64_t cols, uint64_t in_rs,
                                uint64_t out_cs, hipStream_t st)
{
    dim3 tg((unsigned)((rows + TR_T - 1) / TR_T), (unsigned)((cols + TR_T - 1) / TR_T));
    hipLaunchKernelGGL(k_transpose_rm2cm, tg, dim3(256), 0, st, in, out, rows, cols, in_rs, out_cs);
    return (int)hipGetLastError();
}

// ------------------------------------------------------- synthetic data ----
// Seeded uniform values in [0, p): element i = mix(seed + (i+1)*0x9E3779B97F4A7C15),
// re-mixed while >= p (counter-based SplitMix64; the numpy generator matches).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_fill_uniform(uint64_t *__restrict__ out, uint64_t n, uint64_t seed,
                                                      uint64_t start)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t z = mix64(seed + (start + i + 1) * 0x9E3779B97F4A7C15ULL);
        while (z >= GL_P) z = mix64(z ^ 0xD1B54A32D192ED03ULL);
        out[i] = z;
    }
}

extern "C" int ntt_launch_fill_uniform(uint64_t *out, uint64_t n, uint64_t seed, uint64_t start, hipStream_t st)
{
    if (n == 0) return 0;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)blocks), dim3(256), 0, st, out, n, seed, start);
    return (int)hipGetLastError();
}
