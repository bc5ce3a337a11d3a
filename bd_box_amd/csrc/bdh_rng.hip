// Device reproduction of the reference's Gaussian stream.
//
// The reference draws its noise on the host: glibc drand48_r (48-bit LCG, a = 0x5DEECE66D,
// c = 0xB, value X / 2^48) feeding the polar Box-Muller method of gauss_rand
// (src/rand_move.c:132-168): attempt t consumes two uniforms, v = 2u - 1, accepts when
// 0 < v1^2 + v2^2 < 1, and an accepted pair yields (v2*fac, v1*fac) at table positions
// (2p, 2p+1); for an odd table length the last v1*fac is dropped (the pair state is local
// to the call).
//
// Here one workgroup owns one stream.  Each round, thread t jumps the LCG ahead to its
// chunk of K attempts (jump constants A^(2^b), C_b for b < 48 live in constant memory),
// evaluates them, and a workgroup scan of the accept counts gives every accepted attempt
// its pair index p.  The stream state after the call is the state after the attempt that
// produced the last needed pair -- so the integer stream is bit-identical to the host's,
// and repeated calls continue it exactly as repeated gauss_rand calls do.
// v1*v1 + v2*v2 and -2 log(rsq) / rsq are evaluated without FMA contraction, as the
// x86-64 reference does; the device log differs
// from glibc's by an ulp now and then, so Gaussian values agree to ~1e-15 while the
// integer stream (and so which attempts are accepted) is identical.
#include "bdh_common.hpp"
#include "bdh_kernels.hpp"

namespace bdh {

namespace {

constexpr uint64_t LCG_A = 0x5DEECE66DULL;
constexpr uint64_t LCG_C = 0xBULL;
constexpr uint64_t MASK48 = (1ULL << 48) - 1;

__constant__ uint64_t c_jumpA[48];
__constant__ uint64_t c_jumpC[48];

constexpr int RNG_THREADS = 1024;
constexpr int RNG_K = 8;  // attempts per thread per round

__device__ __forceinline__ uint64_t lcg_step(uint64_t x) { return (LCG_A * x + LCG_C) & MASK48; }

__device__ __forceinline__ uint64_t lcg_jump(uint64_t x, uint64_t k)
{
    for (int b = 0; k; ++b, k >>= 1)
        if (k & 1) x = (c_jumpA[b] * x + c_jumpC[b]) & MASK48;
    return x;
}

__device__ __forceinline__ double u48(uint64_t x) { return (double)x * 0x1p-48; }

// log(r) for r in (0.5, 1) with small relative error as log(r) -> 0:
// log(1 + d) = 2 atanh(s), s = d / (2 + d), d = r - 1 exact (Sterbenz); |s| < 1/3 so the
// odd series 2 s (1 + s^2/3 + s^4/5 + ...) is summed to 20 terms (s^40/41 < 1e-20).
__device__ __forceinline__ double log_near_one(double r)
{
#pragma clang fp contract(off)
    const double d = r - 1.0;
    const double s = d / (2.0 + d);
    const double s2 = s * s;
    double p = 1.0 / 41.0;
#pragma unroll
    for (int k = 19; k >= 0; --k) p = 1.0 / (2 * k + 1) + s2 * p;
    return (2.0 * s) * p;
}

__global__ void __launch_bounds__(RNG_THREADS)
gauss_streams_kernel(uint64_t* __restrict__ states, double* __restrict__ out,
                     const long long* __restrict__ offsets, const int* __restrict__ counts,
                     int* __restrict__ status)
{
    // no FMA contraction in this function (the pragma covers only expressions written
    // here, not inlined helpers): v1*v1 + v2*v2 must round exactly as on the host
#pragma clang fp contract(off)
    __shared__ int s_scan[RNG_THREADS];
    __shared__ long long s_last;  // index (within the round) of the attempt giving the last pair
    const int s = blockIdx.x;
    const int t = threadIdx.x;
    const int count = counts[s];
    double* o = out + offsets[s];
    const long long npairs = (count + 1) / 2;
    uint64_t base = states[s];
    long long produced = 0;
    const long long per_round = (long long)RNG_THREADS * RNG_K;

    for (int round = 0; produced < npairs; ++round) {
        if (round > 4096) {  // cannot happen (acceptance pi/4); keeps the loop finite
            if (t == 0) status[0] = 1;
            break;
        }
        // this thread's K attempts
        uint64_t x = lcg_jump(base, 2ull * (uint64_t)t * RNG_K);
        double v1[RNG_K], v2[RNG_K], rsq[RNG_K];
        int acc = 0;
#pragma unroll
        for (int q = 0; q < RNG_K; ++q) {
            x = lcg_step(x);
            v1[q] = 2.0 * u48(x) - 1.0;
            x = lcg_step(x);
            v2[q] = 2.0 * u48(x) - 1.0;
            rsq[q] = v1[q] * v1[q] + v2[q] * v2[q];
            acc += (rsq[q] < 1.0 && rsq[q] != 0.0);
        }
        // inclusive scan of accept counts over the workgroup
        s_scan[t] = acc;
        __syncthreads();
        for (int off = 1; off < RNG_THREADS; off <<= 1) {
            int v = (t >= off) ? s_scan[t - off] : 0;
            __syncthreads();
            s_scan[t] += v;
            __syncthreads();
        }
        const int total = s_scan[RNG_THREADS - 1];
        long long p = produced + (s_scan[t] - acc);
        if (t == 0) s_last = -1;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RNG_K; ++q) {
            if (rsq[q] < 1.0 && rsq[q] != 0.0) {
                if (p < npairs) {
                    // near 1 the library log is accurate only in absolute terms; the series
                    // keeps the relative error at the ulp level where log(rsq) -> 0
                    const double lg = rsq[q] > 0.5 ? log_near_one(rsq[q]) : log(rsq[q]);
                    double fac = sqrt(-2.0 * lg / rsq[q]);
                    o[2 * p] = v2[q] * fac;
                    if (2 * p + 1 < count) o[2 * p + 1] = v1[q] * fac;
                    if (p == npairs - 1) s_last = (long long)t * RNG_K + q;
                }
                ++p;
            }
        }
        __syncthreads();
        if (produced + total >= npairs) {
            base = lcg_jump(base, 2ull * (uint64_t)(s_last + 1));
            produced = npairs;
        } else {
            base = lcg_jump(base, 2ull * (uint64_t)per_round);
            produced += total;
        }
        __syncthreads();
    }
    if (t == 0) states[s] = base;
}

bool g_jumps_ready[64] = {false};

int ensure_jumps()
{
    int dev = 0;
    BDH_HIP(hipGetDevice(&dev));
    if (dev >= 0 && dev < 64 && g_jumps_ready[dev]) return 0;
    uint64_t A[48], C[48];
    A[0] = LCG_A;
    C[0] = LCG_C;
    for (int b = 1; b < 48; ++b) {  // J_{b} = J_{b-1} o J_{b-1}
        A[b] = (A[b - 1] * A[b - 1]) & MASK48;
        C[b] = (A[b - 1] * C[b - 1] + C[b - 1]) & MASK48;
    }
    BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_jumpA), A, sizeof(A)));
    BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_jumpC), C, sizeof(C)));
    if (dev >= 0 && dev < 64) g_jumps_ready[dev] = true;
    return 0;
}

}  // namespace

uint64_t lcg_seed(uint32_t seed) { return (((uint64_t)seed) << 16) | 0x330Eu; }

int gauss_prepare() { return ensure_jumps(); }

int launch_gauss_streams(uint64_t* d_states, double* d_out, const long long* d_offsets,
                         const int* d_counts, int nstreams, int* d_status, hipStream_t st)
{
    if (nstreams <= 0) return 0;
    if (ensure_jumps()) return -1;
    hipLaunchKernelGGL(gauss_streams_kernel, dim3(nstreams), dim3(RNG_THREADS), 0, st,
                       d_states, d_out, d_offsets, d_counts, d_status);
    BDH_LAUNCH_CHECK();
    return 0;
}

}  // namespace bdh
