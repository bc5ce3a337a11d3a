// Device reproduction of the reference's Gaussian stream.
//
// The reference draws its noise on the host: glibc drand48_r (48-bit LCG, a = 0x5DEECE66D,
// c = 0xB, value X / 2^48) feeding the polar Box-Muller method of gauss_rand
// (src/rand_move.c:132-168): attempt t consumes two uniforms, v = 2u - 1, accepts when
// 0 < v1^2 + v2^2 < 1, and an accepted pair yields (v2*fac, v1*fac) at table positions
// (2p, 2p+1); for an odd table length the last v1*fac is dropped (the pair state is local
// to the call).
//
// gauss_streams_kernel: one workgroup owns one stream (the batch's chains, and the remainder
// of a single stream's call below).  Each round, thread t jumps the LCG ahead to its chunk of
// K attempts (jump constants A^(2^b), C_b for b < 48 live in constant memory),
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

// ------------------------------------------------------------------ one stream, many workgroups
// A single long stream (the m steps' noise of one system) in four short launches instead of
// one workgroup's rounds: (1) workgroup w evaluates attempts [w GW_ATT, (w + 1) GW_ATT) of the
// call and counts its accepts; (2) one thread turns the counts into each workgroup's first
// pair index; (3) every workgroup re-evaluates its attempts and writes the accepted pairs
// below the needed count, the thread of the last needed pair setting the stream state after
// its attempt; (4) gauss_streams_kernel finishes the call in the (never observed) case that
// the attempts evaluated held too few accepts, and does nothing otherwise.  The values and the
// final state are those of gauss_streams_kernel, attempt for attempt.
constexpr int GW_THREADS = 256;
constexpr int GW_K = 8;                        // attempts per thread
constexpr int GW_ATT = GW_THREADS * GW_K;      // attempts per workgroup

struct GaussPlan {
    uint64_t base;      // stream state at the start of the call
    long long out_off;  // continuation: index of its first value
    int remaining;      // continuation: values left to draw (0: none)
    int pad;
};

// The GW_K attempts of thread g of the call: uniforms, rsq, accept count.
__device__ __forceinline__ int gw_attempts(uint64_t base, long long g, double (&v1)[GW_K], double (&v2)[GW_K],
                                           double (&rsq)[GW_K])
{
#pragma clang fp contract(off)
    uint64_t x = lcg_jump(base, 2ull * (uint64_t)g * GW_K);
    int acc = 0;
#pragma unroll
    for (int q = 0; q < GW_K; ++q) {
        x = lcg_step(x);
        v1[q] = 2.0 * u48(x) - 1.0;
        x = lcg_step(x);
        v2[q] = 2.0 * u48(x) - 1.0;
        rsq[q] = v1[q] * v1[q] + v2[q] * v2[q];
        acc += (rsq[q] < 1.0 && rsq[q] != 0.0);
    }
    return acc;
}

__global__ void __launch_bounds__(GW_THREADS)
gauss_wide_count_kernel(const uint64_t* __restrict__ state, int* __restrict__ wg_count)
{
    __shared__ int s_red[GW_THREADS / 64];
    double v1[GW_K], v2[GW_K], rsq[GW_K];
    int acc = gw_attempts(*state, (long long)blockIdx.x * GW_THREADS + threadIdx.x, v1, v2, rsq);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
#pragma unroll
        for (int w = 0; w < GW_THREADS / 64; ++w) sum += s_red[w];
        wg_count[blockIdx.x] = sum;
    }
}

// wg_count[w] <- pairs before workgroup w; the continuation when all G workgroups' accepts
// fall short of the call (the stream state then moves past every attempt evaluated).
__global__ void gauss_wide_plan_kernel(uint64_t* __restrict__ state, int* __restrict__ wg_count, int G, int count,
                                       GaussPlan* __restrict__ plan)
{
    if (threadIdx.x != 0) return;
    const uint64_t base = *state;
    const long long npairs = (count + 1) / 2;
    long long run = 0;
    for (int w = 0; w < G; ++w) {
        const int c = wg_count[w];
        wg_count[w] = (int)run;
        run += c;
    }
    plan->base = base;
    if (run >= npairs) {
        plan->remaining = 0;
        plan->out_off = 0;
    } else {
        plan->remaining = count - (int)(2 * run);
        plan->out_off = 2 * run;
        *state = lcg_jump(base, 2ull * (uint64_t)G * GW_ATT);
    }
}

__global__ void __launch_bounds__(GW_THREADS)
gauss_wide_write_kernel(uint64_t* __restrict__ state, double* __restrict__ out, const int* __restrict__ wg_base,
                        int count, const GaussPlan* __restrict__ plan)
{
#pragma clang fp contract(off)
    __shared__ int s_scan[GW_THREADS];
    const int t = threadIdx.x;
    const uint64_t base = plan->base;
    const long long npairs = (count + 1) / 2;
    const long long g = (long long)blockIdx.x * GW_THREADS + t;
    double v1[GW_K], v2[GW_K], rsq[GW_K];
    const int acc = gw_attempts(base, g, v1, v2, rsq);
    s_scan[t] = acc;
    __syncthreads();
    for (int off = 1; off < GW_THREADS; off <<= 1) {
        const int v = (t >= off) ? s_scan[t - off] : 0;
        __syncthreads();
        s_scan[t] += v;
        __syncthreads();
    }
    long long p = wg_base[blockIdx.x] + (s_scan[t] - acc);
#pragma unroll
    for (int q = 0; q < GW_K; ++q) {
        if (rsq[q] < 1.0 && rsq[q] != 0.0) {
            if (p < npairs) {
                const double lg = rsq[q] > 0.5 ? log_near_one(rsq[q]) : log(rsq[q]);
                double fac = sqrt(-2.0 * lg / rsq[q]);
                out[2 * p] = v2[q] * fac;
                if (2 * p + 1 < count) out[2 * p + 1] = v1[q] * fac;
                if (p == npairs - 1) *state = lcg_jump(base, 2ull * (uint64_t)(g * GW_K + q + 1));
            }
            ++p;
        }
    }
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

int gauss_wide_workgroups(long long count)
{
    // attempts: 1.34 per pair (acceptance pi / 4 = 0.785) plus a margin of 2048
    const long long need = (long long)(1.34 * (double)((count + 1) / 2)) + 2048;
    return (int)((need + GW_ATT - 1) / GW_ATT);
}

size_t gauss_wide_scratch_bytes(long long max_count)
{
    return sizeof(GaussPlan) + sizeof(int) * (size_t)gauss_wide_workgroups(max_count);
}

int launch_gauss_wide(uint64_t* d_state, double* d_out, int count, void* d_scratch, size_t scratch_bytes,
                      int* d_status, hipStream_t st)
{
    if (count <= 0) return 0;
    if (ensure_jumps()) return -1;
    const int G = gauss_wide_workgroups(count);
    BDH_CHECK(gauss_wide_scratch_bytes(count) <= scratch_bytes, "gauss: scratch too small for the count");
    GaussPlan* plan = static_cast<GaussPlan*>(d_scratch);
    int* wg = reinterpret_cast<int*>(plan + 1);
    hipLaunchKernelGGL(gauss_wide_count_kernel, dim3(G), dim3(GW_THREADS), 0, st, (const uint64_t*)d_state, wg);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(gauss_wide_plan_kernel, dim3(1), dim3(64), 0, st, d_state, wg, G, count, plan);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(gauss_wide_write_kernel, dim3(G), dim3(GW_THREADS), 0, st, d_state, d_out, (const int*)wg,
                       count, (const GaussPlan*)plan);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(gauss_streams_kernel, dim3(1), dim3(RNG_THREADS), 0, st, d_state, d_out,
                       (const long long*)&plan->out_off, (const int*)&plan->remaining, d_status);
    BDH_LAUNCH_CHECK();
    return 0;
}

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
