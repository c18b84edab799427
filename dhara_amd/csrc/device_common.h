// device_common.h -- device helpers shared by the gfx950 kernels.
//
//  * Mualem-van Genuchten constitutive law in the stable exp/log form (one log, one log1p,
//    three exp/expm1 per cell instead of the reference's seven pow, src/global.cu:66-105).
//  * wave64 shuffle reductions, block reductions through LDS, and a deterministic
//    last-block-done grid reduction (fixed summation order, so results are identical from
//    run to run).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dh {

constexpr int kBlock = 256;        // threads per block of the pointwise kernels
constexpr int kWave = 64;          // CDNA wavefront
constexpr int kMaxBlocks = 4096;   // capacity of the per-block partial buffers
constexpr int kMaxSums = 12;       // values reduced together at most

// Solver state of one Krylov solve, resident in device memory.  Written by the last block
// of each reduction, read by the next kernel (and by the host after a batch).
struct PcgState {
    double rz;        // (r, M^-1 r)
    double pq;        // (p, A p)
    double rr;        // ||r||^2
    double bb;        // ||b||^2 of the outer system
    double alpha, beta;
    double tol2;      // stop when rr <= tol2
    double rr_outer;  // ||r_64||^2 of the last FP64 residual (mixed precision)
    int32_t iter;     // iterations of the current solve
    int32_t maxit;
    int32_t done;     // 1: solve finished (converged, limit, or breakdown)
    int32_t conv;     // 1: converged
    int32_t nan;
    int32_t total;    // iterations summed over the refinement sweeps of a Picard iteration
    int32_t outer_conv;
    int32_t pad;
};

struct PicardState {
    double delta;     // max over cells of the Picard increment (rule units)
    int32_t nan;
    int32_t pad;
};

// ---- Mualem-van Genuchten ------------------------------------------------------------
// theta(psi), K(psi), C = dtheta/dpsi with x = alpha|psi|, t = x^n, m = 1 - 1/n:
//   Se = (1+t)^-m, K = Ks sqrt(Se) (1 - (t/(1+t))^m)^2, C = (ts-tr)(n-1) alpha x^(n-1) (1+t)^(1/n-2)
// 1 - (t/(1+t))^m is evaluated as -expm1(-m log1p(1/t)), which keeps full relative accuracy
// near saturation where the reference's 1 - pow(1 - pow(Se, 1/m), m) cancels.
__device__ __forceinline__ void vg_eval(double psi, double Ks, double alpha, double n, double ts,
                                        double tr, double& theta, double& K, double& C)
{
    if (!(psi < 0.0)) {
        if (psi != psi) { theta = psi; K = psi; C = psi; return; }
        theta = ts; K = Ks; C = 0.0;
        return;
    }
    const double m = (n - 1.0) / n;
    const double lx = log(alpha * (-psi));
    const double t = exp(n * lx);
    const double l1t = log1p(t);
    const double Se = exp(-m * l1t);
    theta = (ts - tr) * Se + tr;
    const double w = (t > 0.0) ? -expm1(-m * log1p(1.0 / t)) : 1.0;
    K = Ks * sqrt(Se) * w * w;
    C = (ts - tr) * (n - 1.0) * alpha * exp((n - 1.0) * lx + (1.0 / n - 2.0) * l1t);
}

// vanGenuchtenInverse, src/global.cu:32-52 (SI).
__device__ __forceinline__ double vg_inverse(double theta, double alpha, double n, double ts,
                                             double tr)
{
    const double m = (n - 1.0) / n;
    if (theta < ts) return -(1.0 / alpha) * pow(pow((ts - tr) / (theta - tr), 1.0 / m) - 1.0, 1.0 / n);
    return 0.0;
}

// ---- reductions ---------------------------------------------------------------------------
__device__ __forceinline__ int lin_tid() { return threadIdx.x + threadIdx.y * blockDim.x; }
__device__ __forceinline__ int lin_nthreads() { return blockDim.x * blockDim.y; }
__device__ __forceinline__ int lin_bid() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ __forceinline__ int lin_nblocks() { return gridDim.x * gridDim.y; }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Block reduction of NV values; the result is valid in every lane of wave 0.
// lds must hold NV * 16 doubles.
template <int NV, bool MAX>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* lds)
{
    const int tid = lin_tid();
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    const int nw = lin_nthreads() / kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = MAX ? wave_max(v[i]) : wave_sum(v[i]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) lds[i * 16 + wid] = v[i];
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double t = (lane < nw) ? lds[i * 16 + lane] : (MAX ? -INFINITY : 0.0);
            v[i] = MAX ? wave_max(t) : wave_sum(t);
        }
    }
}

__device__ __forceinline__ void store_agent(double* p, double v)
{
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double load_agent(const double* p)
{
    unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)u);
}

// Grid-wide reduction of NV values with the last-block-done pattern.  Every block writes
// its partials, the block that arrives last sums all partials in block order (so the result
// does not depend on scheduling) and returns true with the totals in `out` (valid in
// thread 0).  The arrival counter is reset by that block for the next launch.
template <int NV, bool MAX>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], double* part, unsigned* counter,
                                            double (&out)[NV])
{
    __shared__ double lds[NV * 16];
    __shared__ int is_last;
    block_reduce<NV, MAX>(v, lds);
    const int tid = lin_tid();
    const int bid = lin_bid(), nb = lin_nblocks();
    if (tid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) store_agent(part + i * kMaxBlocks + bid, v[i]);
        __threadfence();
        const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == (unsigned)(nb - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = MAX ? -INFINITY : 0.0;
    // fixed order: thread t sums blocks t, t+T, t+2T, ...; then a fixed tree
    for (int b = tid; b < nb; b += lin_nthreads()) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double x = load_agent(part + i * kMaxBlocks + b);
            acc[i] = MAX ? fmax(acc[i], x) : acc[i] + x;
        }
    }
    block_reduce<NV, MAX>(acc, lds);
    if (tid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) out[i] = acc[i];
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

}  // namespace dh
