// device_common.h -- device helpers shared by the gfx950 kernels.
//
//  * Mualem-van Genuchten constitutive law in the stable exp/log form (one log, one log1p,
//    three exp/expm1 per cell instead of the reference's seven pow, src/global.cu:66-105).
//  * wave64 shuffle reductions, block reductions through LDS, and fence-free grid
//    reductions: per-block partials summed by the consumer in a fixed order, so results are
//    identical from run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dh {

constexpr int kBlock = 256;        // threads per block of the pointwise kernels
constexpr int kWave = 64;          // CDNA wavefront
constexpr int kMaxBlocks = 4096;   // capacity of the per-block partial buffers
constexpr int kMaxSums = 12;       // values reduced together at most

// Solver state of one Krylov solve, resident in device memory.  Fields a kernel reads while
// one of its blocks may write are double-buffered by iteration parity (rz), the rest are
// written only by the kernel that ends the solve.  The host reads it after a batch.
struct PcgState {
    double rz[2];     // (r_i, M^-1 r_i), slot i & 1, written by the SpMV kernel of iteration i
    double bb;        // ||b||^2 of the outer system
    double tol2;      // stop when ||r||^2 <= tol2
    double rr;        // ||r||^2 when the solve ended
    double rr_outer;  // ||r_64||^2 of the last FP64 residual (mixed precision)
    double rr_update; // ||r_64||^2 after the last gated reliable update (k_refine_conv)
    double pi[2];     // one-kernel PCG: (p_i, A p_i), slot i & 1, written at the start of iteration i
    int32_t iter;     // iterations done when the solve ended
    int32_t done;     // 1: solve finished (converged, limit, or breakdown)
    int32_t conv;     // 1: converged
    int32_t nan;
    int32_t outer_conv;
    int32_t gate;     // 1: the solve segment ended without convergence or NaN (k_refine_gate);
                      // 2: and the update after it passed the outer test (k_refine_conv)
    int32_t pad[2];
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
    const double x = alpha * (-psi);
    const double lx = log(x);
    const double t = exp(n * lx);
    const double l1t = log1p(t);
    const double Se = exp(-m * l1t);
    theta = (ts - tr) * Se + tr;
    // log1p(1/t) = log1p(t) - log(t); the difference loses ~log(t) t ulps of log1p(1/t), so it
    // is used only for t < 64 (relative error < 1e-13), the direct form beyond
    const double l1it = (t < 64.0) ? l1t - n * lx : log1p(1.0 / t);
    const double w = (t > 0.0) ? -expm1(-m * l1it) : 1.0;
    K = Ks * sqrt(Se) * w * w;
    // dtheta/dpsi = (ts - tr) (n - 1) alpha x^(n-1) (1 + t)^(1/n - 2), with x^(n-1) = t / x and
    // (1 + t)^(1/n - 2) = Se / (1 + t) since Se = (1 + t)^(1/n - 1)
    // (t overflowing to infinity -- |psi| beyond any soil -- gives C = 0, the limit)
    C = (t < 1e300) ? (ts - tr) * (n - 1.0) * alpha * t * Se / (x * (1.0 + t)) : 0.0;
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
// No kernel communicates across blocks: a producer kernel writes one partial per block with a
// plain store, and the consumer -- the next kernel, after the kernel boundary has made the
// stores visible on every XCD, or the host -- sums the partials in block order.  Every
// consuming block runs the same fixed-order sum, so all of them see identical totals and the
// result does not depend on scheduling.  (A last-block-done reduction would need an
// agent-scope release fence per block, i.e. an L2 writeback per block on a multi-XCD part.)
__device__ __forceinline__ int lin_tid() { return threadIdx.x + threadIdx.y * blockDim.x; }
__device__ __forceinline__ int lin_nthreads() { return blockDim.x * blockDim.y; }
__device__ __forceinline__ int lin_bid() { return blockIdx.x + blockIdx.y * gridDim.x; }

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

// The block a kernel body works as: its index among the blocks of the launch and their
// number.  A single-model launch uses the hardware values; a batched launch gives every model
// its own share of blocks (id 0..n-1 within that model).
struct Blk {
    int id, n;
};
__device__ __forceinline__ Blk hw_blk() { return Blk{lin_bid(), (int)(gridDim.x * gridDim.y)}; }

// Producer side: reduce NV values over the block and store them at part[i*kMaxBlocks + bid].
template <int NV, bool MAX>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* part, int bid)
{
    __shared__ double lds[NV * 16];
    block_reduce<NV, MAX>(v, lds);
    if (lin_tid() == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) part[i * kMaxBlocks + bid] = v[i];
    }
}

template <int NV, bool MAX>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* part)
{
    block_partials<NV, MAX>(v, part, lin_bid());
}

// Consumer side: the first nv (<= NV) totals of the nparts partials of the previous kernel,
// returned to every thread of the block.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int nparts, int nv,
                                            double (&out)[NV])
{
    __shared__ double lds[NV * 16];
    __shared__ double res[NV];
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (int b = lin_tid(); b < nparts; b += lin_nthreads()) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
            if (i < nv) acc[i] += part[i * kMaxBlocks + b];
    }
    block_reduce<NV, false>(acc, lds);
    if (lin_tid() == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) res[i] = acc[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = res[i];
}

}  // namespace dh
