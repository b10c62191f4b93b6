// fq_device.h -- device helpers shared by the tile, adjoint and slab kernels.
// Internal to libforqaoa.
#pragma once

#include "fq_internal.h"

#include <hip/hip_runtime.h>

namespace fq {

template <typename R> struct CT;
template <> struct CT<double> { typedef double2 T; };
template <> struct CT<float> { typedef float2 T; };

__device__ __forceinline__ int lds_pad(int e) { return e + (e >> 4); }

// RX(beta) butterfly, the matrix [[c, -i s], [-i s, c]] (src/rx.cu:46-49).
template <typename T, typename R>
__device__ __forceinline__ void rx_pair(T &q0, T &q1, R c, R s)
{
    T o0, o1;
    o0.x = q0.x * c + q1.y * s;
    o0.y = q0.y * c - q1.x * s;
    o1.x = q0.y * s + q1.x * c;
    o1.y = -q0.x * s + q1.y * c;
    q0 = o0;
    q1 = o1;
}

template <typename T, typename R, int A_LOG>
__device__ __forceinline__ void rx_registers(T *v, unsigned mask, R c, R s)
{
    constexpr int A = 1 << A_LOG;
#pragma unroll
    for (int r = 0; r < A_LOG; ++r) {
        if ((mask >> r) & 1u) {
#pragma unroll
            for (int j = 0; j < A; ++j)
                if (!((j >> r) & 1))
                    rx_pair<T, R>(v[j], v[j | (1 << r)], c, s);
        }
    }
}

// RX(beta) up to a factor (rx_tform): the butterfly [[1, -it], [-it, 1]],
// one FMA per output component, half the instructions of rx_pair.
template <typename T, typename R>
__device__ __forceinline__ void rx_pair_t(T &q0, T &q1, R t)
{
    T o0, o1;
    o0.x = fma(t, q1.y, q0.x);
    o0.y = fma(-t, q1.x, q0.y);
    o1.x = fma(t, q0.y, q1.x);
    o1.y = fma(-t, q0.x, q1.y);
    q0 = o0;
    q1 = o1;
}

// RX on the register bits r set in mask (wave-uniform).  Branch-free: an
// ungated bit runs the butterfly with t = 0, which is the identity exactly.
// A branch per bit would join two register assignments after each
// butterfly, and the moves that join them cost about half as many
// instructions as the butterfly itself.
template <typename T, typename R, int A_LOG>
__device__ __forceinline__ void rx_registers_t(T *v, unsigned mask, R t)
{
    constexpr int A = 1 << A_LOG;
#pragma unroll
    for (int r = 0; r < A_LOG; ++r) {
        const R tr = ((mask >> r) & 1u) ? t : R(0);
#pragma unroll
        for (int j = 0; j < A; ++j)
            if (!((j >> r) & 1))
                rx_pair_t<T, R>(v[j], v[j | (1 << r)], tr);
    }
}

// RX on register bits 0..K-1 only (compile time): the last LDS phase of a
// pass whose gated bits are its K lowest register bits, the others filling
// the phase, runs no identity butterflies on the fillers.
template <typename T, typename R, int A_LOG, int K>
__device__ __forceinline__ void rx_registers_low(T *v, R t)
{
    constexpr int A = 1 << A_LOG;
#pragma unroll
    for (int r = 0; r < K; ++r) {
#pragma unroll
        for (int j = 0; j < A; ++j)
            if (!((j >> r) & 1))
                rx_pair_t<T, R>(v[j], v[j | (1 << r)], t);
    }
}

// Exchange of a wave's lanes 32..63 of a with lanes 0..31 of b
// (v_permlane32_swap): for a pair of registers a = v[j], b = v[j | 2^r]
// this exchanges lane bit 5 with register bit r.
__device__ __forceinline__ void lane_xchg32(float &a, float &b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

// The same for lane bit 4 (v_permlane16_swap: the odd 16-lane rows of a
// with the even rows of b).
__device__ __forceinline__ void lane_xchg16(float &a, float &b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

// v *= (c + i s)
template <typename T, typename R>
__device__ __forceinline__ void cmul_by(T &v, R c, R s)
{
    const R x = v.x, y = v.y;
    v.x = x * c - y * s;
    v.y = x * s + y * c;
}

// Value of v in lane `lane` of the wave (lane uniform).
__device__ __forceinline__ double lane_bcast(double v, int lane)
{
    const long long i = __builtin_bit_cast(long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)i, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(i >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ double2 lane_bcast(double2 v, int lane)
{
    return make_double2(lane_bcast(v.x, lane), lane_bcast(v.y, lane));
}

// v, once it has arrived: the select needs the loaded value, so the wait for
// it is issued here and not inside a later loop (where the in-order wait
// counters would also wait for that loop's stores).  Callers pass finite
// values (phase factors, cut coefficients), so the select always returns v.
__device__ __forceinline__ double2 landed(double2 v)
{
    return __builtin_isfinite(v.x) ? v : make_double2(1.0, 0.0);
}

__device__ __forceinline__ double landed(double v)
{
    return __builtin_isfinite(v) ? v : 0.0;
}

// Deposit the low bits of x into the set bits of a uniform mask, one run of
// set bits at a time (a mask with A_LOG bits cleared has at most A_LOG+1 runs).
__device__ __forceinline__ int deposit_runs(int x, unsigned mask)
{
    int e = 0;
    while (mask) {
        const int pos = __builtin_ctz(mask);
        const int len = __builtin_ctz(~(mask >> pos));
        const unsigned run = len >= 32 ? 0xffffffffu : (1u << len) - 1u;
        e |= (int)(((unsigned)x & run) << pos);
        x = len >= 32 ? 0 : (int)((unsigned)x >> len);
        mask &= ~(run << pos);
    }
    return e;
}

// Exponent over gamma of entry gid of a cost launch's phase table (the
// layout of k_phase_table): per tile the coefficients c_i(zH) (i < Q) and
// K(zH) - W/2, then the per-thread and per-register entries and the 16
// register-register entries.  Returns false past the end of the table.
__device__ __forceinline__ bool phase_coef(const TileArgs &a, const PlanDev &p, int a_log, uint64_t gid,
                                           double &coef)
{
    const uint64_t q1 = (uint64_t)a.Q + 1;
    const uint64_t ntile_e = run_tiles(a) * q1;
    const uint64_t nthr_e = (uint64_t)(1 + a_log) << a.nthr;
    if (gid < ntile_e) {
        const uint64_t tile = gid / q1;
        const int i = (int)(gid - tile * q1);
        if (i < a.Q) {
            double acc = 0.0;
            for (int e = p.csr_off[i]; e < p.csr_off[i + 1]; ++e)
                acc += p.csr_w[e] * (double)((tile >> p.csr_h[e]) & 1ull);
            coef = p.cH[i] - 2.0 * acc;
        } else {
            double K = 0.0;
            for (int e = 0; e < p.n_hh; ++e)
                K += p.hh_w[e] * (double)(((tile >> p.hh_u[e]) ^ (tile >> p.hh_v[e])) & 1ull);
            for (int h = 0; h < a.nh; ++h)
                K += p.omega[h] * (double)((tile >> h) & 1ull);
            coef = K - 0.5 * p.wtot;
        }
        return true;
    }
    if (gid < ntile_e + nthr_e) {
        coef = p.thr[gid - ntile_e];
        return true;
    }
    if (gid < ntile_e + nthr_e + FQ_MAX_REGS) {
        coef = p.rr[gid - ntile_e - nthr_e];
        return true;
    }
    return false;
}

// Per-tile part of the cut decomposition, computed by at most two waves:
// c_i(zH) for every tile bit (lanes i < Q) and K(zH) over the tile index's nh
// hidden bits (one wave, butterfly reduction, lane 0 writes).  Published by
// the caller's barrier.
__device__ __forceinline__ void cut_tile_prologue(const PlanDev &p, int Q, int nh, uint64_t tile, double *s_c,
                                                  double *s_K, int t, int nthreads)
{
    for (int i = t; i < Q; i += nthreads) {
        double acc = 0.0;
        for (int e = p.csr_off[i]; e < p.csr_off[i + 1]; ++e)
            acc += p.csr_w[e] * (double)((tile >> p.csr_h[e]) & 1ull);
        s_c[i] = p.cH[i] - 2.0 * acc;
    }
    if (nthreads >= 64) {
        const int kw = nthreads >= 128 ? 64 : 0;
        if (t >= kw && t < kw + 64) {
            const int lane = t - kw;
            double acc = 0.0;
            for (int e = lane; e < p.n_hh; e += 64)
                acc += p.hh_w[e] * (double)(((tile >> p.hh_u[e]) ^ (tile >> p.hh_v[e])) & 1ull);
            for (int h = lane; h < nh; h += 64)
                acc += p.omega[h] * (double)((tile >> h) & 1ull);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1)
                acc += __shfl_xor(acc, off, 64);
            if (lane == 0)
                *s_K = acc;
        }
    } else if (t == 0) {
        double acc = 0.0;
        for (int e = 0; e < p.n_hh; ++e)
            acc += p.hh_w[e] * (double)(((tile >> p.hh_u[e]) ^ (tile >> p.hh_v[e])) & 1ull);
        for (int h = 0; h < nh; ++h)
            acc += p.omega[h] * (double)((tile >> h) & 1ull);
        *s_K = acc;
    }
}

// Barrier that orders LDS accesses only: outstanding global loads (the next
// tile's prefetch) stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Scatter the bits of a packed index to physical positions, as runs:
// index bits [src, src+len) go to positions [dst, dst+len).  Unrolled with a
// fixed bound so the run descriptors stay in scalar registers.
__device__ __forceinline__ uint64_t unpack_runs(uint64_t x, const BitRuns &r)
{
    uint64_t o = 0;
#pragma unroll
    for (int k = 0; k < FQ_MAX_RUNS; ++k) {
        if (k < r.n) {
            const uint32_t d = r.run[k];
            o |= ((x >> (d & 0xff)) & ((1ull << ((d >> 8) & 0xff)) - 1ull)) << (d >> 16);
        }
    }
    return o;
}

// Passes a wave-uniform word through an empty asm in a scalar register: what
// is computed from it after this point is computed there, in scalar
// instructions, instead of once before a tile loop.  Loop-invariant scalars
// held across a tile loop (the 2^A_LOG 64-bit register offsets, the bit-run
// shifts and masks of the tile index, an LDS phase's slots) do not fit the
// scalar register file; the compiler spills them to VGPR lanes, and each use
// then costs one v_readlane (a vector instruction) per 32 bits.
__device__ __forceinline__ void scalar_opaque(uint32_t &x)
{
    asm volatile("" : "+s"(x));
}

// unpack_runs of a wave-uniform index (a tile index) with the run words
// recomputed in scalar registers at each call (scalar_opaque).
__device__ __forceinline__ uint64_t unpack_runs_uniform(uint64_t x, const BitRuns &r)
{
    uint32_t n = (uint32_t)r.n;
    scalar_opaque(n);
    uint64_t o = 0;
#pragma unroll
    for (int k = 0; k < FQ_MAX_RUNS; ++k) {
        if ((uint32_t)k < n) {
            uint32_t d = r.run[k];
            scalar_opaque(d);
            o |= ((x >> (d & 0xff)) & ((1ull << ((d >> 8) & 0xff)) - 1ull)) << (d >> 16);
        }
    }
    return o;
}

// Element offsets of a thread's 2^A_LOG registers: o[j] = sum of 2^pos[r]
// over the set bits r < NB of j, from the packed positions pos_w (pos[r] in
// bits 6r..6r+5), recomputed in scalar registers at each call (scalar_opaque).
template <int A, int NB>
__device__ __forceinline__ void reg_offsets(uint32_t pos_w, uint64_t (&o)[A])
{
    scalar_opaque(pos_w);
#pragma unroll
    for (int j = 0; j < A; ++j) {
        uint64_t x = 0;
#pragma unroll
        for (int r = 0; r < NB; ++r)
            if ((j >> r) & 1)
                x |= 1ull << ((pos_w >> (6 * r)) & 63u);
        o[j] = x;
    }
}

// Tile-bit mask and register gate bits of LDS phase g.  Read from packed
// words with a shift: an indexed kernel-argument array would be loaded from
// memory inside the group loop, and the wait for that load (the counters are
// in issue order) would also wait for the next tile's prefetch and the
// previous tile's stores.
__device__ __forceinline__ unsigned group_mask(const TileArgs &a, int g)
{
    static_assert(FQ_MAX_GROUPS == 5, "four packed masks plus one");
    return g < 4 ? (unsigned)((a.grp_mask_w >> (16 * g)) & 0xffffu) : (unsigned)a.grp_mask[4];
}

__device__ __forceinline__ unsigned group_gates(const TileArgs &a, int g)
{
    return (a.grp_gates_w >> (5 * g)) & 0x1fu;
}

}  // namespace fq
