// fq_serve.hip -- packed execution of many QAOA evaluations of one size N.
//
// No reference counterpart (the reference runs one evaluation per process).
// A slab holds K = 2^s state slots of 2^N amplitudes in one allocation; seen
// as one virtual state of N + s qubits whose top s bits select the slot, a
// single persistent tile-pass launch covers every active slot.  Each slot has
// its own graph, angles and current layer (per-slot parameters in device
// memory), so requests of different depth P advance together one layer per
// step and a request can enter a free slot at any layer boundary.
#include "forqaoa.h"
#include "fq_device.h"
#include "fq_internal.h"

#include <hip/hip_runtime.h>

namespace fq {

// Phase tables of every active slot (layout of k_phase_table per slot, at
// slot * stride).  Grid: (entries, slots).
__global__ void __launch_bounds__(256) k_slab_phase_table(const TileArgs a, int a_log,
                                                          const SlotParams *__restrict__ slots,
                                                          double2 *__restrict__ tabs, size_t stride)
{
    const SlotParams &sp = slots[blockIdx.y];
    if (!sp.active)
        return;
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double coef;
    if (!phase_coef(a, sp.plan, a_log, gid, coef))
        return;
    const double angle = sp.gamma * coef;
    double sn, cs;
    sincos(angle, &sn, &cs);
    tabs[stride * blockIdx.y + gid] = make_double2(cs, sn);
}

// First tile at or after tl (below end) that the pass processes: its slot is
// active and, for a mirror pass, its top local bit is 0.  Skips whole slots.
template <bool MIRROR>
__device__ __forceinline__ uint64_t slab_next_live(const SlotParams *__restrict__ slots, uint64_t tl,
                                                   uint64_t end, int slot_shift)
{
    while (tl < end) {
        const uint64_t slot = tl >> slot_shift;
        if (!slots[slot].active || (MIRROR && ((tl >> (slot_shift - 1)) & 1ull))) {
            tl = (slot + 1) << slot_shift;
            continue;
        }
        return tl;
    }
    return end;
}

// The 2^A_LOG amplitudes of one thread in tile tl (mirror: the upper half
// from the partner tile at complemented offsets).
template <typename T, int A_LOG, bool MIRROR>
__device__ __forceinline__ void slab_load(T *v, const T *__restrict__ sv, const TileArgs &a, uint64_t tl,
                                          uint64_t local_mask, uint64_t toff, uint32_t rpos_w)
{
    constexpr int A = 1 << A_LOG, AH = MIRROR ? A / 2 : A;
    uint64_t roff[A];
    reg_offsets<A, MIRROR ? A_LOG - 1 : A_LOG>(rpos_w, roff);
    const T *p0 = sv + (unpack_runs_uniform(tl, a.hruns) + toff);
#pragma unroll
    for (int j = 0; j < AH; ++j)
        v[j] = p0[roff[j]];
    if (MIRROR) {
        const T *pm = sv + (unpack_runs_uniform(tl ^ local_mask, a.hruns) + a.mirror_pm - toff);
#pragma unroll
        for (int j = AH; j < A; ++j)
            v[j] = pm[-(int64_t)roff[j - AH]];
    }
}

// One tile pass over the slab.  a describes the virtual state (N + s qubits;
// the slot bits are the top high bits); tile index >> slot_shift is the slot.
// MIRROR (half-state slabs, slots of 2^(N-1) amplitudes with z_{N-1} = 0):
// the top register bit is qubit N-1; tile tau (top local bit 0) is paired with
// tau ^ local_mask of the same slot at complemented element offsets, so RX on
// that bit is the butterfly psi(b) <-> psi(~b).
template <typename R, int A_LOG, bool COST, bool MIRROR>
__global__ void __launch_bounds__(1024) k_slab_pass(typename CT<R>::T *__restrict__ sv, const TileArgs a,
                                                    const SlotParams *__restrict__ slots, int slot_shift)
{
    typedef typename CT<R>::T T;
    constexpr int A = 1 << A_LOG;
    constexpr int AH = MIRROR ? A / 2 : A;
    extern __shared__ __align__(16) unsigned char fq_smem_s[];
    const int Q = a.Q, nthr = a.nthr;
    const int t = threadIdx.x;
    const uint64_t ntiles = 1ull << a.nh;
    const uint64_t local_mask = (1ull << slot_shift) - 1ull;
    const uint64_t toff = unpack_runs((uint64_t)t, a.truns);
    // register offsets recomputed where the loads and stores use them
    const uint32_t rpos_w = a.rpos_w;
    const int qmask = (1 << Q) - 1;
    const uint64_t q1 = (uint64_t)Q + 1;
    T *lds = reinterpret_cast<T *>(fq_smem_s);
    const int nat = lds_pad(t);

    // one contiguous range of tiles per work-group; the next live tile's
    // amplitudes load while the current one is in its LDS phases
    const uint64_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint64_t range0 = (uint64_t)blockIdx.x * per;
    const uint64_t range_end = range0 + per < ntiles ? range0 + per : ntiles;
    uint64_t tile = slab_next_live<MIRROR>(slots, range0, range_end, slot_shift);
    T v[A];
    if (tile < range_end && !(COST && slots[tile >> slot_shift].cost_first))
        slab_load<T, A_LOG, MIRROR>(v, sv, a, tile, local_mask, toff, rpos_w);
    while (tile < range_end) {
        const uint64_t next = slab_next_live<MIRROR>(slots, tile + 1, range_end, slot_shift);
        T w[A];
        if (next < range_end && !(COST && slots[next >> slot_shift].cost_first))
            slab_load<T, A_LOG, MIRROR>(w, sv, a, next, local_mask, toff, rpos_w);
        const SlotParams &sp = slots[tile >> slot_shift];
        const uint64_t lt = tile & local_mask;
        T *const p0 = sv + (unpack_runs_uniform(tile, a.hruns) + toff);
        T *const pm = MIRROR ? sv + (unpack_runs_uniform(tile ^ local_mask, a.hruns) + a.mirror_pm - toff) : p0;
        const bool first = COST && sp.cost_first;
        if (COST) {
            const double2 *f_t = sp.ptab + lt * q1;
            const double2 *thr_tab = sp.ptab + (1ull << slot_shift) * q1;
            const double2 *rr_tab = thr_tab + ((size_t)(1 + A_LOG) << nthr);
            const double2 f_thr = thr_tab[t];
            double x = f_thr.x, y = f_thr.y;
            {
                const double2 f = f_t[Q];
                const double nx = x * f.x - y * f.y;
                y = x * f.y + y * f.x;
                x = nx;
            }
            for (int i = 0; i < nthr; ++i) {
                const double2 f = f_t[i];
                const double nx = x * f.x - y * f.y;
                const double ny = x * f.y + y * f.x;
                const bool on = (t >> i) & 1;
                x = on ? nx : x;
                y = on ? ny : y;
            }
            if (first) {
                x *= sp.init_amp;
                y *= sp.init_amp;
            }
            const R c0 = (R)x, s0 = (R)y;
#pragma unroll
            for (int j = 0; j < A; ++j) {
                if (first) {
                    v[j].x = c0;
                    v[j].y = s0;
                } else {
                    cmul_by(v[j], c0, s0);
                }
            }
#pragma unroll
            for (int r = 0; r < A_LOG; ++r) {
                const double2 fr = thr_tab[((1 + r) << nthr) + t];
                const double2 fq = f_t[nthr + r];
                const R fx = (R)(fr.x * fq.x - fr.y * fq.y);
                const R fy = (R)(fr.x * fq.y + fr.y * fq.x);
#pragma unroll
                for (int j = 0; j < A; ++j)
                    if ((j >> r) & 1)
                        cmul_by(v[j], fx, fy);
            }
            if (sp.has_rr) {
#pragma unroll
                for (int j = 0; j < A; ++j) {
                    const double2 f = rr_tab[j];
                    cmul_by(v[j], (R)f.x, (R)f.y);
                }
            }
        }
        const R c = (R)sp.rx_c, s = (R)sp.rx_s;
        rx_registers<T, R, A_LOG>(v, a.gate_mask >> nthr, c, s);
        if (a.ngroups > 0) {
#pragma unroll
            for (int j = 0; j < A; ++j)
                lds[nat + lds_pad(j << nthr)] = v[j];
            lds_barrier();
            for (int g = 0; g < a.ngroups; ++g) {
                const unsigned gm = group_mask(a, g);
                const unsigned rest = (unsigned)qmask & ~gm;
                int et = 0;
                {
                    int k = 0;
                    for (int i = 0; i < Q; ++i)
                        if ((rest >> i) & 1u) {
                            et |= ((t >> k) & 1) << i;
                            ++k;
                        }
                }
                const int etp = lds_pad(et);
                int pr[A_LOG > 0 ? A_LOG : 1];
                unsigned g2 = gm;
#pragma unroll
                for (int r = 0; r < A_LOG; ++r) {
                    pr[r] = __builtin_ctz(g2);
                    g2 &= g2 - 1u;
                }
                int ej[A];
#pragma unroll
                for (int j = 0; j < A; ++j) {
                    int e = 0;
#pragma unroll
                    for (int r = 0; r < A_LOG; ++r)
                        if ((j >> r) & 1)
                            e |= 1 << pr[r];
                    ej[j] = lds_pad(e);
                }
#pragma unroll
                for (int j = 0; j < A; ++j)
                    v[j] = lds[etp + ej[j]];
                rx_registers<T, R, A_LOG>(v, group_gates(a, g), c, s);
#pragma unroll
                for (int j = 0; j < A; ++j)
                    lds[etp + ej[j]] = v[j];
                lds_barrier();
            }
#pragma unroll
            for (int j = 0; j < A; ++j)
                v[j] = lds[nat + lds_pad(j << nthr)];
        }
        {
            uint64_t roff[A];
            reg_offsets<A, MIRROR ? A_LOG - 1 : A_LOG>(rpos_w, roff);
#pragma unroll
            for (int j = 0; j < AH; ++j)
                p0[roff[j]] = v[j];
#pragma unroll
            for (int j = AH; j < A; ++j)
                pm[-(int64_t)roff[j - AH]] = v[j];
        }
#pragma unroll
        for (int j = 0; j < A; ++j)
            v[j] = w[j];
        tile = next;
    }
}

// <C> and sum |psi|^2 of the slots with want_expect; per-tile partials.
template <typename R, int A_LOG>
__global__ void __launch_bounds__(256) k_slab_expect(const typename CT<R>::T *__restrict__ sv,
                                                     const TileArgs a,
                                                     const SlotParams *__restrict__ slots,
                                                     int slot_shift, double *__restrict__ partials)
{
    typedef typename CT<R>::T T;
    constexpr int A = 1 << A_LOG;
    __shared__ double s_c[FQ_MAXQ + 2];
    __shared__ double s_K;
    __shared__ double s_pc[256];
    __shared__ double s_p[256];
    const int t = threadIdx.x;
    const int nthreads = blockDim.x;
    const uint64_t tile = blockIdx.x;
    const SlotParams &sp = slots[tile >> slot_shift];
    if (!sp.want_expect) {
        if (t == 0) {
            partials[2 * tile] = 0.0;
            partials[2 * tile + 1] = 0.0;
        }
        return;
    }
    const PlanDev &p = sp.eplan;
    const uint64_t lt = tile & ((1ull << slot_shift) - 1ull);
    const uint64_t base = unpack_runs(tile, a.hruns) + unpack_runs((uint64_t)t, a.truns);
    cut_tile_prologue(p, a.Q, slot_shift, lt, s_c, &s_K, t, nthreads);
    __syncthreads();
    double cbase = s_K + p.thr[t];
    for (int i = 0; i < a.nthr; ++i)
        cbase += s_c[i] * (double)((t >> i) & 1);
    double cr[A_LOG > 0 ? A_LOG : 1];
#pragma unroll
    for (int r = 0; r < A_LOG; ++r)
        cr[r] = s_c[a.nthr + r] + p.thr[((1 + r) << a.nthr) + t];
    double acc_pc = 0.0, acc_p = 0.0;
#pragma unroll
    for (int j = 0; j < A; ++j) {
        uint64_t o = 0;
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            if ((j >> r) & 1)
                o |= 1ull << a.rpos[r];
        double cut = cbase + p.rr[j];
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            if ((j >> r) & 1)
                cut += cr[r];
        const T v = sv[base + o];
        const double re = (double)v.x, im = (double)v.y;
        const double pr = re * re + im * im;
        acc_pc += pr * cut;
        acc_p += pr;
    }
    s_pc[t] = acc_pc;
    s_p[t] = acc_p;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w && t + w < nthreads) {
            s_pc[t] += s_pc[t + w];
            s_p[t] += s_p[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        partials[2 * tile] = s_pc[0];
        partials[2 * tile + 1] = s_p[0];
    }
}

// One work-group per slot: sum that slot's tile partials in a fixed order.
// Only slots reduced in this step are written, so an earlier result stays
// until it is collected.
__global__ void __launch_bounds__(256) k_slab_finalize(const double *__restrict__ partials,
                                                       uint64_t tiles_per_slot,
                                                       const SlotParams *__restrict__ slots,
                                                       double *__restrict__ out)
{
    if (!slots[blockIdx.x].want_expect)
        return;
    __shared__ double s_a[256];
    __shared__ double s_b[256];
    const int t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * tiles_per_slot;
    double x = 0.0, y = 0.0;
    for (uint64_t i = t; i < tiles_per_slot; i += 256) {
        x += partials[2 * (base + i)];
        y += partials[2 * (base + i) + 1];
    }
    s_a[t] = x;
    s_b[t] = y;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s_a[t] += s_a[t + w];
            s_b[t] += s_b[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[2 * blockIdx.x] = s_a[0];
        out[2 * blockIdx.x + 1] = s_b[0];
    }
}

// ---- launchers -----------------------------------------------------------------

hipError_t launch_slab_phase_table(const TileArgs &local, int a_log, const SlotParams *slots, int nslots,
                                   double2 *tabs, size_t stride, hipStream_t stream)
{
    const size_t n = phase_table_entries(local, a_log);
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)nslots);
    hipLaunchKernelGGL(k_slab_phase_table, grid, dim3(256), 0, stream, local, a_log, slots, tabs, stride);
    return hipGetLastError();
}

template <typename R, int A_LOG, bool COST, bool MIRROR>
static hipError_t launch_slab_pass_t(void *sv, const TileArgs &a, const SlotParams *slots, int slot_shift,
                                     hipStream_t stream)
{
    typedef typename CT<R>::T T;
    const size_t lds = a.ngroups > 0 ? tile_lds_bytes(sizeof(R) == 8 ? 0 : 1, a.Q) : 0;
    const void *fn = reinterpret_cast<const void *>(&k_slab_pass<R, A_LOG, COST, MIRROR>);
    if (lds > 65536) {
        if (hipError_t e = allow_large_lds(fn))
            return e;
    }
    const unsigned threads = 1u << a.nthr;
    const int per_cu = occupancy_per_cu(fn, (int)threads, lds);
    const uint64_t tiles = 1ull << a.nh;
    const uint64_t resident = (uint64_t)per_cu * device_cu_count();
    const uint64_t grid = tiles < resident ? tiles : resident;
    hipLaunchKernelGGL((k_slab_pass<R, A_LOG, COST, MIRROR>), dim3((unsigned)grid), dim3(threads), lds, stream,
                       reinterpret_cast<T *>(sv), a, slots, slot_shift);
    return hipGetLastError();
}

hipError_t launch_slab_pass(int dtype, int a_log, bool cost, bool mirror, void *sv, const TileArgs &a,
                            const SlotParams *slots, int slot_shift, hipStream_t stream)
{
    if (mirror && (cost || slot_shift < 1))
        return hipErrorInvalidValue;
#define FQ_SP(R, AL) \
    return mirror ? launch_slab_pass_t<R, AL, false, true>(sv, a, slots, slot_shift, stream) \
           : cost ? launch_slab_pass_t<R, AL, true, false>(sv, a, slots, slot_shift, stream) \
                  : launch_slab_pass_t<R, AL, false, false>(sv, a, slots, slot_shift, stream);
    if (dtype == 0) {
        switch (a_log) {
            case 1: FQ_SP(double, 1)
            case 2: FQ_SP(double, 2)
            case 3: FQ_SP(double, 3)
            default: return hipErrorInvalidValue;
        }
    }
    switch (a_log) {
        case 1: FQ_SP(float, 1)
        case 2: FQ_SP(float, 2)
        case 3: FQ_SP(float, 3)
        case 4: FQ_SP(float, 4)
        default: return hipErrorInvalidValue;
    }
#undef FQ_SP
}

hipError_t launch_slab_expect(int dtype, int a_log, const void *sv, const TileArgs &a, const SlotParams *slots,
                              int slot_shift, int nslots, double *partials, double *out, hipStream_t stream)
{
    const uint64_t tiles = 1ull << a.nh;
    const dim3 g((unsigned)tiles), b(1u << a.nthr);
#define FQ_SE(R, AL)                                                                                  \
    hipLaunchKernelGGL((k_slab_expect<R, AL>), g, b, 0, stream,                                      \
                       reinterpret_cast<const typename CT<R>::T *>(sv), a, slots, slot_shift, partials); \
    break;
    if (dtype == 0) {
        switch (a_log) {
            case 1: FQ_SE(double, 1)
            case 2: FQ_SE(double, 2)
            case 3: FQ_SE(double, 3)
            case 4: FQ_SE(double, 4)
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (a_log) {
            case 1: FQ_SE(float, 1)
            case 2: FQ_SE(float, 2)
            case 3: FQ_SE(float, 3)
            case 4: FQ_SE(float, 4)
            default: return hipErrorInvalidValue;
        }
    }
#undef FQ_SE
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_slab_finalize, dim3(nslots), dim3(256), 0, stream, partials,
                       tiles / (uint64_t)nslots, slots, out);
    return hipGetLastError();
}

}  // namespace fq
