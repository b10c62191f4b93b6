// fq_kernels.hip -- gfx950 kernels of the QAOA Max-Cut layer loop.
//
// Everything on this path is HBM streaming: each kernel reads and writes every
// amplitude once per launch (the block transposition moves (1 - 2^-k) of
// them), so the kernels are shaped for coalesced 64-lane access, LDS staging
// and few VALU operations per byte.  No MFMA: there is no dense contraction.
//
//  k_tile    one "tile pass": a work-group owns 2^Q amplitudes whose index
//            bits at the Q tile positions vary (the high bits are fixed by
//            the tile index).  Optional cost phase (Max-Cut value per basis
//            index from a per-tile / per-thread decomposition of the
//            quadratic form, then exp(i*gamma*s/2)), optional launch-control
//            write (no read), optional RX on any subset of the tile bits.
//            Register bits are butterflied in registers; the other gate bits
//            are remapped into registers through padded LDS phases.
//            Replaces _rotationCompressionWeighted_conti (src/rzz.cu:121-156)
//            and _multirotateX_unroll (src/rx.cu:56-87).
//  k_swap_tiled / k_swap_generic
//            in-place qubit-block exchange [out,out+k) <-> [in,in+k);
//            replaces _swap_gate_conti (src/swap.cu:76-102) without the
//            host-built pair table: 2^t x 2^t sub-tiles are transposed
//            through LDS so both sides are read and written coalesced.
//  k_expect / k_finalize   <C> and sum|psi|^2, FP64, fixed reduction order.
//  k_topk    per-block top-16 candidates (bitonic sort in LDS).
#include "fq_device.h"
#include "fq_internal.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <tuple>

namespace fq {

// Phase table of one cost-phase launch (the tile kernel does no transcendental
// work).  Layout: for each tile, Q+1 factors
//   tab[tile*(Q+1) + i] = exp(i*gamma*c_i(zH))          i < Q
//   tab[tile*(Q+1) + Q] = exp(i*gamma*(K(zH) - W/2))
// then, per thread t of the tile kernel and register bit r,
//   tab[ntiles*(Q+1) + t]                  = exp(i*gamma*thr[t])
//   tab[ntiles*(Q+1) + ((1+r) << nthr) + t] = exp(i*gamma*d_r(t))
// and finally the 16 register-register factors exp(i*gamma*rr[j]),
// so that exp(i*gamma*s/2) = exp(i*gamma*(Cut - W/2)) of every amplitude is a
// product of table entries.  If cut is not null it receives the exponents
// divided by gamma at the same indices (c_i, K - W/2, thr, d_r, rr): their
// sums give s/2 = Cut - W/2 of every amplitude.
__global__ void __launch_bounds__(256) k_phase_table(const TileArgs a, int a_log,
                                                     double2 *__restrict__ tab,
                                                     double *__restrict__ cut)
{
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double coef;
    if (!phase_coef(a, a.plan, a_log, gid, coef))
        return;
    double sn, cs;
    sincos(a.gamma * coef, &sn, &cs);
    tab[gid] = make_double2(cs, sn);
    if (cut)
        cut[gid] = coef;
}

__device__ __forceinline__ void zmul(double &x, double &y, double2 f)
{
    const double nx = x * f.x - y * f.y;
    y = x * f.y + y * f.x;
    x = nx;
}

// exp(i*theta) with theta = s*gamma/2 as src/rzz.cu:146 (FP64 sincos).
__device__ __forceinline__ void phase_of(double theta, double &c, double &s)
{
    sincos(theta, &s, &c);
}
// complex64: range-reduce theta in FP64, evaluate in FP32.
__device__ __forceinline__ void phase_of(double theta, float &c, float &s)
{
    const double two_pi_hi = 6.283185307179586232;
    const double two_pi_lo = 2.4492935982947064e-16;
    const double n = rint(theta * 0.15915494309189535);
    double r = fma(-n, two_pi_hi, theta);
    r = fma(-n, two_pi_lo, r);
    sincosf((float)r, &s, &c);
}

__device__ __forceinline__ uint64_t tile_base(const TileArgs &a, uint64_t tile)
{
    return unpack_runs_uniform(tile, a.hruns);
}

// deposit the low bits of x into the set bits of mask (q-bit masks)
template <int QT>
__device__ __forceinline__ int deposit(int x, unsigned mask, int q)
{
    int e = 0, k = 0;
    const int qq = QT > 0 ? QT : q;
#pragma unroll
    for (int i = 0; i < (QT > 0 ? QT : FQ_MAXQ); ++i) {
        if (i < qq && ((mask >> i) & 1u)) {
            e |= ((x >> k) & 1) << i;
            ++k;
        }
    }
    return e;
}

// k-th tile of this work-group in a persistent pass of G = gridDim.x groups:
//   order 0: k*G + b (the groups sweep the tiles together),
//   order 1: the groups of one XCD (dispatched round-robin, b mod 8) take
//            consecutive tiles of each sweep,
//   order 2: one contiguous range of tiles per group,
//   order 3: as 0 with one work-group per tile (no persistence).
// Returns ntiles or more once the group has no tile left.
__device__ __forceinline__ uint64_t tile_at(const TileArgs &a, uint64_t ntiles, uint64_t k)
{
    const uint64_t G = gridDim.x, b = blockIdx.x;
    if (a.order == 1)
        return k * G + (b & 7) * (G >> 3) + (b >> 3);
    if (a.order == 2) {
        const uint64_t per = (ntiles + G - 1) / G;
        return k < per ? b * per + k : ntiles;
    }
    return k * G + b;
}

// Hides the value of an LDS base index from the compiler between a phase's
// reads and its write-back, so the write-back recomputes its 2^A_LOG
// addresses (one add each) instead of keeping them live in VGPRs through the
// phase's butterflies.  Used where the VGPRs run short: the complex64 passes
// with the cost phase (with the addresses kept, the boundary pass spilled 36
// bytes per lane, three of them prefetched amplitudes whose spill waited for
// every load in flight).
__device__ __forceinline__ void lds_opaque(int &x)
{
    asm volatile("" : "+v"(x));
}

// RX on the gate positions of a tile held in the natural layout (register j
// of thread t = tile element t + j * 2^nthr), returned in the same layout:
// register bits in registers, thread bits through the LDS remapping phases.
// The caller's previous LDS access was a read of its own natural slots (or
// none), so the first write here needs no barrier.
template <typename T, typename R, int A_LOG>
__device__ __forceinline__ void mix_natural(T (&v)[1 << A_LOG], const TileArgs &a, const int *etp, int nat,
                                            int nthr, R t)
{
    constexpr int A = 1 << A_LOG;
    extern __shared__ __align__(16) unsigned char fq_smem[];
    rx_registers_t<T, R, A_LOG>(v, a.gate_mask >> nthr, t);
    if (a.ngroups == 0)
        return;
    T *lds = reinterpret_cast<T *>(fq_smem);
#pragma unroll
    for (int j = 0; j < A; ++j)
        lds[nat + lds_pad(j << nthr)] = v[j];
    lds_barrier();
#pragma unroll 1
    for (int g = 0; g < a.ngroups; ++g) {
        int eg = g == 0 ? etp[0] : g == 1 ? etp[1] : g == 2 ? etp[2] : g == 3 ? etp[3] : etp[4];
        const unsigned gm = group_mask(a, g);
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
            v[j] = lds[eg + ej[j]];
        lds_opaque(eg);
        rx_registers_t<T, R, A_LOG>(v, group_gates(a, g), t);
#pragma unroll
        for (int j = 0; j < A; ++j)
            lds[eg + ej[j]] = v[j];
        lds_barrier();
    }
#pragma unroll
    for (int j = 0; j < A; ++j)
        v[j] = lds[nat + lds_pad(j << nthr)];
}

// Persistent tile pass.  Each work-group walks tiles blockIdx.x,
// blockIdx.x + gridDim.x, ...; while a tile goes through its cost phase and
// LDS phases, the next tile's amplitudes are already loading into registers.
// QT > 0 fixes the tile width at compile time (the fused passes' full-width
// tiles); QT == 0 reads it from the arguments.
// PRE: the pass that ends one layer and opens the next applies RX of the
// previous layer (pre_t) on the same gate positions before the cost phase
// (COST, not FIRST, MIRROR 0 only).
// RX is applied as the one-FMA butterfly of rx_tform, i.e. up to a factor per
// gate; the host passes the product of those factors (and of any earlier
// passes that left theirs pending) as a.scale_re + i a.scale_im, which a cost
// pass folds into its FP64 phase factor and a pass without the cost phase
// multiplies in before its stores when a.apply_scale is set.
// KL > 0: the last LDS phase's gated bits are its KL lowest register bits
// (the host launches this instance only then), and its RX runs on those
// alone instead of running identity butterflies on the filler bits.
// LS (the complex64 layer-boundary pass at 1024 x 16, a.lane_swap): thread
// bits 4 and 5, both gated, are exchanged with register bits 1 and 0 by
// v_permlane16_swap / v_permlane32_swap right after the natural layout's RX,
// get their RX there, and go to the LDS in that exchanged layout; the LDS
// phases take the other gated thread bits, and after the second RX set the
// swaps bring the natural layout back for the stores.  At N=34 that turns
// the strided boundary pass's two LDS phases (4 + 2 bits) into one (strided
// boundary pass 65.6 -> 58.3 ms).  The plain passes do not use it: with
// their last phase's registers on other qubits their direct stores measured
// far slower (N=34 plain strided pass 51.2 -> 75.6 ms).
template <typename R, int A_LOG, int QT, bool COST, bool FIRST, bool GATES, int MIRROR = 0, bool PRE = false,
          int KL = 0, bool LS = false>
__global__ void __launch_bounds__(GATES ? (A_LOG >= 5 || (sizeof(R) == 8 && A_LOG >= 4) ? 512 : 1024) : 256)
    k_tile(typename CT<R>::T *__restrict__ sv, const TileArgs a, const double2 *__restrict__ ptab)
{
    typedef typename CT<R>::T T;
    constexpr int A = 1 << A_LOG;
    constexpr int MAXG = FQ_MAX_GROUPS;
    extern __shared__ __align__(16) unsigned char fq_smem[];

    const int Q = QT > 0 ? QT : a.Q;
    const int nthr = QT > 0 ? QT - A_LOG : a.nthr;
    const int t = threadIdx.x;
    // MIRROR: iterate tile pairs (tau, ~tau); the upper half of the register
    // index addresses the complementary tile at complemented offsets.
    // MIRROR == 2 also writes the upper half of the full state
    // (psi(x | 2^(N-1)) = psi(~x)) from the same registers.
    const uint64_t ntiles = run_tiles(a);
    const uint64_t hmask = (1ull << a.nh) - 1ull;
    constexpr int AH = MIRROR ? (1 << A_LOG) / 2 : (1 << A_LOG);

    const uint64_t toff = unpack_runs((uint64_t)t, a.truns);
    // element offset of register j, recomputed where the loads and stores use it
    const uint32_t rpos_w = a.rpos_w;
    constexpr int NB = MIRROR ? A_LOG - 1 : A_LOG;
    // per-thread LDS slots of every remapping phase (padded), once per launch;
    // lds_pad is additive over disjoint bit sets: pad(x | y) = pad(x) + pad(y)
    int etp[MAXG];
    const int nat = LS ? lds_pad((t & ~0x30) | (((t >> 5) & 1) << (QT - A_LOG)) | (((t >> 4) & 1) << (QT - A_LOG + 1)))
                       : lds_pad(t);
    if (GATES) {
        const unsigned qmask = (1u << Q) - 1u;
#pragma unroll
        for (int g = 0; g < MAXG; ++g)
            etp[g] = g < a.ngroups ? lds_pad(deposit<QT>(t, qmask & ~(unsigned)a.grp_mask[g], Q)) : 0;
    }
    // Direct store: when the last LDS phase holds none of the six lane bits,
    // the tile is stored straight from that phase's registers (no write-back
    // and re-read of the natural layout).  Element e of the tile lives at
    // tile_base + phys(e); phys is additive over disjoint bits: the thread's
    // part (pe_last) is computed once per launch, the part of each register
    // (the phase's register bits sit at the positions packed in a.dpos_w) at
    // the stores, as reg_offsets does for the natural layout.
    // (full-width tiles only, and not in the read-write cost passes, whose
    // registers are full with the prefetch and the phase factors)
    const bool direct = GATES && MIRROR == 0 && QT > 0 && (FIRST || !COST) && a.direct_store && a.ngroups > 0;
    uint64_t pe_last = 0;
    if (direct) {
        const unsigned qmask = (1u << Q) - 1u;
        const unsigned gm = (unsigned)a.grp_mask[a.ngroups - 1];
        const unsigned tmask = (1u << nthr) - 1u;
        const unsigned et = (unsigned)deposit<QT>(t, qmask & ~gm, Q);
        pe_last = unpack_runs((uint64_t)(et & tmask), a.truns);
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            if ((et >> (nthr + r)) & 1u)
                pe_last |= 1ull << a.rpos[r];
    }
    const uint64_t q1 = (uint64_t)Q + 1;
    const double2 *thr_tab = COST ? ptab + (ntiles * q1) : nullptr;
    const double2 *rr_tab = COST ? thr_tab + ((size_t)(1 + A_LOG) << nthr) : nullptr;

    uint64_t kt = 0;
    uint64_t tile = tile_at(a, ntiles, 0);
    // compile-time tile width: lane i of each wave holds factor i of the
    // tile's phase-table row, loaded one tile ahead with the amplitudes
    // (not in the complex64 read-write cost passes: the boundary pass
    // with the row and the register-register factors in VGPRs spilled 92
    // bytes per lane, and both passes fit their 128 VGPRs without spills
    // when they read the row with scalar loads instead)
    constexpr bool ROWPF = COST && QT > 0 && !(sizeof(R) == 4 && !FIRST);
    // When the next tile's loads are issued (measured on MI355X; a tile's
    // loads and the previous tile's stores issued back to back left the
    // memory pipe idle through the compute between them):
    // * the complex64 boundary pass at 1024 threads issues half of them
    //   before its cost phase and half after it (N=34: strided boundary
    //   pass 72.5 -> 66.2 ms, 8.31 -> 8.56 layers/s);
    // * the passes without the cost phase issue them after their first LDS
    //   phase instead of at the top of the iteration (N=34 c64 8.58 -> 8.82
    //   layers/s, N=30 c64 130.2 -> 131.7, half state 247 -> 250; complex128
    //   unchanged).
    constexpr bool SPLITPF = PRE && sizeof(R) == 4 && A_LOG == 4;
    constexpr bool LATEPF = !COST;
    const int row_lane = (t & 63) % (QT + 1);
    double2 fcur = make_double2(0.0, 0.0), fnext = make_double2(0.0, 0.0);
    if (ROWPF && tile < ntiles)
        fcur = ptab[tile * q1 + row_lane];
    // the thread's own factors do not depend on the tile: loaded once, so no
    // load issued after a tile's stores has to wait for them
    double2 f_thr = make_double2(1.0, 0.0);
    double2 f_reg[A_LOG > 0 ? A_LOG : 1];
    // compile-time widths have at least 64 threads: lane j holds
    // register-register factor j (broadcast with readlane)
    double2 f_rr = make_double2(1.0, 0.0);
    if (COST) {
        f_thr = landed(thr_tab[t]);
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            f_reg[r] = landed(thr_tab[((1 + r) << nthr) + t]);
        if (ROWPF)
            f_rr = landed(rr_tab[(t & 63) % A]);
    }
    T v[A];
    if (!FIRST && tile < ntiles) {
        uint64_t roff[A];
        reg_offsets<A, NB>(rpos_w, roff);
        const T *p0 = sv + (tile_base(a, tile) + toff);
#pragma unroll
        for (int j = 0; j < AH; ++j)
            v[j] = p0[roff[j]];
        if (MIRROR) {
            const T *pm = sv + (tile_base(a, ~tile & hmask) + a.mirror_pm - toff);
#pragma unroll
            for (int j = AH; j < A; ++j)
                v[j] = pm[-(int64_t)roff[j - AH]];
        }
    }
    while (tile < ntiles) {
        const uint64_t next = tile_at(a, ntiles, ++kt);

        // Cost phase: exp(i*gamma*s/2), s = 2*Cut - W, of register j is the
        // product of the thread's base factor, one factor per register bit
        // set in j and the register-register term; all factors come from the
        // phase table, whose thread and register bits are those of the layout
        // the registers hold when it is applied (the natural layout, or in the
        // merged boundary pass the last LDS phase's).
        auto apply_cost = [&]() {
            const double2 *f_t = ptab + tile * q1;
            auto row = [&](int i) -> double2 { return ROWPF ? lane_bcast(fcur, i) : f_t[i]; };
            double x = f_thr.x, y = f_thr.y;
            zmul(x, y, row(Q));
            // complex64: thread bits from an opaque copy of t, so the
            // per-bit lane masks of the selects below are recomputed per tile
            // (two vector instructions each) instead of held in scalar
            // registers across the tile loop, where they spill (the
            // complex128 passes hold them without spilling)
            int tb = t;
            if (sizeof(R) == 4)
                lds_opaque(tb);
#pragma unroll
            for (int i = 0; i < (QT > 0 ? QT - A_LOG : FQ_MAXQ); ++i) {
                if (QT > 0 || i < nthr) {
                    const double2 f = row(i);
                    const double nx = x * f.x - y * f.y;
                    const double ny = x * f.y + y * f.x;
                    const bool on = (tb >> i) & 1;
                    x = on ? nx : x;
                    y = on ? ny : y;
                }
            }
            // the pass's RX factor (and the launch-control amplitude)
            {
                const double sr = FIRST ? a.init_amp * a.scale_re : a.scale_re;
                const double si = FIRST ? a.init_amp * a.scale_im : a.scale_im;
                const double nx = x * sr - y * si;
                y = x * si + y * sr;
                x = nx;
            }
            const R c0 = (R)x, s0 = (R)y;
#pragma unroll
            for (int j = 0; j < A; ++j) {
                if (FIRST) {
                    v[j].x = c0;
                    v[j].y = s0;
                } else {
                    cmul_by(v[j], c0, s0);
                }
            }
#pragma unroll
            for (int r = 0; r < A_LOG; ++r) {
                double fx = f_reg[r].x, fy = f_reg[r].y;
                zmul(fx, fy, row(nthr + r));
                const R cf = (R)fx, sf = (R)fy;
#pragma unroll
                for (int j = 0; j < A; ++j)
                    if ((j >> r) & 1)
                        cmul_by(v[j], cf, sf);
            }
            if (a.has_rr) {
#pragma unroll
                for (int j = 0; j < A; ++j) {
                    const double2 f = ROWPF ? lane_bcast(f_rr, j) : rr_tab[j];
                    cmul_by(v[j], (R)f.x, (R)f.y);
                }
            }
        };
        T w[A];
        // the next tile's loads (issued after the cost phase: few registers live)
        // registers [j0, j1) of the next tile (part of prefetch)
        auto prefetch_part = [&](int j0, int j1) {
            if (ROWPF && j0 == 0 && next < ntiles)
                fnext = ptab[next * q1 + row_lane];
            if (!FIRST && next < ntiles) {
                uint64_t roff[A];
                reg_offsets<A, NB>(rpos_w, roff);
                const T *p1 = sv + (tile_base(a, next) + toff);
#pragma unroll
                for (int j = 0; j < AH; ++j)
                    if (j >= j0 && j < j1)
                        w[j] = p1[roff[j]];
                if (MIRROR && j1 > AH) {
                    const T *pm = sv + (tile_base(a, ~next & hmask) + a.mirror_pm - toff);
#pragma unroll
                    for (int j = AH; j < A; ++j)
                        if (j >= j0 && j < j1)
                            w[j] = pm[-(int64_t)roff[j - AH]];
                }
            }
        };
        auto prefetch = [&]() {
            if (ROWPF && next < ntiles)
                fnext = ptab[next * q1 + row_lane];
            if (!FIRST && next < ntiles) {
                uint64_t roff[A];
                reg_offsets<A, NB>(rpos_w, roff);
                const T *p1 = sv + (tile_base(a, next) + toff);
#pragma unroll
                for (int j = 0; j < AH; ++j)
                    w[j] = p1[roff[j]];
                if (MIRROR) {
                    const T *pm = sv + (tile_base(a, ~next & hmask) + a.mirror_pm - toff);
#pragma unroll
                    for (int j = AH; j < A; ++j)
                        w[j] = pm[-(int64_t)roff[j - AH]];
                }
            }
        };
        T *lds = reinterpret_cast<T *>(fq_smem);
        // LDS slots of LDS phase g (padded): eg for this thread, ej[j] for register j
        int ej[A];
        auto phase_slots = [&](int g) -> int {
            uint32_t gm = group_mask(a, g);
            scalar_opaque(gm);
            int pr[A_LOG > 0 ? A_LOG : 1];
            unsigned g2 = gm;
#pragma unroll
            for (int r = 0; r < A_LOG; ++r) {
                pr[r] = __builtin_ctz(g2);
                g2 &= g2 - 1u;
            }
#pragma unroll
            for (int j = 0; j < A; ++j) {
                int e = 0;
#pragma unroll
                for (int r = 0; r < A_LOG; ++r)
                    if ((j >> r) & 1)
                        e |= 1 << pr[r];
                ej[j] = lds_pad(e);
            }
            return g == 0 ? etp[0] : g == 1 ? etp[1] : g == 2 ? etp[2] : g == 3 ? etp[3] : etp[4];
        };
        // read LDS phase g into the registers and apply RX(tt) on its gated bits
        auto phase_read_rx = [&](int g, int &eg, R tt) {
            eg = phase_slots(g);
#pragma unroll
            for (int j = 0; j < A; ++j)
                v[j] = lds[eg + ej[j]];
            if (COST && sizeof(R) == 4)
                lds_opaque(eg);
            rx_registers_t<T, R, A_LOG>(v, group_gates(a, g), tt);
        };
        // the last LDS phase of a KL instance
        auto phase_read_last = [&](int g, int &eg, R tt) {
            eg = phase_slots(g);
#pragma unroll
            for (int j = 0; j < A; ++j)
                v[j] = lds[eg + ej[j]];
            if (COST && sizeof(R) == 4)
                lds_opaque(eg);
            rx_registers_low<T, R, A_LOG, (KL > 0 ? KL : A_LOG)>(v, tt);
        };
        auto phase_write = [&](int eg) {
#pragma unroll
            for (int j = 0; j < A; ++j)
                lds[eg + ej[j]] = v[j];
            lds_barrier();
        };
        // LS: element (t, j) of the exchanged layout is tile element
        // (t without bits 4, 5) | t4 << (nthr + 1) | t5 << nthr | j0 << 5 |
        // j1 << 4 | (j >> 2) << (nthr + 2); lds_pad is additive over disjoint
        // bits, so the thread's part (nat) and register j's (a constant) add
        auto reg_slot = [&](int j) -> int {
            if constexpr (LS)
                return lds_pad(((j & 1) << 5) | (((j >> 1) & 1) << 4) | ((j >> 2) << (QT - A_LOG + 2)));
            else
                return lds_pad(j << nthr);
        };
        auto natural_write = [&]() {
            // own slots: the previous tile's final read used the same slots,
            // so no barrier is needed before this write
#pragma unroll
            for (int j = 0; j < A; ++j)
                lds[nat + reg_slot(j)] = v[j];
            lds_barrier();
        };
        auto natural_read = [&]() {
#pragma unroll
            for (int j = 0; j < A; ++j)
                v[j] = lds[nat + reg_slot(j)];
        };
        // LS: exchange thread bit 5 with register bit 0 and thread bit 4 with
        // register bit 1 (each swap is its own inverse, and the two commute)
        auto lane_swap = [&]() {
            if constexpr (LS) {
#pragma unroll
                for (int j = 0; j < A; ++j)
                    if (!(j & 1)) {
                        lane_xchg32(v[j].x, v[j | 1].x);
                        lane_xchg32(v[j].y, v[j | 1].y);
                    }
#pragma unroll
                for (int j = 0; j < A; ++j)
                    if (!(j & 2)) {
                        lane_xchg16(v[j].x, v[j | 2].x);
                        lane_xchg16(v[j].y, v[j | 2].y);
                    }
            }
        };

        if constexpr (PRE && sizeof(R) == 4) {
            // complex64 layer-boundary pass: RX(pre) through its LDS phases, the cost
            // phase applied in the last phase's layout (with the phase table
            // of that layout; the natural one without LDS phases), RX of this
            // layer on that phase's bits while the registers still hold it,
            // then the other phases and the natural register bits: two LDS
            // round trips fewer than finishing RX(pre) in the natural layout
            // and starting again.
            const R pt = a.pre_tf, rt = a.rx_tf;
            const int G = a.ngroups;
            rx_registers_t<T, R, A_LOG>(v, a.gate_mask >> nthr, pt);
            if constexpr (LS) {
                lane_swap();
                rx_registers_low<T, R, A_LOG, 2>(v, pt);
            }
            int eg = 0;
            if (G > 0) {
                natural_write();
                if constexpr (KL > 0) {
#pragma unroll 1
                    for (int g = 0; g + 1 < G; ++g) {
                        phase_read_rx(g, eg, pt);
                        phase_write(eg);
                    }
                    phase_read_last(G - 1, eg, pt);
                } else {
#pragma unroll 1
                    for (int g = 0; g < G; ++g) {
                        phase_read_rx(g, eg, pt);
                        if (g + 1 == G)
                            break;
                        phase_write(eg);
                    }
                }
            }
            if constexpr (SPLITPF)
                prefetch_part(0, A / 2);
            apply_cost();
            if constexpr (SPLITPF)
                prefetch_part(A / 2, A);
            else
                prefetch();
            if (G > 0) {
                if constexpr (KL > 0)
                    rx_registers_low<T, R, A_LOG, KL>(v, rt);
                else
                    rx_registers_t<T, R, A_LOG>(v, group_gates(a, G - 1), rt);
                eg = phase_slots(G - 1);
                phase_write(eg);
#pragma unroll 1
                for (int g = 0; g + 1 < G; ++g) {
                    phase_read_rx(g, eg, rt);
                    phase_write(eg);
                }
                natural_read();
            }
            if constexpr (LS) {
                rx_registers_low<T, R, A_LOG, 2>(v, rt);
                lane_swap();
            }
            rx_registers_t<T, R, A_LOG>(v, a.gate_mask >> nthr, rt);
        } else {
            // (the complex128 boundary pass, at 128 VGPRs with its row
            // prefetch, spills in the merged order: it finishes RX(pre) in
            // the natural layout instead)
            if (PRE)
                mix_natural<T, R, A_LOG>(v, a, etp, nat, nthr, (R)a.pre_t);
            if (COST)
                apply_cost();
            if (!LATEPF)
                prefetch();
            // LATEPF: after the first LDS phase's write-back, or after the
            // RX where that phase is stored directly or there is none
            bool pf_left = LATEPF;
            auto prefetch_rest = [&]() {
                if (LATEPF && pf_left) {
                    prefetch();
                    pf_left = false;
                }
            };
            if (GATES) {
                const R rt = sizeof(R) == 4 ? (R)a.rx_tf : (R)a.rx_t;
                rx_registers_t<T, R, A_LOG>(v, a.gate_mask >> nthr, rt);
                if (a.ngroups > 0) {
                    natural_write();
                    // rolled loop: unrolling lets the compiler hoist every
                    // phase's addresses out of the tile loop and spill
                    int eg = 0;
                    if constexpr (KL > 0) {
                        const int G = a.ngroups;
#pragma unroll 1
                        for (int g = 0; g + 1 < G; ++g) {
                            phase_read_rx(g, eg, rt);
                            phase_write(eg);
                            if (g == 0)
                                prefetch_rest();
                        }
                        // (KL instances run with two or more phases: the
                        // loads went out after the first)
                        phase_read_last(G - 1, eg, rt);
                        if (!direct)
                            phase_write(eg);
                    } else {
#pragma unroll 1
                        for (int g = 0; g < a.ngroups; ++g) {
                            phase_read_rx(g, eg, rt);
                            if (direct && g == a.ngroups - 1)
                                break;
                            phase_write(eg);
                            if (g == 0)
                                prefetch_rest();
                        }
                    }
                    if (direct) {
                        // every thread has read its last-phase elements before
                        // the next tile's first LDS writes
                        lds_barrier();
                    } else {
                        natural_read();
                    }
                }
            }
            prefetch_rest();
        }

        if (!COST && a.apply_scale) {
#pragma unroll
            for (int j = 0; j < A; ++j)
                cmul_by(v[j], (R)a.scale_re, (R)a.scale_im);
        }

        if (direct) {
            uint64_t doff[A];
            reg_offsets<A, A_LOG>(a.dpos_w, doff);
            T *const ps = sv + (tile_base(a, tile) + pe_last);
#pragma unroll
            for (int j = 0; j < A; ++j)
                ps[doff[j]] = v[j];
        } else {
            uint64_t roff[A];
            reg_offsets<A, NB>(rpos_w, roff);
            T *const p0 = sv + (tile_base(a, tile) + toff);
#pragma unroll
            for (int j = 0; j < AH; ++j)
                p0[roff[j]] = v[j];
            if (MIRROR) {
                T *const pm = sv + (tile_base(a, ~tile & hmask) + a.mirror_pm - toff);
#pragma unroll
                for (int j = AH; j < A; ++j)
                    pm[-(int64_t)roff[j - AH]] = v[j];
                if (MIRROR == 2) {
                    const uint64_t half = 1ull << (a.N - 1);
#pragma unroll
                    for (int j = 0; j < AH; ++j)
                        p0[half + roff[j]] = v[j + AH];
#pragma unroll
                    for (int j = AH; j < A; ++j)
                        pm[half - roff[j - AH]] = v[j - AH];
                }
            }
        }
        if (!FIRST && next < ntiles) {
#pragma unroll
            for (int j = 0; j < A; ++j)
                v[j] = w[j];
        }
        if (ROWPF)
            fcur = fnext;
        tile = next;
    }
}

// ---- <C> reduction ---------------------------------------------------------

template <typename R, int A_LOG>
__global__ void __launch_bounds__(256) k_expect(const typename CT<R>::T *__restrict__ sv,
                                                const TileArgs a, double *__restrict__ partials)
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

    const uint64_t toff = unpack_runs((uint64_t)t, a.truns);
    const T *p0 = sv + (tile_base(a, tile) + toff);
    T v[A];
#pragma unroll
    for (int j = 0; j < A; ++j) {
        uint64_t o = 0;
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            if ((j >> r) & 1)
                o |= 1ull << a.rpos[r];
        v[j] = p0[o];
    }
    cut_tile_prologue(a.plan, a.Q, a.nh, tile, s_c, &s_K, t, nthreads);
    __syncthreads();
    double cbase = s_K + a.plan.thr[t];
    for (int i = 0; i < a.nthr; ++i)
        cbase += s_c[i] * (double)((t >> i) & 1);
    double cr[A_LOG > 0 ? A_LOG : 1];
#pragma unroll
    for (int r = 0; r < A_LOG; ++r)
        cr[r] = s_c[a.nthr + r] + a.plan.thr[((1 + r) << a.nthr) + t];
    double acc_pc = 0.0, acc_p = 0.0;
#pragma unroll
    for (int j = 0; j < A; ++j) {
        double cut = cbase + a.plan.rr[j];
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            if ((j >> r) & 1)
                cut += cr[r];
        const double re = (double)v[j].x, im = (double)v[j].y;
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

__global__ void __launch_bounds__(1024) k_finalize(const double *__restrict__ partials,
                                                   uint64_t n, double *__restrict__ out)
{
    __shared__ double s_a[1024];
    __shared__ double s_b[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (uint64_t i = t; i < n; i += 1024) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    s_a[t] = a;
    s_b[t] = b;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            s_a[t] += s_a[t + w];
            s_b[t] += s_b[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = s_a[0];
        out[1] = s_b[0];
    }
}

// ---- block transposition ---------------------------------------------------

__device__ __forceinline__ uint64_t deposit_rest(uint64_t r, const SwapArgs &a)
{
    uint64_t o = 0;
    for (int i = 0; i < a.nr; ++i)
        o |= ((r >> i) & 1ull) << a.rpos[i];
    return o;
}

// out == 0: the low block is contiguous.  Work-group = one pair of 2^t x 2^t
// sub-tiles (X <= Y) of the 2^k x 2^k transpose for one rest index.
template <typename T>
__global__ void __launch_bounds__(256) k_swap_tiled(T *__restrict__ sv, const SwapArgs a)
{
    extern __shared__ __align__(16) unsigned char fq_smem[];
    const int side = 1 << a.t;
    const int stride = side + 1;
    T *la = reinterpret_cast<T *>(fq_smem);
    T *lb = la + side * stride;
    const uint64_t lin = blockIdx.x;
    const uint64_t p = lin % a.npairs;
    const uint64_t r = lin / a.npairs;
    // p -> (X <= Y): p = Y(Y+1)/2 + X
    uint64_t Y = (uint64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (Y * (Y + 1) / 2 > p)
        --Y;
    while ((Y + 1) * (Y + 2) / 2 <= p)
        ++Y;
    const uint64_t X = p - Y * (Y + 1) / 2;
    const uint64_t base = deposit_rest(r, a);
    const int n = side * side;
    const int s = a.in;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int xr = i >> a.t, yc = i & (side - 1);
        la[xr * stride + yc] = sv[base + ((X * side + xr) << s) + (Y * side + yc)];
        if (X != Y)
            lb[xr * stride + yc] = sv[base + ((Y * side + xr) << s) + (X * side + yc)];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int xr = i >> a.t, yc = i & (side - 1);
        if (X != Y) {
            sv[base + ((X * side + xr) << s) + (Y * side + yc)] = lb[yc * stride + xr];
            sv[base + ((Y * side + xr) << s) + (X * side + yc)] = la[yc * stride + xr];
        } else {
            sv[base + ((X * side + xr) << s) + (Y * side + yc)] = la[yc * stride + xr];
        }
    }
}

// Pair index p -> (X <= Y) with p = Y(Y+1)/2 + X.
__device__ __forceinline__ void pair_of(uint64_t p, uint64_t &X, uint64_t &Y)
{
    uint64_t y = (uint64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (y * (y + 1) / 2 > p)
        --y;
    while ((y + 1) * (y + 2) / 2 <= p)
        ++y;
    Y = y;
    X = p - y * (y + 1) / 2;
}

// Unit u of the pipelined transpose: the two tile origins and whether the
// pair is on the diagonal (one tile, transposed in place).
__device__ __forceinline__ void swap_unit(const SwapArgs &a, uint64_t u, uint64_t &pa, uint64_t &pb, bool &diag)
{
    uint64_t p, r;
    if (a.rest_fast) {
        r = u & ((1ull << a.nr) - 1ull);
        p = u >> a.nr;
    } else {
        p = u % a.npairs;
        r = u / a.npairs;
    }
    uint64_t X, Y;
    pair_of(p, X, Y);
    const uint64_t base = deposit_rest(r, a);
    pa = base + ((Y << a.t) << a.in) + (X << a.t);
    pb = base + ((X << a.t) << a.in) + (Y << a.t);
    diag = X == Y;
}

// out == 0, persistent and pipelined.  A unit is a pair of tiles (X <= Y)
// of 2^t x 2^m x 2^t amplitudes for one rest index: tile A has block-0 index
// X*2^t + xc, rest bits [k, k+m) = rl and block-s index Y*2^t + yc; tile B
// the same with X and Y exchanged.  Each side is read as 2^t runs of 2^(t+m)
// contiguous amplitudes (m > 0 only when the whole block is in the tile,
// t == k), exchanged through LDS (row stride 2^(m+t) + 1: conflict-free for
// the transposed read) and written back to the other tile.  The next unit's
// loads are issued before the current unit's LDS exchange.  256 threads,
// E amplitudes per thread per tile.
template <typename T> struct NativeVec;
template <> struct NativeVec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct NativeVec<float2> { typedef float type __attribute__((ext_vector_type(2))); };

// (register arrays of the clang vector type stay in VGPRs across the loop;
// arrays of HIP_vector_type<double, 2> went to scratch here)
template <typename T0, int E>
__global__ void __launch_bounds__(256) k_swap_pipe(T0 *__restrict__ sv0, const SwapArgs a, uint64_t nunits)
{
    typedef typename NativeVec<T0>::type T;
    T *__restrict__ sv = reinterpret_cast<T *>(sv0);
    extern __shared__ __align__(16) unsigned char fq_smem[];
    const int t = a.t, m = a.m, s = a.in, k = a.k;
    const int side = 1 << t;
    const int R = (side << m) + 1;
    T *la = reinterpret_cast<T *>(fq_smem);
    T *lb = la + (size_t)side * R;
    const int tid = threadIdx.x;

    // per-thread element coordinates (xc fastest, then rl, then yc)
    uint64_t off_g[E];
    int off_l[E], off_t[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = tid + e * 256;
        const int xc = i & (side - 1);
        const int rl = (i >> t) & ((1 << m) - 1);
        const int yc = i >> (t + m);
        off_g[e] = ((uint64_t)yc << s) + ((uint64_t)rl << k) + (uint64_t)xc;
        off_l[e] = yc * R + rl * side + xc;  // natural slot
        off_t[e] = xc * R + rl * side + yc;  // transposed slot
    }


    uint64_t u = blockIdx.x;
    T va[E], vb[E];
    uint64_t pa = 0, pb = 0;
    bool diag = true;
    if (u < nunits) {
        swap_unit(a, u, pa, pb, diag);
#pragma unroll
        for (int e = 0; e < E; ++e)
            va[e] = sv[pa + off_g[e]];
        if (!diag) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                vb[e] = sv[pb + off_g[e]];
        }
    }
    while (u < nunits) {
        const uint64_t next = u + gridDim.x;
        uint64_t na = 0, nb = 0;
        bool ndiag = true;
        T wa[E], wb[E];
        if (next < nunits) {
            swap_unit(a, next, na, nb, ndiag);
#pragma unroll
            for (int e = 0; e < E; ++e)
                wa[e] = sv[na + off_g[e]];
            if (!ndiag) {
#pragma unroll
                for (int e = 0; e < E; ++e)
                    wb[e] = sv[nb + off_g[e]];
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e)
            la[off_l[e]] = va[e];
        if (!diag) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                lb[off_l[e]] = vb[e];
        }
        lds_barrier();
        if (diag) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                sv[pa + off_g[e]] = la[off_t[e]];
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e)
                sv[pa + off_g[e]] = lb[off_t[e]];
#pragma unroll
            for (int e = 0; e < E; ++e)
                sv[pb + off_g[e]] = la[off_t[e]];
        }
        lds_barrier();  // the LDS reads finish before the next unit's writes
#pragma unroll
        for (int e = 0; e < E; ++e) {
            va[e] = wa[e];
            vb[e] = wb[e];
        }
        u = next;
        pa = na;
        pb = nb;
        diag = ndiag;
    }
}

// Any out: one thread per (rest, x, y); the x < y threads swap their pair.
template <typename T>
__global__ void __launch_bounds__(256) k_swap_generic(T *__restrict__ sv, const SwapArgs a,
                                                      uint64_t total)
{
    const uint64_t lin = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= total)
        return;
    const uint64_t km = (1ull << a.k) - 1;
    const uint64_t y = lin & km;
    const uint64_t x = (lin >> a.k) & km;
    if (x >= y)
        return;
    const uint64_t r = lin >> (2 * a.k);
    const uint64_t base = deposit_rest(r, a);
    const uint64_t i0 = base + (x << a.in) + (y << a.out);
    const uint64_t i1 = base + (y << a.in) + (x << a.out);
    const T tmp = sv[i0];
    sv[i0] = sv[i1];
    sv[i1] = tmp;
}

// ---- top-k -----------------------------------------------------------------

__device__ __forceinline__ bool better(double pa, uint64_t ia, double pb, uint64_t ib)
{
    return pa > pb || (pa == pb && ia < ib);
}

// Wave-wide best (p, i) by butterfly; every lane receives it and the lane
// it came from.
__device__ __forceinline__ void wave_best(double &p, uint64_t &i, int &from)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double op = __shfl_xor(p, off, 64);
        const uint64_t oi = ((uint64_t)(uint32_t)__shfl_xor((int)(uint32_t)i, off, 64)) |
                            ((uint64_t)(uint32_t)__shfl_xor((int)(i >> 32), off, 64) << 32);
        const int of = __shfl_xor(from, off, 64);
        if (better(op, oi, p, i)) {
            p = op;
            i = oi;
            from = of;
        }
    }
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)x, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(x >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t x, int d)
{
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)x, d, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(x >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Per-block top-16 (probability, index) candidates.  Each wave keeps its 16
// best in lanes 0..15 (one entry per lane, descending).  Amplitudes stream 8
// per thread per step (8 loads in flight); an amplitude is a candidate only if
// it beats the wave's 16th (one compare and a ballot), and candidates are
// inserted one at a time with lane shuffles.  One wave then merges the 4
// waves' 64 entries into the block's 16.
template <typename T>
__global__ void __launch_bounds__(256) k_topk(const T *__restrict__ sv, uint64_t n,
                                              double *__restrict__ cand_p,
                                              uint64_t *__restrict__ cand_i)
{
    constexpr int K = FQ_TOPK_MAX;
    constexpr int U = 8;
    __shared__ double s_p[4 * K];
    __shared__ uint64_t s_i[4 * K];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    double lp = -1.0;       // this lane's list entry (lanes < K)
    uint64_t li = ~0ull;
    double thr_p = -1.0;    // the wave's K-th best
    uint64_t thr_i = ~0ull;
    const uint64_t step = (uint64_t)gridDim.x * 256 * U;
    for (uint64_t base = (uint64_t)blockIdx.x * 256 * U + t; base < n; base += step) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t b = base + (uint64_t)u * 256;
            if (b < n)
                v[u] = sv[b];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t b = base + (uint64_t)u * 256;
            const double re = (double)v[u].x, im = (double)v[u].y;
            const double pb = re * re + im * im;
            uint64_t mask = __ballot(b < n && better(pb, b, thr_p, thr_i));
            while (mask) {
                const int src = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                const double cp = __shfl(pb, src, 64);
                const uint64_t ci = shfl_u64(b, src);
                const uint64_t gt = __ballot(lane < K && better(cp, ci, lp, li));
                if (gt) {
                    const int pos = __ffsll((unsigned long long)gt) - 1;
                    const double up = __shfl_up(lp, 1, 64);
                    const uint64_t ui = shfl_up_u64(li, 1);
                    if (lane > pos && lane < K) {
                        lp = up;
                        li = ui;
                    }
                    if (lane == pos) {
                        lp = cp;
                        li = ci;
                    }
                    thr_p = __shfl(lp, K - 1, 64);
                    thr_i = shfl_u64(li, K - 1);
                }
            }
        }
    }
    if (lane < K) {
        s_p[wave * K + lane] = lp;
        s_i[wave * K + lane] = li;
    }
    __syncthreads();
    if (wave == 0) {
        double cp = s_p[lane];
        uint64_t ci = s_i[lane];
        for (int r = 0; r < K; ++r) {
            double bp = cp;
            uint64_t bi = ci;
            int from = lane;
            wave_best(bp, bi, from);
            if (lane == 0) {
                cand_p[(uint64_t)blockIdx.x * K + r] = bp;
                cand_i[(uint64_t)blockIdx.x * K + r] = bi;
            }
            if (lane == from) {
                cp = -2.0;
                ci = ~0ull;
            }
        }
    }
}

// ---- small helpers -------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) k_gather(const T *__restrict__ sv,
                                                const uint64_t *__restrict__ idx, uint64_t n,
                                                double *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const T v = sv[idx[i]];
        out[2 * i] = (double)v.x;
        out[2 * i + 1] = (double)v.y;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_zero_one(T *__restrict__ sv, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        T v;
        v.x = (i == 0) ? 1 : 0;
        v.y = 0;
        sv[i] = v;
    }
}

// ---- launchers ---------------------------------------------------------------

int tile_max_q(int dtype) { return dtype == 0 ? 13 : 14; }

size_t tile_lds_bytes(int dtype, int Q)
{
    const size_t elem = dtype == 0 ? 16 : 8;
    const size_t n = ((size_t)1 << Q) + ((size_t)1 << Q) / 16;
    return n * elem;
}

int device_cu_count()
{
    static std::atomic<int> cus{0};
    int c = cus.load(std::memory_order_relaxed);
    if (c == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
            c = 256;
        cus.store(c, std::memory_order_relaxed);
    }
    return c;
}

static int device_cus() { return device_cu_count(); }

hipError_t allow_large_lds(const void *fn)
{
    static std::mutex mu;
    static std::map<const void *, hipError_t> done;
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(fn);
    if (it != done.end())
        return it->second;
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    done[fn] = e;
    return e;
}

int occupancy_per_cu(const void *fn, int threads, size_t lds)
{
    struct Key {
        const void *fn;
        int threads;
        size_t lds;
        bool operator<(const Key &o) const
        {
            return std::tie(fn, threads, lds) < std::tie(o.fn, o.threads, o.lds);
        }
    };
    static std::mutex mu;
    static std::map<Key, int> cache;
    const Key key{fn, threads, lds};
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end())
            return it->second;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) != hipSuccess || per_cu < 1) {
        hipGetLastError();
        per_cu = 1;
    }
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = per_cu;
    return per_cu;
}

template <typename R, int A_LOG, int QT, bool COST, bool FIRST, bool GATES, bool PRE = false, int KL = 0,
          bool LS = false>
static hipError_t launch_tile_t(void *sv, const TileArgs &a, const double2 *ptab, hipStream_t stream)
{
    typedef typename CT<R>::T T;
    const unsigned threads = 1u << a.nthr;
    const uint64_t tiles = run_tiles(a);
    const size_t lds = GATES && a.ngroups > 0 ? tile_lds_bytes(sizeof(R) == 8 ? 0 : 1, a.Q) : 0;
    const void *fn = reinterpret_cast<const void *>(&k_tile<R, A_LOG, QT, COST, FIRST, GATES, 0, PRE, KL, LS>);
    if (lds > 65536) {
        if (hipError_t e = allow_large_lds(fn))
            return e;
    }
    // persistent grid: as many work-groups as can be resident at once
    const int per_cu = occupancy_per_cu(fn, (int)threads, lds);
    const uint64_t resident = (uint64_t)per_cu * device_cus();
    // order 3: one tile per work-group (the dispatcher keeps a sliding window)
    const uint64_t grid = tiles < resident || a.order == 3 ? tiles : resident;
    TileArgs la = a;
    if (la.order == 1 && grid % 8 != 0)
        la.order = 0;
    hipLaunchKernelGGL((k_tile<R, A_LOG, QT, COST, FIRST, GATES, 0, PRE, KL, LS>), dim3((unsigned)grid), dim3(threads),
                       lds, stream, reinterpret_cast<T *>(sv), la, ptab);
    return hipGetLastError();
}

template <typename R, int A_LOG, int QT, int MIRROR>
static hipError_t launch_mirror_m(void *sv, const TileArgs &a, hipStream_t stream)
{
    typedef typename CT<R>::T T;
    const unsigned threads = 1u << a.nthr;
    const uint64_t tiles = a.ntiles_run;
    const size_t lds = a.ngroups > 0 ? tile_lds_bytes(sizeof(R) == 8 ? 0 : 1, a.Q) : 0;
    const void *fn = reinterpret_cast<const void *>(&k_tile<R, A_LOG, QT, false, false, true, MIRROR>);
    if (lds > 65536) {
        if (hipError_t e = allow_large_lds(fn))
            return e;
    }
    const int per_cu = occupancy_per_cu(fn, (int)threads, lds);
    const uint64_t resident = (uint64_t)per_cu * device_cus();
    const uint64_t grid = tiles < resident || a.order == 3 ? tiles : resident;
    TileArgs la = a;
    if (la.order == 1 && grid % 8 != 0)
        la.order = 0;
    hipLaunchKernelGGL((k_tile<R, A_LOG, QT, false, false, true, MIRROR>), dim3((unsigned)grid), dim3(threads),
                       lds, stream, reinterpret_cast<T *>(sv), la, nullptr);
    return hipGetLastError();
}

template <typename R, int A_LOG, int QT>
static hipError_t launch_mirror_t(bool expand, void *sv, const TileArgs &a, hipStream_t stream)
{
    return expand ? launch_mirror_m<R, A_LOG, QT, 2>(sv, a, stream) : launch_mirror_m<R, A_LOG, QT, 1>(sv, a, stream);
}

hipError_t launch_tile_mirror(int dtype, int a_log, bool expand, void *sv, const TileArgs &a, hipStream_t stream)
{
    if (a.ntiles_run == 0)
        return hipErrorInvalidValue;
    if (dtype == 0) {
        if (a_log == 3 && a.Q == 13)
            return launch_mirror_t<double, 3, 13>(expand, sv, a, stream);
        switch (a_log) {
            case 1: return launch_mirror_t<double, 1, 0>(expand, sv, a, stream);
            case 2: return launch_mirror_t<double, 2, 0>(expand, sv, a, stream);
            case 3: return launch_mirror_t<double, 3, 0>(expand, sv, a, stream);
            default: return hipErrorInvalidValue;
        }
    }
    if (a_log == 5 && a.Q == 14)
        return launch_mirror_t<float, 5, 14>(expand, sv, a, stream);
    if (a_log == 4 && a.Q == 14)
        return launch_mirror_t<float, 4, 14>(expand, sv, a, stream);
    switch (a_log) {
        case 1: return launch_mirror_t<float, 1, 0>(expand, sv, a, stream);
        case 2: return launch_mirror_t<float, 2, 0>(expand, sv, a, stream);
        case 3: return launch_mirror_t<float, 3, 0>(expand, sv, a, stream);
        case 4: return launch_mirror_t<float, 4, 0>(expand, sv, a, stream);
        default: return hipErrorInvalidValue;
    }
}

// Cost phase without gates (the reference's separate cost kernel): 256
// threads x up to 16 registers, no LDS.
template <typename R, int A_LOG, int QT>
static hipError_t launch_cost_only(bool first, void *sv, const TileArgs &a, const double2 *ptab,
                                   hipStream_t stream)
{
    return first ? launch_tile_t<R, A_LOG, QT, true, true, false>(sv, a, ptab, stream)
                 : launch_tile_t<R, A_LOG, QT, true, false, false>(sv, a, ptab, stream);
}

// The last LDS phase gates its two lowest register bits only (the complex64
// strided passes' 14 = 4 + 4 + 4 + 2 and 10 = 4 + 4 + 2 gates): the KL = 2
// instances skip the butterflies of its two filler bits (N=34 complex64:
// strided boundary pass 67.9 -> 64.8 ms, plain strided pass 52.5 -> 50.8 ms).
static bool last_phase_low2(const TileArgs &a)
{
    return a.ngroups >= 2 && a.grp_gates[a.ngroups - 1] == 3u;
}

// Gated passes, with or without the cost phase.
template <typename R, int A_LOG, int QT>
static hipError_t launch_gated(bool cost, bool first, void *sv, const TileArgs &a, const double2 *ptab,
                               hipStream_t stream)
{
    if (!cost) {
        if (a.lane_swap)
            return hipErrorInvalidValue;
        if constexpr (sizeof(R) == 4 && A_LOG == 4 && QT == 14)
            if (last_phase_low2(a))
                return launch_tile_t<R, A_LOG, QT, false, false, true, false, 2>(sv, a, ptab, stream);
        return launch_tile_t<R, A_LOG, QT, false, false, true>(sv, a, ptab, stream);
    }
    return first ? launch_tile_t<R, A_LOG, QT, true, true, true>(sv, a, ptab, stream)
                 : launch_tile_t<R, A_LOG, QT, true, false, true>(sv, a, ptab, stream);
}

template <typename R>
static hipError_t launch_tile_r(int a_log, bool cost, bool first, bool gates, void *sv,
                                const TileArgs &a, const double2 *ptab, hipStream_t stream)
{
    if (!gates) {
        if (a_log == 4 && a.Q == 12)
            return launch_cost_only<R, 4, 12>(first, sv, a, ptab, stream);
        switch (a_log) {
            case 1: return launch_cost_only<R, 1, 0>(first, sv, a, ptab, stream);
            case 2: return launch_cost_only<R, 2, 0>(first, sv, a, ptab, stream);
            case 3: return launch_cost_only<R, 3, 0>(first, sv, a, ptab, stream);
            case 4: return launch_cost_only<R, 4, 0>(first, sv, a, ptab, stream);
            default: return hipErrorInvalidValue;
        }
    }
    // compile-time tile widths for the full-width tiles of the fused passes;
    // anything else takes the generic instance.  Register bits per thread:
    // up to 3 for complex128, 4 for complex64 (5 in the 14-qubit tiles).
    if constexpr (sizeof(R) == 8) {
        if (a_log == 3 && a.Q == 13)
            return launch_gated<R, 3, 13>(cost, first, sv, a, ptab, stream);
        if (a_log == 4 && a.Q == 13 && cost && first)
            return launch_tile_t<R, 4, 13, true, true, true>(sv, a, ptab, stream);
        if (a_log == 3 && a.Q == 12)
            return launch_gated<R, 3, 12>(cost, first, sv, a, ptab, stream);
        switch (a_log) {
            case 1: return launch_gated<R, 1, 0>(cost, first, sv, a, ptab, stream);
            case 2: return launch_gated<R, 2, 0>(cost, first, sv, a, ptab, stream);
            case 3: return launch_gated<R, 3, 0>(cost, first, sv, a, ptab, stream);
            default: return hipErrorInvalidValue;
        }
    } else {
        if (a_log == 5 && a.Q == 14)
            return launch_gated<R, 5, 14>(cost, first, sv, a, ptab, stream);
        if (a_log == 4 && a.Q == 14)
            return launch_gated<R, 4, 14>(cost, first, sv, a, ptab, stream);
        if (a_log == 4 && a.Q == 13)
            return launch_gated<R, 4, 13>(cost, first, sv, a, ptab, stream);
        switch (a_log) {
            case 1: return launch_gated<R, 1, 0>(cost, first, sv, a, ptab, stream);
            case 2: return launch_gated<R, 2, 0>(cost, first, sv, a, ptab, stream);
            case 3: return launch_gated<R, 3, 0>(cost, first, sv, a, ptab, stream);
            case 4: return launch_gated<R, 4, 0>(cost, first, sv, a, ptab, stream);
            default: return hipErrorInvalidValue;
        }
    }
}

hipError_t launch_tile(int dtype, int a_log, bool cost, bool first, bool gates, void *sv,
                       const TileArgs &a, const double2 *ptab, hipStream_t stream)
{
    if (!cost && !gates)
        return hipSuccess;
    if (dtype == 0)
        return launch_tile_r<double>(a_log, cost, first, gates, sv, a, ptab, stream);
    return launch_tile_r<float>(a_log, cost, first, gates, sv, a, ptab, stream);
}

template <typename R>
static hipError_t launch_pre_r(int a_log, void *sv, const TileArgs &a, const double2 *ptab, hipStream_t stream)
{
    if constexpr (sizeof(R) == 8) {
        if (a_log == 3 && a.Q == 13)
            return launch_tile_t<R, 3, 13, true, false, true, true>(sv, a, ptab, stream);
        if (a_log == 3 && a.Q == 12)
            return launch_tile_t<R, 3, 12, true, false, true, true>(sv, a, ptab, stream);
    } else {
        if (a_log == 5 && a.Q == 14)
            return launch_tile_t<R, 5, 14, true, false, true, true>(sv, a, ptab, stream);
        if (a_log == 4 && a.Q == 14) {
            if (a.lane_swap)
                return launch_tile_t<R, 4, 14, true, false, true, true, 0, true>(sv, a, ptab, stream);
            return last_phase_low2(a) ? launch_tile_t<R, 4, 14, true, false, true, true, 2>(sv, a, ptab, stream)
                                      : launch_tile_t<R, 4, 14, true, false, true, true>(sv, a, ptab, stream);
        }
        if (a_log == 4 && a.Q == 13)
            return launch_tile_t<R, 4, 13, true, false, true, true>(sv, a, ptab, stream);
    }
    switch (a_log) {
        case 1: return launch_tile_t<R, 1, 0, true, false, true, true>(sv, a, ptab, stream);
        case 2: return launch_tile_t<R, 2, 0, true, false, true, true>(sv, a, ptab, stream);
        case 3: return launch_tile_t<R, 3, 0, true, false, true, true>(sv, a, ptab, stream);
        case 4:
            if constexpr (sizeof(R) == 4)
                return launch_tile_t<R, 4, 0, true, false, true, true>(sv, a, ptab, stream);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tile_pre(int dtype, int a_log, void *sv, const TileArgs &a, const double2 *ptab,
                           hipStream_t stream)
{
    if (dtype == 0)
        return launch_pre_r<double>(a_log, sv, a, ptab, stream);
    return launch_pre_r<float>(a_log, sv, a, ptab, stream);
}

// Streaming probe (fq_stream_probe): mode 0 scales n 8-byte elements in
// place, mode 1 reads them (sum kept live through sink), mode 2 fills them.
// One work-group per block of 1024 x 16 elements (128 KiB), no persistence:
// the form that streamed fastest among the shapes tried on MI355X.
__global__ void __launch_bounds__(1024) k_stream(float2 *__restrict__ p, uint64_t n, int mode, float scale,
                                                 float *__restrict__ sink)
{
    const uint64_t base = (uint64_t)blockIdx.x * (1024ull * 16) + threadIdx.x;
    float2 v[16];
    if (mode != 2) {
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (base + 1024ull * j < n)
                v[j] = p[base + 1024ull * j];
    }
    if (mode == 1) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (base + 1024ull * j < n)
                acc += v[j].x + v[j].y;
        if (acc == 1.0f)
            sink[0] = acc;
        return;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (base + 1024ull * j < n)
            p[base + 1024ull * j] = mode == 2 ? make_float2(scale, scale)
                                              : make_float2(v[j].x * scale, v[j].y * scale);
}

hipError_t launch_stream(void *p, size_t bytes, int mode, float *sink, hipStream_t stream)
{
    const uint64_t n = bytes / 8;
    const uint64_t blocks = (n + 1024 * 16 - 1) / (1024 * 16);
    if (blocks > 0xffffffffull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stream, dim3((unsigned)blocks), dim3(1024), 0, stream, reinterpret_cast<float2 *>(p), n,
                       mode, mode == 2 ? 0.0f : 1.0f, sink);
    return hipGetLastError();
}

size_t phase_table_entries(const TileArgs &a, int a_log)
{
    return (size_t)run_tiles(a) * (size_t)(a.Q + 1) + ((size_t)(1 + a_log) << a.nthr) + FQ_MAX_REGS;
}

hipError_t launch_phase_table(const TileArgs &a, int a_log, double2 *tab, double *cut,
                              hipStream_t stream)
{
    const size_t n = phase_table_entries(a, a_log);
    const size_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_phase_table, dim3((unsigned)blocks), dim3(256), 0, stream, a, a_log, tab,
                       cut);
    return hipGetLastError();
}

template <typename T, int E>
static hipError_t launch_swap_pipe(void *sv, const SwapArgs &a, hipStream_t stream)
{
    const int side = 1 << a.t;
    const size_t lds = 2 * (size_t)side * (((size_t)side << a.m) + 1) * sizeof(T);
    const void *fn = reinterpret_cast<const void *>(&k_swap_pipe<T, E>);
    if (lds > 65536) {
        if (hipError_t e = allow_large_lds(fn))
            return e;
    }
    const uint64_t nunits = a.npairs << a.nr;
    const int per_cu = occupancy_per_cu(fn, 256, lds);
    const uint64_t resident = (uint64_t)per_cu * device_cus();
    const uint64_t grid = nunits < resident ? nunits : resident;
    hipLaunchKernelGGL((k_swap_pipe<T, E>), dim3((unsigned)grid), dim3(256), lds, stream,
                       reinterpret_cast<T *>(sv), a, nunits);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_swap_t(void *sv, const SwapArgs &a, hipStream_t stream)
{
    if (a.pipe) {
        // tiles of 2^10 complex128 / 2^12 complex64 amplitudes
        constexpr int E = sizeof(T) == 16 ? 4 : 16;
        if ((1 << (2 * a.t + a.m)) != 256 * E)
            return hipErrorInvalidValue;
        return launch_swap_pipe<T, E>(sv, a, stream);
    }
    if (a.t >= 0) {
        const uint64_t nrest = 1ull << a.nr;
        const uint64_t blocks = a.npairs * nrest;
        const int side = 1 << a.t;
        const size_t lds = 2 * (size_t)side * (side + 1) * sizeof(T);
        hipLaunchKernelGGL(k_swap_tiled<T>, dim3((unsigned)blocks), dim3(256), lds, stream,
                           reinterpret_cast<T *>(sv), a);
    } else {
        const uint64_t total = 1ull << (a.nr + 2 * a.k);
        const uint64_t blocks = (total + 255) / 256;
        hipLaunchKernelGGL(k_swap_generic<T>, dim3((unsigned)blocks), dim3(256), 0, stream,
                           reinterpret_cast<T *>(sv), a, total);
    }
    return hipGetLastError();
}

hipError_t launch_swap(int dtype, void *sv, const SwapArgs &a, hipStream_t stream)
{
    return dtype == 0 ? launch_swap_t<double2>(sv, a, stream) : launch_swap_t<float2>(sv, a, stream);
}

template <typename R, int A_LOG>
static hipError_t launch_expect_t(const void *sv, const TileArgs &a, double *partials, double *out2,
                                  hipStream_t stream)
{
    typedef typename CT<R>::T T;
    const uint64_t tiles = run_tiles(a);
    hipLaunchKernelGGL((k_expect<R, A_LOG>), dim3((unsigned)tiles), dim3(1u << a.nthr), 0, stream,
                       reinterpret_cast<const T *>(sv), a, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, partials, tiles, out2);
    return hipGetLastError();
}

hipError_t launch_finalize(const double *partials, uint64_t n, double *out2, hipStream_t stream)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, partials, n, out2);
    return hipGetLastError();
}

hipError_t launch_expect(int dtype, int a_log, const void *sv, const TileArgs &a, double *partials,
                         double *out2, hipStream_t stream)
{
#define FQ_EXP(R)                                                              \
    switch (a_log) {                                                           \
        case 1: return launch_expect_t<R, 1>(sv, a, partials, out2, stream);   \
        case 2: return launch_expect_t<R, 2>(sv, a, partials, out2, stream);   \
        case 3: return launch_expect_t<R, 3>(sv, a, partials, out2, stream);   \
        case 4: return launch_expect_t<R, 4>(sv, a, partials, out2, stream);   \
        default: return hipErrorInvalidValue;                                  \
    }
    if (dtype == 0) {
        FQ_EXP(double)
    } else {
        FQ_EXP(float)
    }
#undef FQ_EXP
}

hipError_t launch_topk(int dtype, const void *sv, int N, int nblocks, double *cand_p,
                       uint64_t *cand_i, hipStream_t stream)
{
    const uint64_t n = 1ull << N;
    if (dtype == 0)
        hipLaunchKernelGGL(k_topk<double2>, dim3(nblocks), dim3(256), 0, stream,
                           reinterpret_cast<const double2 *>(sv), n, cand_p, cand_i);
    else
        hipLaunchKernelGGL(k_topk<float2>, dim3(nblocks), dim3(256), 0, stream,
                           reinterpret_cast<const float2 *>(sv), n, cand_p, cand_i);
    return hipGetLastError();
}

hipError_t launch_gather(int dtype, const void *sv, const uint64_t *idx, size_t n, double *out,
                         hipStream_t stream)
{
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (n == 0)
        return hipSuccess;
    if (dtype == 0)
        hipLaunchKernelGGL(k_gather<double2>, dim3(blocks), dim3(256), 0, stream,
                           reinterpret_cast<const double2 *>(sv), idx, (uint64_t)n, out);
    else
        hipLaunchKernelGGL(k_gather<float2>, dim3(blocks), dim3(256), 0, stream,
                           reinterpret_cast<const float2 *>(sv), idx, (uint64_t)n, out);
    return hipGetLastError();
}

hipError_t launch_fill_zero_one(int dtype, void *sv, int N, hipStream_t stream)
{
    const uint64_t n = 1ull << N;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    if (dtype == 0)
        hipLaunchKernelGGL(k_zero_one<double2>, dim3((unsigned)blocks), dim3(256), 0, stream,
                           reinterpret_cast<double2 *>(sv), n);
    else
        hipLaunchKernelGGL(k_zero_one<float2>, dim3((unsigned)blocks), dim3(256), 0, stream,
                           reinterpret_cast<float2 *>(sv), n);
    return hipGetLastError();
}

}  // namespace fq
