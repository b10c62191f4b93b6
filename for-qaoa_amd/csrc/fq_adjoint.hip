// fq_adjoint.hip -- adjoint (reverse-uncompute) gradient of <C> for the fused
// schedule.  No reference counterpart: the reference computes no expectation
// and no gradient (src/Weighted.cpp:153-163 prints timings only).
//
// Forward: psi_k = M(beta_k) U(gamma_k) psi_{k-1}, psi_0 = |+>^N,
//   U(g) = exp(i g S/2), S = diag(2 Cut - W)   (src/rzz.cu:146)
//   M(b) = exp(-i b/2 B), B = sum_j X_j          (src/rx.cu:46-49)
// E = <psi_P| Cut |psi_P>.  With lambda = Cut psi_P and both states walked
// back layer by layer (psi <- U^+ M^+ psi, lambda <- U^+ M^+ lambda):
//   dE/dbeta_k  =  Im <lambda_k | B | psi_k>
//   dE/dgamma_k = -Im <lambda'_k | S | phi_k>,   lambda'_k = M^+ lambda_k,
//                                                phi_k = M^+ psi_k.
// <lambda|X_j|psi> is invariant under any RX applied to both states, so each
// term is accumulated while qubit j's pairs sit in registers during the
// uncompute pass; <lambda|S|phi> commutes with the diagonal U and is
// accumulated in the pass that applies U^+.  FP64 accumulators, fixed-order
// reductions.
#include "fq_device.h"
#include "fq_internal.h"

#include <hip/hip_runtime.h>

namespace fq {

// Im(conj(l0) p1 + conj(l1) p0): the pair's contribution to Im<lambda|X_j|psi>,
// in the amplitude precision (products of two stored amplitudes)
template <typename T, typename R>
__device__ __forceinline__ R im_x_pair(const T &l0, const T &l1, const T &p0, const T &p1)
{
    return (l0.x * p1.y - l0.y * p1.x) + (l1.x * p0.y - l1.y * p0.x);
}

// RX on the register gate bits of both states; the A/2 pair terms of each
// gate bit are summed in the amplitude precision and added to the FP64
// accumulator once per bit.
template <typename T, typename R, int A_LOG>
__device__ __forceinline__ void rx_both(T *vp, T *vl, unsigned mask, R c, R s, double &acc)
{
    constexpr int A = 1 << A_LOG;
#pragma unroll
    for (int r = 0; r < A_LOG; ++r) {
        if ((mask >> r) & 1u) {
            R part = 0;
#pragma unroll
            for (int j = 0; j < A; ++j) {
                if (!((j >> r) & 1)) {
                    const int j1 = j | (1 << r);
                    part += im_x_pair<T, R>(vl[j], vl[j1], vp[j], vp[j1]);
                    rx_pair<T, R>(vp[j], vp[j1], c, s);
                    rx_pair<T, R>(vl[j], vl[j1], c, s);
                }
            }
            acc += (double)part;
        }
    }
}


// One uncompute pass over a tile geometry for psi and lambda together:
// RX(rx angle) on the gate bits (accumulating Im<lambda|X_j|psi>), then, if
// COSTBW, Im<lambda|S|phi> from the per-amplitude cut value and the diagonal
// phase on both.  Phase factors (built for -gamma) and cut coefficients come
// from the two tables of k_phase_table, same layout: s/2 of an amplitude is
// the sum of its tile's K entry, its tile coefficients c_i for the set thread
// and register bits, the thread's entry, its register entries and the
// register-register entry, and its phase is the product of the same entries
// of ptab.  partials[2*block] and [2*block+1] receive the work-group's two
// sums.
// Half states (run_tiles < 2^nh): the tiles with qubit N-1 = 0 only; MIRROR:
// qubit N-1 is the top register bit and the upper register half addresses
// tile ~tau at complemented offsets (psi(b) = psi(~b)), as in k_tile.  Every
// element then stands for itself and its complement: the caller doubles the
// sums.
// BOUNDARY (with COSTBW): the pass that undoes a layer-boundary pass of the
// pipelined forward (RX(beta_p), cost phase gamma_{p+1}, RX(beta_{p+1}) on its
// qubits): RX(rx angle) = RX(-beta_{p+1}), the conjugate phase, then RX with
// the coefficients a.pre_c, a.pre_s = RX(-beta_p) on the same gate bits,
// whose Im<lambda|X_j|psi> sum goes to partials[2*(gridDim.x + block)].
template <typename R, int A_LOG, bool COSTBW, bool MIRROR, bool BOUNDARY = false>
__global__ void __launch_bounds__(512) k_pair_pass(typename CT<R>::T *__restrict__ psi,
                                                    typename CT<R>::T *__restrict__ lam,
                                                    const TileArgs a,
                                                    const double2 *__restrict__ ptab,
                                                    const double *__restrict__ ctab,
                                                    double *__restrict__ partials)
{
    typedef typename CT<R>::T T;
    constexpr int A = 1 << A_LOG;
    extern __shared__ __align__(16) unsigned char fq_smem_a[];
    __shared__ double s_red[3][8];

    const int Q = a.Q, nthr = a.nthr;
    const int t = threadIdx.x;
    const int nthreads = blockDim.x;
    static_assert(!(COSTBW && MIRROR), "the cost pass is never the mirror pass");
    static_assert(!BOUNDARY || COSTBW, "a boundary pair pass has the cost phase");
    constexpr int AH = MIRROR ? A / 2 : A;
    const uint64_t ntiles = run_tiles(a);
    const uint64_t hmask = (1ull << a.nh) - 1ull;
    const uint64_t toff = unpack_runs((uint64_t)t, a.truns);
    // register offsets recomputed where the loads and stores use them
    const uint32_t rpos_w = a.rpos_w;
    constexpr int NB = MIRROR ? A_LOG - 1 : A_LOG;
    // both states' elements of a tile in registers (upper register half of a
    // mirror pass: the complementary tile at complemented offsets)
    auto load_tile = [&](uint64_t tl, T *p, T *l) {
        uint64_t roff[A];
        reg_offsets<A, NB>(rpos_w, roff);
        const uint64_t b0 = unpack_runs_uniform(tl, a.hruns) + toff;
#pragma unroll
        for (int j = 0; j < AH; ++j) {
            p[j] = psi[b0 + roff[j]];
            l[j] = lam[b0 + roff[j]];
        }
        if (MIRROR) {
            const uint64_t bm = unpack_runs_uniform(~tl & hmask, a.hruns) + a.mirror_pm - toff;
#pragma unroll
            for (int j = AH; j < A; ++j) {
                p[j] = psi[bm - roff[j - AH]];
                l[j] = lam[bm - roff[j - AH]];
            }
        }
    };
    const size_t lds_elems = ((size_t)1 << Q) + ((size_t)1 << Q) / 16;
    T *ldsP = reinterpret_cast<T *>(fq_smem_a);
    T *ldsL = ldsP + lds_elems;
    const R c = (R)a.rx_c, s = (R)a.rx_s;
    const int qmask = (1 << Q) - 1;
    const uint64_t q1 = (uint64_t)Q + 1;
    const uint64_t thr0 = ntiles * q1;                          // per-thread entries
    const uint64_t rr0 = thr0 + ((uint64_t)(1 + A_LOG) << nthr);  // register-register
    const int lane = t & 63;
    // A tile's Q+1 table entries sit in lanes 0..Q and are broadcast with
    // readlane when the wave has that many lanes (and A for the register-
    // register entries); otherwise they are read from the tables.
    const bool rowpf = (1 << nthr) >= (Q + 1 > A ? Q + 1 : A);
    // tile-independent entries of this thread, loaded once
    double c_thr = 0.0, c_rr = 0.0;
    double2 f_thr = make_double2(1.0, 0.0), f_rr = make_double2(1.0, 0.0);
    double c_reg[A_LOG > 0 ? A_LOG : 1];
    double2 f_reg[A_LOG > 0 ? A_LOG : 1];
    if (COSTBW) {
        c_thr = landed(ctab[thr0 + t]);
        f_thr = landed(ptab[thr0 + t]);
#pragma unroll
        for (int r = 0; r < A_LOG; ++r) {
            c_reg[r] = landed(ctab[thr0 + ((uint64_t)(1 + r) << nthr) + t]);
            f_reg[r] = landed(ptab[thr0 + ((uint64_t)(1 + r) << nthr) + t]);
        }
        if (rowpf) {
            c_rr = landed(ctab[rr0 + lane % A]);
            f_rr = landed(ptab[rr0 + lane % A]);
        }
    }
    // the tile's row of both tables (lane i: entry i), and the next tile's
    double crow = 0.0, cnext = 0.0;
    double2 frow = make_double2(1.0, 0.0), fnext = make_double2(1.0, 0.0);
    const bool row_lane = rowpf && lane <= Q;

    double accX = 0.0, accG = 0.0, accX2 = 0.0;
    // one contiguous range of tiles per work-group; the next tile's
    // amplitudes load while the current one is in its LDS phases (not in the
    // complex128 cost pass, whose registers are full without it)
    constexpr bool PREFETCH = !(COSTBW && sizeof(R) == 8);
    const uint64_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint64_t first_tile = (uint64_t)blockIdx.x * per;
    const uint64_t end_tile = first_tile + per < ntiles ? first_tile + per : ntiles;
    T vp[A], vl[A];
    if (first_tile < end_tile) {
        load_tile(first_tile, vp, vl);
        if (COSTBW && row_lane) {
            crow = ctab[first_tile * q1 + lane];
            frow = ptab[first_tile * q1 + lane];
        }
    }
    for (uint64_t tile = first_tile; tile < end_tile; ++tile) {
        const uint64_t base = unpack_runs_uniform(tile, a.hruns) + toff;
        T wp[A], wl[A];
        if (!PREFETCH && tile != first_tile) {
            load_tile(tile, vp, vl);
            if (COSTBW && row_lane) {
                crow = ctab[tile * q1 + lane];
                frow = ptab[tile * q1 + lane];
            }
        }
        // the next tile's loads go out after the first LDS phase (after the
        // register RX where there is none), not right after the previous
        // tile's stores: measured on MI355X, N=26 P=8 complex64 adjoint
        // 109.3 -> 112.8 gradients/s (pair passes 3.77 -> 3.93 TB/s)
        bool pf_left = PREFETCH && tile + 1 < end_tile;
        auto prefetch = [&]() {
            if (pf_left) {
                load_tile(tile + 1, wp, wl);
                if (COSTBW && row_lane) {
                    cnext = ctab[(tile + 1) * q1 + lane];
                    fnext = ptab[(tile + 1) * q1 + lane];
                }
                pf_left = false;
            }
        };
        const int nat = lds_pad(t);
        // RX(cc, ss) on every gate bit of both states, its pair sums into
        // acc: the register bits, then the LDS phases (the next tile's loads
        // after the first one when `loads`); returns in the natural layout.
        // Each thread's natural slots are its own, so a second set needs no
        // barrier before its first write.
        auto rx_set = [&](R cc, R ss, double &acc, bool loads) {
        rx_both<T, R, A_LOG>(vp, vl, a.gate_mask >> nthr, cc, ss, acc);
        if (a.ngroups > 0) {
#pragma unroll
            for (int j = 0; j < A; ++j) {
                ldsP[nat + lds_pad(j << nthr)] = vp[j];
                ldsL[nat + lds_pad(j << nthr)] = vl[j];
            }
        }
        lds_barrier();
        for (int g = 0; g < a.ngroups; ++g) {
            const unsigned gm = group_mask(a, g);
            const int etp = lds_pad(deposit_runs(t, (unsigned)qmask & ~gm));
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
            for (int j = 0; j < A; ++j) {
                vp[j] = ldsP[etp + ej[j]];
                vl[j] = ldsL[etp + ej[j]];
            }
            rx_both<T, R, A_LOG>(vp, vl, group_gates(a, g), cc, ss, acc);
#pragma unroll
            for (int j = 0; j < A; ++j) {
                ldsP[etp + ej[j]] = vp[j];
                ldsL[etp + ej[j]] = vl[j];
            }
            lds_barrier();
            if (loads && g == 0)
                prefetch();
        }
        if (loads)
            prefetch();
        if (a.ngroups > 0) {
#pragma unroll
            for (int j = 0; j < A; ++j) {
                vp[j] = ldsP[nat + lds_pad(j << nthr)];
                vl[j] = ldsL[nat + lds_pad(j << nthr)];
            }
        }
        };
        rx_set(c, s, accX, true);
        if (COSTBW) {
            const double *c_t = ctab + tile * q1;
            const double2 *f_t = ptab + tile * q1;
            auto crow_at = [&](int i) -> double { return rowpf ? lane_bcast(crow, i) : c_t[i]; };
            auto frow_at = [&](int i) -> double2 { return rowpf ? lane_bcast(frow, i) : f_t[i]; };
            // s/2 and the phase factor of register 0: tile K entry, the
            // thread's entry and the tile coefficients of the set thread bits
            double half = crow_at(Q) + c_thr;
            double x = f_thr.x, y = f_thr.y;
            {
                const double2 f = frow_at(Q);
                const double nx = x * f.x - y * f.y;
                y = x * f.y + y * f.x;
                x = nx;
            }
            for (int i = 0; i < nthr; ++i) {
                const double ci = crow_at(i);
                const double2 f = frow_at(i);
                const bool on = (t >> i) & 1;
                const double nx = x * f.x - y * f.y;
                const double ny = x * f.y + y * f.x;
                half += on ? ci : 0.0;
                x = on ? nx : x;
                y = on ? ny : y;
            }
            double cr[A_LOG > 0 ? A_LOG : 1];
            double2 fr[A_LOG > 0 ? A_LOG : 1];
#pragma unroll
            for (int r = 0; r < A_LOG; ++r) {
                cr[r] = crow_at(nthr + r) + c_reg[r];
                const double2 fq = frow_at(nthr + r);
                fr[r] = make_double2(f_reg[r].x * fq.x - f_reg[r].y * fq.y,
                                     f_reg[r].x * fq.y + f_reg[r].y * fq.x);
            }
#pragma unroll
            for (int j = 0; j < A; ++j) {
                double h = half + (rowpf ? lane_bcast(c_rr, j) : ctab[rr0 + j]);
#pragma unroll
                for (int r = 0; r < A_LOG; ++r)
                    if ((j >> r) & 1)
                        h += cr[r];
                // s = 2 Cut - W = 2 h;  Im(conj(l) * s * p)
                const R im = vl[j].x * vp[j].y - vl[j].y * vp[j].x;
                accG += 2.0 * h * (double)im;
                cmul_by(vp[j], (R)x, (R)y);
                cmul_by(vl[j], (R)x, (R)y);
            }
#pragma unroll
            for (int r = 0; r < A_LOG; ++r) {
#pragma unroll
                for (int j = 0; j < A; ++j)
                    if ((j >> r) & 1) {
                        cmul_by(vp[j], (R)fr[r].x, (R)fr[r].y);
                        cmul_by(vl[j], (R)fr[r].x, (R)fr[r].y);
                    }
            }
            if (a.has_rr) {
#pragma unroll
                for (int j = 0; j < A; ++j) {
                    const double2 f = rowpf ? lane_bcast(f_rr, j) : ptab[rr0 + j];
                    cmul_by(vp[j], (R)f.x, (R)f.y);
                    cmul_by(vl[j], (R)f.x, (R)f.y);
                }
            }
        }
        if constexpr (BOUNDARY)
            rx_set((R)a.pre_c, (R)a.pre_s, accX2, false);
        {
            uint64_t roff[A];
            reg_offsets<A, NB>(rpos_w, roff);
#pragma unroll
            for (int j = 0; j < AH; ++j) {
                psi[base + roff[j]] = vp[j];
                lam[base + roff[j]] = vl[j];
            }
            if (MIRROR) {
                const uint64_t bm = unpack_runs_uniform(~tile & hmask, a.hruns) + a.mirror_pm - toff;
#pragma unroll
                for (int j = AH; j < A; ++j) {
                    psi[bm - roff[j - AH]] = vp[j];
                    lam[bm - roff[j - AH]] = vl[j];
                }
            }
        }
        // the next tile's LDS phases start after every thread has read its
        // amplitudes of this one back
        lds_barrier();
        if (PREFETCH) {
#pragma unroll
            for (int j = 0; j < A; ++j) {
                vp[j] = wp[j];
                vl[j] = wl[j];
            }
            crow = cnext;
            frow = fnext;
        }
    }
    // fixed-order work-group reduction: waves by shuffle, then 16 slots
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        accX += __shfl_xor(accX, off, 64);
        accG += __shfl_xor(accG, off, 64);
        if (BOUNDARY)
            accX2 += __shfl_xor(accX2, off, 64);
    }
    const int wave = t >> 6;
    if (lane == 0) {
        s_red[0][wave] = accX;
        s_red[1][wave] = accG;
        s_red[2][wave] = accX2;
    }
    __syncthreads();
    if (t == 0) {
        double x = 0.0, g = 0.0, x2 = 0.0;
        for (int w = 0; w < (nthreads + 63) / 64; ++w) {
            x += s_red[0][w];
            g += s_red[1][w];
            x2 += s_red[2][w];
        }
        partials[2 * blockIdx.x] = x;
        partials[2 * blockIdx.x + 1] = g;
        if (BOUNDARY) {
            partials[2 * (gridDim.x + blockIdx.x)] = x2;
            partials[2 * (gridDim.x + blockIdx.x) + 1] = 0.0;
        }
    }
}

// lambda = Cut * psi, and the partial sums of sum |psi|^2 Cut and sum |psi|^2.
template <typename R, int A_LOG>
__global__ void __launch_bounds__(256) k_cut_mul(const typename CT<R>::T *__restrict__ psi,
                                                 typename CT<R>::T *__restrict__ lam, const TileArgs a,
                                                 double *__restrict__ partials)
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
    const uint64_t base = unpack_runs(tile, a.hruns) + unpack_runs((uint64_t)t, a.truns);
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
        uint64_t o = 0;
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            if ((j >> r) & 1)
                o |= 1ull << a.rpos[r];
        double cut = cbase + a.plan.rr[j];
#pragma unroll
        for (int r = 0; r < A_LOG; ++r)
            if ((j >> r) & 1)
                cut += cr[r];
        const T v = psi[base + o];
        const double re = (double)v.x, im = (double)v.y;
        const double pr = re * re + im * im;
        acc_pc += pr * cut;
        acc_p += pr;
        T w;
        w.x = (R)(re * cut);
        w.y = (R)(im * cut);
        lam[base + o] = w;
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

size_t pair_lds_bytes(int dtype, int Q)
{
    const size_t elem = dtype == 0 ? 16 : 8;
    return 2 * (((size_t)1 << Q) + ((size_t)1 << Q) / 16) * elem;
}

// Persistent grid: as many work-groups as can be resident (one partial sum
// pair each, at most the tile count).
template <typename R, int A_LOG, bool COSTBW, bool MIRROR, bool BOUNDARY>
static hipError_t launch_pair_t(void *psi, void *lam, const TileArgs &a, const double2 *ptab,
                                const double *ctab, double *partials, unsigned *grid_out,
                                hipStream_t stream)
{
    typedef typename CT<R>::T T;
    const size_t lds = pair_lds_bytes(sizeof(R) == 8 ? 0 : 1, a.Q);
    const void *fn = reinterpret_cast<const void *>(&k_pair_pass<R, A_LOG, COSTBW, MIRROR, BOUNDARY>);
    if (lds > 65536) {
        if (hipError_t e = allow_large_lds(fn))
            return e;
    }
    const unsigned threads = 1u << a.nthr;
    const int per_cu = occupancy_per_cu(fn, (int)threads, lds);
    const uint64_t tiles = run_tiles(a);
    const uint64_t resident = (uint64_t)per_cu * (uint64_t)device_cu_count();
    const unsigned grid = (unsigned)(tiles < resident ? tiles : resident);
    *grid_out = grid;
    hipLaunchKernelGGL((k_pair_pass<R, A_LOG, COSTBW, MIRROR, BOUNDARY>), dim3(grid), dim3(threads), lds, stream,
                       reinterpret_cast<T *>(psi), reinterpret_cast<T *>(lam), a, ptab, ctab, partials);
    return hipGetLastError();
}

hipError_t launch_pair(int dtype, int a_log, bool costbw, bool mirror, bool boundary, void *psi, void *lam,
                       const TileArgs &a, const double2 *ptab, const double *ctab, double *partials,
                       unsigned *grid, hipStream_t stream)
{
    if ((costbw && mirror) || (boundary && !costbw))
        return hipErrorInvalidValue;
#define FQ_PAIR(R, AL)                                                                                        \
    return boundary ? launch_pair_t<R, AL, true, false, true>(psi, lam, a, ptab, ctab, partials, grid, stream)  \
           : costbw ? launch_pair_t<R, AL, true, false, false>(psi, lam, a, ptab, ctab, partials, grid, stream) \
           : mirror ? launch_pair_t<R, AL, false, true, false>(psi, lam, a, ptab, ctab, partials, grid, stream) \
                    : launch_pair_t<R, AL, false, false, false>(psi, lam, a, ptab, ctab, partials, grid, stream);
    if (dtype == 0) {
        switch (a_log) {
            case 1: FQ_PAIR(double, 1)
            case 2: FQ_PAIR(double, 2)
            case 3: FQ_PAIR(double, 3)
            default: return hipErrorInvalidValue;
        }
    }
    switch (a_log) {
        case 1: FQ_PAIR(float, 1)
        case 2: FQ_PAIR(float, 2)
        case 3: FQ_PAIR(float, 3)
        default: return hipErrorInvalidValue;
    }
#undef FQ_PAIR
}

hipError_t launch_cut_mul(int dtype, int a_log, const void *psi, void *lam, const TileArgs &a,
                          double *partials, hipStream_t stream)
{
    const uint64_t tiles = run_tiles(a);
    const dim3 g((unsigned)tiles), b(1u << a.nthr);
#define FQ_CM(R, AL)                                                                                       \
    hipLaunchKernelGGL((k_cut_mul<R, AL>), g, b, 0, stream,                                               \
                       reinterpret_cast<const typename CT<R>::T *>(psi),                                 \
                       reinterpret_cast<typename CT<R>::T *>(lam), a, partials);                         \
    return hipGetLastError();
    if (dtype == 0) {
        switch (a_log) {
            case 1: { FQ_CM(double, 1) }
            case 2: { FQ_CM(double, 2) }
            case 3: { FQ_CM(double, 3) }
            case 4: { FQ_CM(double, 4) }
            default: return hipErrorInvalidValue;
        }
    }
    switch (a_log) {
        case 1: { FQ_CM(float, 1) }
        case 2: { FQ_CM(float, 2) }
        case 3: { FQ_CM(float, 3) }
        case 4: { FQ_CM(float, 4) }
        default: return hipErrorInvalidValue;
    }
#undef FQ_CM
}

}  // namespace fq
