// fq_internal.h -- structures shared by the gfx950 kernels and the host side of
// libforqaoa.  Not part of the public ABI (include/forqaoa.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstddef>

#define FQ_MAXN 40         // qubits (2^40 complex64 = 8 TiB: the HBM limit bites first)
#define FQ_MAXQ 14         // tile qubits held in LDS (2^14 complex64 = 128 KiB)
#define FQ_MAX_GROUPS 5    // LDS re-mapping phases per tile pass
#define FQ_MAX_RUNS 8      // bit runs of a scattered index
#define FQ_TOPK_MAX 16
#define FQ_MAX_REG_LOG 5   // register bits per thread of a tile pass
#define FQ_MAX_REGS (1 << FQ_MAX_REG_LOG)

namespace fq {

// Device view of a cost plan: the Max-Cut quadratic form split for one tile
// geometry.  For a tile (2^Q amplitudes whose index bits at tile positions
// vary) with the other N-Q "high" bits fixed by the tile index:
//   Cut(z) = K(zH) + sum_{tile bit i} c_i(zH) z_i + Cut_tile(z_tile)
//   c_i(zH) = cH[i] - 2 sum_{(h,w) in csr(i)} w zH_h
//   K(zH)  = sum_{hh edges} w [zH_u != zH_v] + sum_h omega[h] zH_h
// The tile bits split into nthr thread bits and A_LOG register bits; the
// thread-dependent, tile-independent part comes from per-thread tables
//   thr[t]                   = edges inside the thread bits + register-thread
//                              edges' constant parts
//   thr[(1+r)*2^nthr + t]    = d_r(t) (coefficient of register bit r)
// and the register-register edges from rr[j] (j = register index).
struct PlanDev {
    const double *cH;      // [Q]
    const int *csr_off;    // [Q+1]
    const int *csr_h;      // [nnz] high-bit index
    const double *csr_w;   // [nnz]
    const int *hh_u;       // [n_hh] high-bit index
    const int *hh_v;       // [n_hh]
    const double *hh_w;    // [n_hh]
    const double *omega;   // [nh]
    const double *thr;     // [(1 + A_LOG) << nthr]
    int n_hh;
    int pad_;
    double wtot;           // sum of all weights
    double rr[FQ_MAX_REGS];  // register-register cut contributions
};

// Index bits [src, src+len) go to physical positions [dst, dst+len); one run
// packed per word as src | len << 8 | dst << 16.
struct BitRuns {
    int n;
    uint32_t run[FQ_MAX_RUNS];
};

// One tile pass: load 2^Q amplitudes per work-group (tile bits tpos[], high
// bits hpos[]), optionally apply the cost phase, RX on the gate bits, store.
struct TileArgs {
    int N, Q, nthr, nh;
    BitRuns hruns;           // tile index -> high positions
    BitRuns truns;           // thread index -> positions of the thread bits
    uint8_t rpos[8];         // positions of the register bits
    uint32_t rpos_w;         // the same packed, 6 bits each (rpos[r] at bit 6r)
    uint64_t ntiles_run;     // tiles processed (0: all 2^nh)
    uint64_t mirror_pm;      // mirror pass: sum of 2^p over the real tile positions
    uint32_t order;          // tile order of the persistent loop (tile_at)
    uint32_t gate_mask;      // over tile bits
    int ngroups;
    uint16_t grp_mask[FQ_MAX_GROUPS];  // tile-bit mask of each LDS phase (A_LOG bits)
    uint8_t grp_gates[FQ_MAX_GROUPS];  // register bits of that phase that receive RX
    uint64_t grp_mask_w;     // grp_mask[0..3], 16 bits each (read with a shift, not an index)
    uint32_t grp_gates_w;    // grp_gates[0..4], 5 bits each
    uint32_t direct_store;   // last LDS phase holds no lane bit: store from its layout
    uint32_t lane_swap;      // RX on thread bits 4 and 5 by cross-lane swaps, not in an LDS phase
    uint32_t dpos_w;         // direct_store: position of the last phase's register bit r (bit 6r)
    double rx_c, rx_s;       // cos(beta/2), sin(beta/2) (host-computed, src/rx.cu:121)
    double pre_c, pre_s;     // layer-boundary pass: RX of the previous layer, applied first
    // k_tile applies RX up to a factor as the butterfly [[1, -it], [-it, 1]]
    // (rx_tform), one FMA per output component
    double rx_t, pre_t;
    float rx_tf, pre_tf;     // the same in FP32 (scalar operands of the complex64 butterflies)
    double scale_re, scale_im;  // factor of the pass: folded into the cost phase,
    int apply_scale;            // or (apply_scale) multiplied in before the stores
    double gamma, init_amp;  // cost phase angle, launch-control amplitude
    int has_rr;              // any register-register edge in the plan
    PlanDev plan;
};

// Tiles a launch processes (a Z2-reduced state runs the lower half only).
__host__ __device__ inline uint64_t run_tiles(const TileArgs &a)
{
    return a.ntiles_run ? a.ntiles_run : (1ull << a.nh);
}

// Per-slot parameters of a packed (slab) step, in device memory.
struct SlotParams {
    double rx_c, rx_s;       // RX(beta) of the slot's current layer
    double gamma;            // cost angle of the slot's current layer
    double init_amp;         // (1/sqrt 2)^N
    int active;              // the slot runs this step
    int cost_first;          // layer 0: launch-control write, no read
    int has_rr;
    int want_expect;         // the slot finishes this step: reduce <C>
    const double2 *ptab;     // the slot's phase table for this step
    PlanDev plan;            // cut plan of the first (contiguous) tile pass
    PlanDev eplan;           // cut plan of the <C> reduction geometry
};

struct SwapArgs {
    int N, out, in, k;
    int t;                   // tiled: sub-tile log2 side (-1: one thread per pair)
    int m;                   // pipelined: rest bits [k, k+m) inside the tile (t == k)
    int pipe;                // pipelined persistent kernel (1) or one pair per work-group (0)
    int rest_fast;           // pipelined: consecutive units step the rest index first
    int nr;                  // number of rest bits (outside the tile)
    uint8_t rpos[FQ_MAXN];   // rest bit -> physical position
    uint64_t npairs;         // tiled: number of (X <= Y) sub-tile pairs
};

// Resident work-groups per CU of kernel fn at this block size and dynamic
// LDS (at least 1), cached per (fn, threads, lds): the runtime query costs
// tens of microseconds and the persistent launchers ask on every launch.
int occupancy_per_cu(const void *fn, int threads, size_t lds);
// Lets kernel fn take up to 159 KiB of dynamic LDS (set once per kernel).
hipError_t allow_large_lds(const void *fn);
// Compute units of the current device (cached).
int device_cu_count();

// Launchers (fq_kernels.hip).  All enqueue on `stream` and return the HIP
// launch status.
hipError_t launch_tile(int dtype, int a_log, bool cost, bool first, bool gates, void *sv,
                       const TileArgs &a, const double2 *phase_tab, hipStream_t stream);
// Layer-boundary pass: RX(pre) on the pass's qubits, cost phase, RX(rx).
hipError_t launch_tile_pre(int dtype, int a_log, void *sv, const TileArgs &a, const double2 *ptab,
                           hipStream_t stream);
// Mirror pass of a Z2-reduced state: the top register bit of the virtual tile
// selects the complementary tile (complemented element offsets) and RX on it
// is the butterfly of the global spin flip.  expand: also write the upper
// half of the full state.
hipError_t launch_tile_mirror(int dtype, int a_log, bool expand, void *sv, const TileArgs &a,
                              hipStream_t stream);
// Phase factors of a cost-phase launch into tab (size phase_table_entries);
// with cut not null, their exponents over gamma into cut (same size).
hipError_t launch_phase_table(const TileArgs &a, int a_log, double2 *tab, double *cut,
                              hipStream_t stream);
size_t phase_table_entries(const TileArgs &a, int a_log);
hipError_t launch_swap(int dtype, void *sv, const SwapArgs &a, hipStream_t stream);
hipError_t launch_expect(int dtype, int a_log, const void *sv, const TileArgs &a, double *partials,
                         double *out2, hipStream_t stream);
hipError_t launch_topk(int dtype, const void *sv, int N, int nblocks, double *cand_p,
                       uint64_t *cand_i, hipStream_t stream);
hipError_t launch_gather(int dtype, const void *sv, const uint64_t *idx, size_t n, double *out,
                         hipStream_t stream);
hipError_t launch_fill_zero_one(int dtype, void *sv, int N, hipStream_t stream);
// streaming probe over bytes (a multiple of 8): mode 0 in-place scale,
// 1 read-only, 2 fill; sink: one float
hipError_t launch_stream(void *p, size_t bytes, int mode, float *sink, hipStream_t stream);
hipError_t launch_finalize(const double *partials, uint64_t n, double *out2, hipStream_t stream);
// adjoint (fq_adjoint.hip)
// grid receives the number of work-groups (= partial sum pairs written).
// With costbw, ptab and ctab are the phase and cut tables of one
// launch_phase_table call.
// mirror: the half-state mirror pass (qubit N-1 the top register bit).
// boundary (with costbw): after the conjugate phase, RX with a.pre_c, a.pre_s
// on the same gate bits; its sums go to partials[2 * (*grid) ...], the
// work-group's second pair of partial sums.
hipError_t launch_pair(int dtype, int a_log, bool costbw, bool mirror, bool boundary, void *psi, void *lam,
                       const TileArgs &a, const double2 *ptab, const double *ctab, double *partials,
                       unsigned *grid, hipStream_t stream);
hipError_t launch_cut_mul(int dtype, int a_log, const void *psi, void *lam, const TileArgs &a,
                          double *partials, hipStream_t stream);
size_t pair_lds_bytes(int dtype, int Q);
// packed slabs (fq_serve.hip)
hipError_t launch_slab_phase_table(const TileArgs &local, int a_log, const SlotParams *slots, int nslots,
                                   double2 *tabs, size_t stride, hipStream_t stream);
hipError_t launch_slab_pass(int dtype, int a_log, bool cost, bool mirror, void *sv, const TileArgs &a,
                            const SlotParams *slots, int slot_shift, hipStream_t stream);
hipError_t launch_slab_expect(int dtype, int a_log, const void *sv, const TileArgs &a, const SlotParams *slots,
                              int slot_shift, int nslots, double *partials, double *out, hipStream_t stream);
int tile_max_q(int dtype);
size_t tile_lds_bytes(int dtype, int Q);

// RX(beta) coefficients c = cos(beta/2), s = sin(beta/2) (src/rx.cu:121) as
// the kernels of amplitude type dtype apply them.  FQ_C128: the FP64 values.
// FQ_C64: an FP32 pair (returned as doubles, exact when cast back) chosen so
// that c^2 + s^2 = 1 as nearly as FP32 allows.  Rounding cos and sin to FP32
// independently leaves c^2 + s^2 - 1 up to ~1e-7, and RX scales the norm of
// the state by exactly c^2 + s^2, so that error compounds over every qubit
// and layer into a drift of N*P*(c^2 + s^2 - 1) in <C>.
void rx_coeffs(double beta, int dtype, double *c, double *s);
// RX(beta) as sigma * B(t), B(t) = [[1, -it], [-it, 1]] with t rounded to the
// amplitude type of dtype and sigma (FP64, from the rounded t) making
// sigma * B(t) exactly unitary: RX of an angle within one rounding of beta.
// t = tan(beta/2), sigma = cos(beta/2)-like, real.  With flip and
// |sin(beta/2)| > |cos(beta/2)| instead RX(beta) = -i sin(beta/2) X B(t') with
// t' = tan((beta - pi)/2) = -cot(beta/2): the X is dropped and sigma is
// -i/sqrt(1 + t'^2) (signed).  Dropping X is exact only where the caller
// applies RX(beta) to every qubit of an X^(x)N-symmetric state (a QAOA layer:
// the dropped X's form X^(x)N, which commutes with every layer and leaves the
// state unchanged).  |t| <= 1 with flip; without it |t| = |tan(beta/2)|.
void rx_tform(double beta, int dtype, bool flip, double *t, double *sigma_re, double *sigma_im);

}  // namespace fq
