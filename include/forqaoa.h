/*
 * forqaoa.h -- C ABI of the MI355X (gfx950) QAOA Max-Cut state-vector library.
 *
 * The reference (FOR-QAOA) has no FFI: its hot path is a set of C++ free
 * functions called by src/Weighted.cpp.  Each entry point below names the
 * reference function it replaces (file:line, relative to the FOR-QAOA root).
 * Plain pointers and sizes only; every call returns an int status
 * (FQ_OK = 0) and fq_last_error() gives the message of the calling thread's
 * last failure.  The reference exits the process on any error
 * (include/helper_cuda.hpp:588-596); the `weighted` driver maps a non-zero
 * status to exit(1) to keep that CLI behaviour.
 *
 * Conventions (as the reference):
 *  - amplitude index b = sum_i z_i 2^i: physical qubit i is bit i;
 *  - amplitudes are interleaved {re, im}: complex128 = 2 doubles,
 *    complex64 = 2 floats;
 *  - the cost phase multiplies psi(b) by exp(i*gamma*s(b)/2) with
 *    s(b) = sum_{i<j} W_ij * (z_i != z_j ? +1 : -1)  (src/rzz.cu:131-155);
 *  - the mixer is RX(beta) = exp(-i*beta/2*X) on each target
 *    (src/rx.cu:46-49, 121).
 *
 * Threading: all work of a state is queued on that state's HIP stream; calls
 * return without waiting except where noted (reductions, copies, fq_sync).
 * One host thread per state handle; distinct handles may be used from
 * distinct threads.  A graph may be shared by states used from one thread.
 */
#ifndef FORQAOA_H
#define FORQAOA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FQ_OK 0
#define FQ_ERR_ARG 1         /* invalid argument */
#define FQ_ERR_OOM 2         /* device or host allocation failed */
#define FQ_ERR_BUDGET 3      /* request exceeds the configured memory budget */
#define FQ_ERR_DEVICE 4      /* HIP runtime / kernel error */
#define FQ_ERR_UNSUPPORTED 5 /* valid in the reference but not on one GPU (D > 0) */

/* amplitude types */
#define FQ_C128 0 /* complex128, as the reference (include/state.h:23-31) */
#define FQ_C64 1  /* complex64 amplitudes; cut values, phases, reductions in FP64 */

/* layer schedules for fq_run */
#define FQ_SCHED_REFERENCE 0 /* exactly src/Weighted.cpp:115-137: cost, R x RX chunk, R-1 x transposition */
#define FQ_SCHED_FUSED 1     /* cost fused into the first tile pass; high qubits by strided tiles, no transposition */
#define FQ_SCHED_Z2 2        /* FQ_SCHED_FUSED on the half state z_{N-1} = 0 (psi(b) = psi(~b) for Max-Cut
                                QAOA from |+>); RX on the top qubit pairs complementary tiles; the full
                                state is written back after the last layer.  N >= 3. */

/* kernel classes for timing */
#define FQ_K_COST 0
#define FQ_K_RX 1
#define FQ_K_SWAP 2
#define FQ_K_FUSED 3
#define FQ_K_REDUCE 4
#define FQ_K_TOPK 5
#define FQ_K_ADJOINT 6
#define FQ_K_COUNT 7

typedef struct fq_state fq_state;
typedef struct fq_graph fq_graph;

/* Message of this thread's last failed call ("" if none). */
const char *fq_last_error(void);
/* ABI version (major*100 + minor). */
int fq_version(void);

/* ---- state ------------------------------------------------------------ */

/* createMPIState / createState + prepare (src/state.cpp:31-86, 89-159,
 * 162-196).  N qubits, D log2 devices (must be 0: one GPU), B log2 amplitudes
 * of the inter-device transfer buffer (validated B <= N-D as README.md:31
 * states; no buffer is needed on one GPU), C chunk qubits of the RX tile
 * (C <= 13 for FQ_C128, C <= 14 for FQ_C64: the 2^C tile lives in LDS).
 * The amplitudes are left uninitialised: the first cost layer writes them. */
int fq_state_create(int N, int D, int B, int C, int dtype, fq_state **out);
/* destroyState (include/state.h:69; never called by the reference). */
int fq_state_destroy(fq_state *st);
/* Geometry query: any pointer may be NULL. */
int fq_state_info(const fq_state *st, int *N, int *C, int *dtype, size_t *bytes);
/* Raw device pointer of the amplitudes (for zero-copy interop). */
int fq_state_device_ptr(const fq_state *st, void **dptr);
/* Run the state's work on an external HIP stream (NULL: its own stream). */
int fq_state_set_stream(fq_state *st, void *hip_stream);
/* stateInitDeviceState (src/state.cpp:198-219): |0...0>.  Also resets the
 * qubit map to the identity. */
int fq_state_init_zero(fq_state *st);
/* initH (src/rzz.cu:26-28): the launch-control amplitude (1/sqrt 2)^N used
 * by the first cost layer.  Optional: fq_state_create computes it too. */
int fq_init_h(fq_state *st);

/* ---- graph ------------------------------------------------------------ */

/* init2DGraph (src/graph.cpp:46-52): dense N x N upper-triangular weights. */
int fq_graph_create(int N, fq_graph **out);
/* addEdgeTo2DGraph (src/graph.cpp:54-64): W[min][max] = w; a repeated edge
 * overwrites; w = 0 means no edge (skipped as src/rzz.cu:134). */
int fq_graph_add_edge(fq_graph *g, int u, int v, double w);
/* free2DGraph (src/graph.cpp:66-68). */
int fq_graph_destroy(fq_graph *g);
/* Copy the current N*N weight matrix (current labels) to W. */
int fq_graph_weights(const fq_graph *g, double *W);
/* mapping_after_swap_conti (src/swap.cu:27-50): relabel after a block swap. */
int fq_graph_relabel(fq_graph *g, int swap_out, int swap_in, int num_swap);

/* ---- gate operations (one layer's building blocks) --------------------- */

/* rotationCompressionWeighted (src/rzz.cu:173-206): psi(b) *= exp(i*gamma*
 * s(b)/2) with W in the state's physical labels; is_first != 0 is the
 * launch-control layer: psi(b) = (1/sqrt2)^N exp(i*gamma*s(b)/2), no read. */
int fq_cost_layer(fq_state *st, fq_graph *g, double gamma, int is_first);
/* multiRotateX (src/rx.cu:118-174): RX(beta) on each listed physical qubit.
 * Targets below C run in one 2^C-amplitude LDS tile pass as the reference;
 * any other distinct targets < N are also accepted (strided tiles). */
int fq_multi_rx(fq_state *st, const int *targets, int gate_size, double beta);
/* swap_gate_conti, local branch (src/swap.cu:104-105, 166-191): exchange
 * qubit blocks [swap_out, swap_out+k) and [swap_in, swap_in+k) in HBM;
 * requires swap_out + k <= swap_in and swap_in + k <= N.  Updates the
 * state's physical->logical qubit map. */
int fq_swap_conti(fq_state *st, int swap_out, int swap_in, int num_swap);

/* The driver loop (src/Weighted.cpp:97-137) for P layers with per-layer
 * angles gammas[p], betas[p].  FQ_SCHED_REFERENCE also relabels g after each
 * transposition exactly as the driver does (g ends in the state's physical
 * labels); FQ_SCHED_FUSED and FQ_SCHED_Z2 leave the layout and g unchanged.
 * With two or more passes per layer (three for FQ_SCHED_Z2) the fused
 * schedules pipeline the layers: the pass that ends layer p also applies the
 * cost phase and RX of layer p+1 on its qubits, so every layer after the
 * first runs one pass fewer (environment FQ_LAYER_PIPE=0 turns it off). */
int fq_run(fq_state *st, fq_graph *g, int P, const double *gammas, const double *betas,
           int schedule);

/* ---- observables (no reference counterpart: the driver prints timings only,
 *      src/Weighted.cpp:153-163) ---------------------------------------- */

/* <C> = sum_b |psi_b|^2 Cut(b), W in physical labels; FP64 accumulation,
 * deterministic order.  norm2 (may be NULL) receives sum |psi|^2.  Waits. */
int fq_expect_cut(fq_state *st, fq_graph *g, double *expect, double *norm2);
/* fq_expect_cut without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that must not call hipMalloc).  Its kernels keep
 * their scratch (partial sums, cut plan, result) in a fixed 4.3 MB device
 * array of the library, one per GPU, held until the process ends; states of
 * N <= 30 only (FQ_ERR_ARG beyond).  Calls that overlap on one GPU share
 * that array and race.  Waits. */
int fq_expect_cut_static(fq_state *st, fq_graph *g, double *expect, double *norm2);
/* The k <= 16 most probable basis states, best first (ties: smaller physical
 * index); idx receives LOGICAL indices (via the qubit map).  Waits. */
int fq_topk(fq_state *st, int k, uint64_t *idx, double *prob);
/* Physical -> logical qubit map: map[p] = logical qubit at physical bit p. */
int fq_qubit_map(const fq_state *st, int *map);
/* Copy the amplitudes to host memory (bytes = 2^N * element size);
 * logical != 0 reorders into logical qubit order.  Waits. */
int fq_state_copy_out(fq_state *st, void *host, size_t bytes, int logical);
/* Upload amplitudes (physical order) from host memory.  Waits. */
int fq_state_copy_in(fq_state *st, const void *host, size_t bytes);
/* Read n amplitudes as complex128 pairs out[2i], out[2i+1]; idx are logical
 * indices if logical != 0, else physical.  Waits. */
int fq_state_gather(fq_state *st, const uint64_t *idx, size_t n, double *out, int logical);
/* Copy amplitudes and qubit map of src into dst (same N and dtype), device to
 * device on dst's stream after src's queued work; work queued on src later
 * starts after the copy.  Does not wait. */
int fq_state_clone(fq_state *dst, const fq_state *src);
/* Wait for all queued work of the state. */
int fq_sync(fq_state *st);

/* ---- adjoint gradient (no reference counterpart: the reference computes no
 *      expectation and no gradient) --------------------------------------- */

/* d<C>/dgamma_k and d<C>/dbeta_k of the P-layer evaluation by reverse
 * uncompute (no checkpoints): st must hold the state fq_run(st, g, P, gammas,
 * betas, FQ_SCHED_FUSED or FQ_SCHED_Z2) produced (identity qubit map).  The state is consumed
 * (its amplitudes are unspecified afterwards: for N >= 3 the walk back runs on
 * the half z_{N-1} = 0, psi(b) = psi(~b)); a second state-sized buffer is
 * allocated on first use.  expect (may be NULL) receives <C>.  FP64
 * accumulation, fixed reduction order.  Waits. */
int fq_backward(fq_state *st, fq_graph *g, int P, const double *gammas, const double *betas,
                double *expect, double *grad_gammas, double *grad_betas);

/* ---- memory budget and packed serving (no reference counterpart) -------- */

/* Cap on the HBM held by states and slabs of this process (0: no cap).
 * Creating a state, a slab or the adjoint buffer beyond it fails with
 * FQ_ERR_BUDGET. */
int fq_set_memory_budget(size_t bytes);
int fq_memory_in_use(size_t *bytes);

/* A slab packs nslots (a power of two) evaluations of the same N into one
 * allocation: its slots form one virtual state whose top bits select the
 * slot, so each tile pass is one launch over every active slot.  Slots are
 * assigned requests (graph, P, angles) independently; fq_slab_step advances
 * every active slot by one fused layer (slots at different layers and with
 * different P run together); a slot finishing its last layer gets <C> reduced
 * and is collected with fq_slab_collect, after which it is free for the next
 * request.  Calls on one slab run in order on the slab's own HIP stream. */
typedef struct fq_slab fq_slab;
int fq_slab_create(int N, int nslots, int dtype, fq_slab **out);
/* As fq_slab_create with a layer schedule: FQ_SCHED_FUSED (what
 * fq_slab_create uses) or FQ_SCHED_Z2 (each slot holds the half state
 * z_{N-1} = 0: half the memory and traffic; N >= 3). */
int fq_slab_create_sched(int N, int nslots, int dtype, int schedule, fq_slab **out);
int fq_slab_destroy(fq_slab *sb);
int fq_slab_info(const fq_slab *sb, int *N, int *nslots, int *free_slots, int *active);
int fq_slab_assign(fq_slab *sb, int slot, fq_graph *g, int P, const double *gammas,
                   const double *betas);
/* Enqueue one layer for every active slot (waits for the previous step first);
 * nfinishing (may be NULL) receives the number of slots finishing with it. */
int fq_slab_step(fq_slab *sb, int *nfinishing);
/* Layers a step runs per slot in a slab of large slots (2^10 or more
 * first-pass tiles per slot, whose slots run fq_run's tile passes one after
 * the other): up to `layers` (>= 1, default 1), fewer for a slot that has
 * fewer left.  Slabs of small slots (one packed launch per pass) keep one
 * layer per step. */
int fq_slab_set_layers_per_step(fq_slab *sb, int layers);
/* busy = 1 while the last step is still running (non-blocking). */
int fq_slab_poll(fq_slab *sb, int *busy);
/* Finished slots: indices, <C>, sum |psi|^2; frees them.  Waits. */
int fq_slab_collect(fq_slab *sb, int *slots, double *expect, double *norm2, int *n);
/* Amplitudes of one slot (logical = physical order).  Waits. */
int fq_slab_copy_out(fq_slab *sb, int slot, void *host, size_t bytes);

/* The tile passes one layer of FQ_SCHED_FUSED or FQ_SCHED_Z2 runs for N
 * qubits (no counterpart in the reference; host-only, no device needed):
 * *npass passes of *tile_q positions each; pass p's tile positions (thread
 * bits first, the last ones register bits) in tiles[16*p .. 16*p+15] (-1
 * padded) and its RX targets as the bit mask gates[p].  In layer order pass
 * 0 applies every layer's cost phase; with pipelined layers it applies layer
 * 0's, and pass 0 and the last pass (FQ_SCHED_FUSED) or pass 1 (FQ_SCHED_Z2)
 * alternately end a layer and apply the next one's.  max_pass bounds the
 * output arrays. */
int fq_layer_passes(int N, int dtype, int schedule, int max_pass, int *npass, int *tile_q, int *tiles,
                    uint64_t *gates);

/* ---- timing (src/rx.cu:139-157 accumulates rx_gate_time_ms the same way) */

/* Enable/disable per-launch HIP event timing on the state's stream. */
int fq_timing_enable(fq_state *st, int on);
/* Accumulated device time (ms), launch count and algorithmic HBM bytes of a
 * kernel class since the last reset.  Waits for the recorded launches. */
int fq_timing_get(fq_state *st, int kind, double *ms, int64_t *count, double *bytes);
int fq_timing_reset(fq_state *st);

/* ---- device streaming rates (no counterpart in the reference: a yardstick
 * for the passes above, which all stream the state through HBM) */

/* Allocates `bytes` (>= 1 MiB, a multiple of 8; counted against the memory
 * budget), then times `reps` launches each of an in-place scale (read and
 * write every byte), a read-only sum and a fill over it, and frees it.
 * Rates in GB/s of HBM traffic (2 x bytes per in-place launch). */
int fq_stream_probe(size_t bytes, int reps, double *inplace_gbs, double *read_gbs, double *write_gbs);

#ifdef __cplusplus
}
#endif
#endif
