/*
 * fq_oracle.h -- CPU restatement of FOR-QAOA's single-device QAOA Max-Cut layer
 * loop, in complex128 (interleaved {re, im} doubles, little-endian qubit order).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against, and the "reference CPU path" timed beside the GPU in bench.py.  It is
 * never linked into, loaded by, or called from the product library
 * (for-qaoa_amd/csrc) or the `weighted` driver.  Only tests/, the smoke check in
 * __graft_entry__.py and bench.py's cpu_baseline leg may use it.
 *
 * Every function names the reference lines (paths relative to the FOR-QAOA
 * repository root) whose behaviour it restates.  How each function is pinned
 * (dense textbook-gate simulator, Hamming-weight closed form, analytic P=1
 * expectation, brute-force relabel, bitwise permutation checks) is listed in
 * tests/test_oracle_pins.py.
 */
#ifndef FQ_ORACLE_H
#define FQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double re, im;
} oq_cplx;

/* initH (src/rzz.cu:26-28): initialAmp = pow(1/sqrt(2), N). */
double oq_init_amp(int N);

/* UpperTri_to_1D (src/rzz.cu:158-171): flatten W[i*N+j], i<j, i-major. */
void oq_flatten_upper(int N, const double *W, double *flat);

/* _rotationCompressionWeighted_conti (src/rzz.cu:121-156) over the whole
 * state (D=0, b_offset=0).  first!=0 is the launch-control branch
 * (src/rzz.cu:148-150): the state is written, not read. */
void oq_cost_layer(int N, oq_cplx *sv, const double *W, double gamma, int first,
                   double init_amp);

/* multiRotateX + _multirotateX_unroll + rx_gate_unroll (src/rx.cu:24-87,
 * 118-126): RX(beta) on targets[0..g) inside every 2^C-amplitude chunk.
 * Requires every target < min(C, N).  chunk = min(C, N) (see the note on
 * N < C in the .c file). */
int oq_multi_rx(int N, int C, oq_cplx *sv, const int *targets, int g, double beta);

/* _swap_gate_conti (src/swap.cu:76-102) with the pair table of
 * swap_gate_conti (src/swap.cu:173-186): exchange qubit blocks
 * [out, out+k) and [in, in+k).  Requires out+k <= in and in+k <= N. */
int oq_swap_conti(int N, oq_cplx *sv, int out, int in, int k);

/* mapping_after_swap_conti (src/swap.cu:27-50): relabel the upper-triangular
 * weight matrix W (N*N, W[min*N+max]) after oq_swap_conti. */
int oq_relabel(int N, double *W, int out, int in, int k);

/* The driver's layer loop (src/Weighted.cpp:93-137) generalised to per-layer
 * angles.  W is relabelled in place exactly as the driver does; qmap[phys]
 * receives the logical qubit held at each physical position at the end. */
int oq_run_weighted(int N, int C, int P, double *W, const double *gammas,
                    const double *betas, oq_cplx *sv, int *qmap);

/* oq_run_weighted, and *seconds receives the monotonic-clock time of the
 * P-layer loop alone, where MEASURET_START / MEASURET_END sit in the driver
 * (src/Weighted.cpp:113,140). */
int oq_run_weighted_timed(int N, int C, int P, double *W, const double *gammas,
                          const double *betas, oq_cplx *sv, int *qmap, double *seconds);

/* Cut(z) = sum_{i<j} W[i][j] * [z_i != z_j] for one basis index. */
double oq_cut_value(int N, const double *W, uint64_t b);

/* <C> = sum_b |psi_b|^2 Cut(b) with W in the state's (physical) labels. */
double oq_expect_cut(int N, const oq_cplx *sv, const double *W);

/* sum_b |psi_b|^2 */
double oq_norm2(int N, const oq_cplx *sv);

/* Largest k probabilities (ties: smaller index first); indices physical. */
void oq_topk(int N, const oq_cplx *sv, int k, uint64_t *idx, double *prob);

/* Physical -> logical reordering: out[L(b)] = sv[b] with bit p of b moved to
 * bit qmap[p]. */
void oq_to_logical(int N, const oq_cplx *sv, const int *qmap, oq_cplx *out);

/* Number of OpenMP threads the oracle uses (1 without OpenMP). */
int oq_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
