/*
 * fq_oracle.c -- CPU restatement of FOR-QAOA's single-device QAOA layer loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, and the timed CPU baseline).  Never
 * linked into the product library or driver.  Built with -ffp-contract=off so
 * that every product and sum is rounded exactly as written here.
 *
 * Deliberate divergence from the reference, kept in the restatement because the
 * reference's behaviour is a launch bug rather than an algorithm: for N < C the
 * reference's multiRotateX launches 2^N / 2^C = 0 blocks (src/rx.cu:129) and
 * silently applies no mixer.  Here the chunk is min(C, N), so the mixer is
 * applied; the product library does the same.
 */
#include "fq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int oq_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* src/rzz.cu:26-28 */
double oq_init_amp(int N) { return pow(1. / sqrt(2.), N); }

/* src/rzz.cu:158-171 */
void oq_flatten_upper(int N, const double *W, double *flat)
{
    int cnt = 0;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++)
            flat[cnt++] = W[i * N + j];
}

/* src/rzz.cu:121-156: per amplitude, visit every pair (i<j) in flatten order,
 * skip zero weights, add +w if the bits differ and -w if they agree, then
 * sincos(sum*angle/2).  Layer 0 writes init_amp*(cos, sin) without reading. */
void oq_cost_layer(int N, oq_cplx *sv, const double *W, double gamma, int first,
                   double init_amp)
{
    const int npairs = N * (N - 1) / 2;
    double *flat = (double *)malloc(sizeof(double) * (npairs > 0 ? npairs : 1));
    oq_flatten_upper(N, W, flat);
    const int64_t n = (int64_t)1 << N;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n; t++) {
        const uint64_t b = (uint64_t)t;
        double sum = 0;
        int cnt = 0;
        for (int i = 0; i < N; i++) {
            for (int j = i + 1; j < N; j++) {
                if (flat[cnt] != 0) {
                    if (((b >> i) & 1u) ^ ((b >> j) & 1u))
                        sum += flat[cnt];
                    else
                        sum -= flat[cnt];
                }
                cnt++;
            }
        }
        double s, c;
        sincos(sum * gamma / 2, &s, &c);
        if (first) {
            sv[t].re = c * init_amp;
            sv[t].im = s * init_amp;
        } else {
            const oq_cplx in = sv[t];
            sv[t].re = in.re * c - in.im * s;
            sv[t].im = in.re * s + in.im * c;
        }
    }
    free(flat);
}

/* src/rx.cu:24-28 bit_string: insert a zero bit at position target. */
static inline uint64_t insert_zero(uint64_t task, int target)
{
    const uint64_t mask = (1ull << target) - 1;
    return ((task >> target) << (target + 1)) | (task & mask);
}

/* src/rx.cu:32-53, 56-87, 118-126: for every 2^C chunk, apply RX(beta) to the
 * targets in list order; cos/sin of beta/2 as computed on the host at
 * src/rx.cu:121.  Butterfly arithmetic as src/rx.cu:46-49. */
int oq_multi_rx(int N, int C, oq_cplx *sv, const int *targets, int g, double beta)
{
    const int chunk = C < N ? C : N;
    for (int i = 0; i < g; i++)
        if (targets[i] < 0 || targets[i] >= chunk)
            return 1;
    double s, c;
    sincos(beta / 2, &s, &c);
    const int64_t nchunks = (int64_t)1 << (N - chunk);
    const int64_t half = (int64_t)1 << (chunk - 1);
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nchunks; ch++) {
        oq_cplx *base = sv + (ch << chunk);
        for (int gi = 0; gi < g; gi++) {
            const int target = targets[gi];
            for (int64_t u = 0; u < half; u++) {
                const uint64_t up = insert_zero((uint64_t)u, target);
                const uint64_t lo = up + (1ull << target);
                const oq_cplx q0 = base[up];
                const oq_cplx q1 = base[lo];
                oq_cplx o0, o1;
                o0.re = q0.re * c + q1.im * s;
                o0.im = q0.im * c - q1.re * s;
                o1.re = q0.im * s + q1.re * c;
                o1.im = -q0.re * s + q1.im * c;
                base[up] = o0;
                base[lo] = o1;
            }
        }
    }
    return 0;
}

/* src/swap.cu:65-73 bit_string_multi: insert `amount` zero bits at targ. */
static inline uint64_t insert_zeros(uint64_t task, int targ, int amount)
{
    const uint64_t mask = (1ull << targ) - 1;
    return ((task >> targ) << (targ + amount)) | (task & mask);
}

/* src/swap.cu:76-102 with the (s1<s2) pair enumeration of src/swap.cu:180-184:
 * for every rest index and pair, swap psi[s1<<in | s2<<out | g] with
 * psi[s1<<out | s2<<in | g]. */
int oq_swap_conti(int N, oq_cplx *sv, int out, int in, int k)
{
    if (k <= 0)
        return 0;
    if (out < 0 || out + k > in || in + k > N)
        return 1;
    const int64_t nrest = (int64_t)1 << (N - 2 * k);
    const int64_t ns = (int64_t)1 << k;
#pragma omp parallel for collapse(2) schedule(dynamic, 64)
    for (int64_t r = 0; r < nrest; r++) {
        for (int64_t s1 = 0; s1 < ns; s1++) {
            uint64_t gidx = insert_zeros((uint64_t)r, out, k);
            gidx = insert_zeros(gidx, in, k);
            for (int64_t s2 = s1 + 1; s2 < ns; s2++) {
                const uint64_t up = ((uint64_t)s1 << in) + ((uint64_t)s2 << out) + gidx;
                const uint64_t lo = ((uint64_t)s1 << out) + ((uint64_t)s2 << in) + gidx;
                const oq_cplx tmp = sv[up];
                sv[up] = sv[lo];
                sv[lo] = tmp;
            }
        }
    }
    return 0;
}

/* src/swap.cu:27-50: for each t<k, exchange labels out+t and in+t in the
 * upper-triangular storage W[min][max]; the (out+t, in+t) entry itself stays. */
int oq_relabel(int N, double *W, int out, int in, int k)
{
    if (!(in > out) || in + k > N || out + k > in)
        return 1;
    for (int t = 0; t < k; t++) {
        const int a = out + t, b = in + t;
        for (int i = 0; i < N; i++) {
            if (i == a || i == b)
                continue;
            double *pa = (i < a) ? &W[i * N + a] : &W[a * N + i];
            double *pb = (i < b) ? &W[i * N + b] : &W[b * N + i];
            const double tmp = *pa;
            *pa = *pb;
            *pb = tmp;
        }
    }
    return 0;
}

/* src/Weighted.cpp:93-137: subcircuits [C, C, ..., N mod C]; per layer one cost
 * phase, then for each subcircuit RX on physical qubits 0..|sub_i|-1 followed
 * (except after the last) by the block swap [0,k) <-> [s, s+k) and the relabel. */
int oq_run_weighted(int N, int C, int P, double *W, const double *gammas,
                    const double *betas, oq_cplx *sv, int *qmap)
{
    return oq_run_weighted_timed(N, C, P, W, gammas, betas, sv, qmap, NULL);
}

/* MEASURET_START / MEASURET_END (include/utils.h:29-37, src/Weighted.cpp:113,
 * 140) bracket the P-layer loop only; a monotonic clock here. */
static double oq_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int oq_run_weighted_timed(int N, int C, int P, double *W, const double *gammas,
                          const double *betas, oq_cplx *sv, int *qmap, double *seconds)
{
    if (N < 1 || C < 1 || P < 0)
        return 1;
    const double init_amp = oq_init_amp(N);
    int sizes[64];
    int R = 0;
    for (int i = 0; i < N; i += C)
        sizes[R++] = (i + C <= N) ? C : N - i;
    int targets[64];
    for (int i = 0; i < 64; i++)
        targets[i] = i;
    for (int p = 0; p < N; p++)
        qmap[p] = p;
    const double t0 = oq_now();
    for (int p = 0; p < P; p++) {
        oq_cost_layer(N, sv, W, gammas[p], p == 0, init_amp);
        int swap_in = 0;
        for (int i = 0; i < R; i++) {
            if (oq_multi_rx(N, C, sv, targets, sizes[i], betas[p]))
                return 2;
            if (i < R - 1) {
                swap_in += sizes[i];
                const int k = sizes[i + 1];
                if (oq_swap_conti(N, sv, 0, swap_in, k))
                    return 3;
                if (oq_relabel(N, W, 0, swap_in, k))
                    return 4;
                for (int t = 0; t < k; t++) {
                    const int tmp = qmap[t];
                    qmap[t] = qmap[swap_in + t];
                    qmap[swap_in + t] = tmp;
                }
            }
        }
    }
    if (seconds)
        *seconds = oq_now() - t0;
    return 0;
}

double oq_cut_value(int N, const double *W, uint64_t b)
{
    double cut = 0;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++)
            if (W[i * N + j] != 0 && (((b >> i) ^ (b >> j)) & 1u))
                cut += W[i * N + j];
    return cut;
}

/* Fixed-size partial sums combined in index order, so the result does not
 * depend on the thread count. */
#define OQ_RED_CHUNK 4096

double oq_expect_cut(int N, const oq_cplx *sv, const double *W)
{
    int E = 0;
    int eu[2080], ev[2080];
    double ew[2080];
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++)
            if (W[i * N + j] != 0) {
                eu[E] = i;
                ev[E] = j;
                ew[E] = W[i * N + j];
                E++;
            }
    const int64_t n = (int64_t)1 << N;
    const int64_t nch = (n + OQ_RED_CHUNK - 1) / OQ_RED_CHUNK;
    double *part = (double *)calloc((size_t)nch, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nch; ch++) {
        double acc = 0;
        const int64_t lo = ch * OQ_RED_CHUNK;
        const int64_t hi = lo + OQ_RED_CHUNK < n ? lo + OQ_RED_CHUNK : n;
        for (int64_t b = lo; b < hi; b++) {
            double cut = 0;
            for (int e = 0; e < E; e++)
                if ((((uint64_t)b >> eu[e]) ^ ((uint64_t)b >> ev[e])) & 1u)
                    cut += ew[e];
            acc += (sv[b].re * sv[b].re + sv[b].im * sv[b].im) * cut;
        }
        part[ch] = acc;
    }
    double tot = 0;
    for (int64_t ch = 0; ch < nch; ch++)
        tot += part[ch];
    free(part);
    return tot;
}

double oq_norm2(int N, const oq_cplx *sv)
{
    const int64_t n = (int64_t)1 << N;
    const int64_t nch = (n + OQ_RED_CHUNK - 1) / OQ_RED_CHUNK;
    double *part = (double *)calloc((size_t)nch, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nch; ch++) {
        double acc = 0;
        const int64_t lo = ch * OQ_RED_CHUNK;
        const int64_t hi = lo + OQ_RED_CHUNK < n ? lo + OQ_RED_CHUNK : n;
        for (int64_t b = lo; b < hi; b++)
            acc += sv[b].re * sv[b].re + sv[b].im * sv[b].im;
        part[ch] = acc;
    }
    double tot = 0;
    for (int64_t ch = 0; ch < nch; ch++)
        tot += part[ch];
    free(part);
    return tot;
}

/* Keep (prob, idx) lists ordered best-first: larger prob, then smaller idx. */
static inline int oq_better(double pa, uint64_t ia, double pb, uint64_t ib)
{
    return pa > pb || (pa == pb && ia < ib);
}

static void oq_insert(int k, int *cnt, uint64_t *idx, double *prob, uint64_t i, double p)
{
    if (*cnt == k && !oq_better(p, i, prob[k - 1], idx[k - 1]))
        return;
    int pos = (*cnt < k) ? (*cnt)++ : k - 1;
    while (pos > 0 && oq_better(p, i, prob[pos - 1], idx[pos - 1])) {
        prob[pos] = prob[pos - 1];
        idx[pos] = idx[pos - 1];
        pos--;
    }
    prob[pos] = p;
    idx[pos] = i;
}

void oq_topk(int N, const oq_cplx *sv, int k, uint64_t *idx, double *prob)
{
    const int64_t n = (int64_t)1 << N;
    if (k > n)
        k = (int)n;
    const int64_t nch = (n + 65535) / 65536;
    uint64_t *cidx = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(nch * k));
    double *cprob = (double *)malloc(sizeof(double) * (size_t)(nch * k));
    int *ccnt = (int *)calloc((size_t)nch, sizeof(int));
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nch; ch++) {
        const int64_t lo = ch * 65536;
        const int64_t hi = lo + 65536 < n ? lo + 65536 : n;
        for (int64_t b = lo; b < hi; b++) {
            const double p = sv[b].re * sv[b].re + sv[b].im * sv[b].im;
            oq_insert(k, &ccnt[ch], cidx + ch * k, cprob + ch * k, (uint64_t)b, p);
        }
    }
    int cnt = 0;
    for (int64_t ch = 0; ch < nch; ch++)
        for (int j = 0; j < ccnt[ch]; j++)
            oq_insert(k, &cnt, idx, prob, cidx[ch * k + j], cprob[ch * k + j]);
    free(cidx);
    free(cprob);
    free(ccnt);
}

void oq_to_logical(int N, const oq_cplx *sv, const int *qmap, oq_cplx *out)
{
    const int64_t n = (int64_t)1 << N;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < n; b++) {
        uint64_t L = 0;
        for (int p = 0; p < N; p++)
            L |= (((uint64_t)b >> p) & 1u) << qmap[p];
        out[L] = sv[b];
    }
}
