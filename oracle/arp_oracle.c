/*
 * CPU oracle for the delta-encoded Ising annealing detector's kernel path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: tests, the smoke check and the CPU-baseline leg of
 * the benchmark call it; the product path (arpeggio_amd) never links or calls it.
 *
 * Plain C restatement of the reference's integer semantics (paths relative to the reference repository):
 *   - Philox4x32-10 (Random123): arpeggio/mimo/chordfx.py:39-56, arpeggio/sim/local_code_a59/arpeggio_ref59.py:14-21
 *   - balanced start pattern:   arpeggio_ref59.py:31-38
 *   - Arpeggio first-accept ('arpeggio') and parallel ('par') passes: arpeggio_ref59.py:41-85
 *     (GPU form with identical semantics: arpeggio/sim/arpeggio_fx.cu:35-100; host API arpeggio/sim/arpeggiofx.py:23-44)
 *   - Chord snapshot commit with copy phases, Onsager c_k and pinning q: arpeggio/mimo/chordfx.py:80-118
 *   - the production engine's lane-shared Chord (one field per coordinate, P copies, every accepted lane of the active
 *     copy committed from one snapshot) with its xorshift32 lane streams or Philox: arpeggio/sim/x3model/chord_ref.py:
 *     17-45 (fmix32, xs32, lane_seeds) and 73-123 (chord_ref); pinned against the reference's own cycle model
 *     (arpeggio/sim/x3model/engine_model.cpp, built into oracle/_ref by this directory's Makefile)
 *   - FP32/FP64 float twins of both anneals with Philox uniforms r = (U - 1/2) 4T (the reference's float solvers,
 *     arpeggio/sim/arpx55.py:29-99 rule 'r' and arpeggio/mimo/mimo.py:204-249, draw from torch generators; this
 *     oracle feeds them the fixed-point path's Philox words instead).
 * Every anneal is independent; the batch entry points spread (frame, anneal) pairs over nthreads POSIX threads.
 *
 * Pins (tests/test_oracle_*.py): Random123 known-answer vectors; golden vectors written by importing the reference's
 * own pure-Python/PyTorch functions in the build container (tests/golden/make_golden.py); B200 kernel output hashes
 * from the reference's data/ber/equivalence/kernel_b200_v2.json; brute-force energy bookkeeping on tiny problems; and
 * the float64 twin reproducing the integer trajectories exactly on dyadic inputs.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MASK32 0xFFFFFFFFu

/* ------------------------------------------------------------------ Philox4x32-10 (Salmon et al., SC'11) */
static inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int i = 0; i < 10; i++) {
        if (i) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t a = (uint64_t)0xD2511F53u * c[0];
        uint64_t b = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t x0 = (uint32_t)(b >> 32) ^ c[1] ^ k0;
        uint32_t x2 = (uint32_t)(a >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)b;
        c[3] = (uint32_t)a;
        c[0] = x0;
        c[2] = x2;
    }
}

static inline uint32_t philox_word(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, int w)
{
    uint32_t c[4] = {c0, c1, c2, c3};
    philox10(c, k0, k1);
    return c[w & 3];
}

/* in: n rows of (counter[4], key[2]); out: n rows of 4 words. */
int orc_philox_kat(const uint32_t *in, uint32_t *out, int n)
{
    for (int t = 0; t < n; t++) {
        uint32_t c[4] = {in[6 * t], in[6 * t + 1], in[6 * t + 2], in[6 * t + 3]};
        philox10(c, in[6 * t + 4], in[6 * t + 5]);
        memcpy(out + 4 * t, c, sizeof c);
    }
    return 0;
}

/* The i-th P-bit word with P/2 ones, in increasing numeric order (arpeggio_ref59.py:31-38). -1 if i is out of range. */
int orc_bal_pat(int P, uint32_t i)
{
    uint32_t n = 0;
    for (int v = 0; v < (1 << P); v++) {
        if (__builtin_popcount((unsigned)v) == P / 2) {
            if (n == i) return v;
            n++;
        }
    }
    return -1;
}

static int bal_count(int P) { return P == 2 ? 2 : (P == 4 ? 6 : 20); }

/* ------------------------------------------------------------------ thread pool over independent items */
typedef void (*item_fn)(void *ctx, long item);
typedef struct {
    item_fn fn;
    void *ctx;
    long n;
    long next;
    pthread_mutex_t mu;
} pool_t;

static void *pool_worker(void *arg)
{
    pool_t *p = (pool_t *)arg;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        long it = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (it >= p->n) break;
        p->fn(p->ctx, it);
    }
    return NULL;
}

static void run_pool(item_fn fn, void *ctx, long n, int nthreads)
{
    if (nthreads <= 1 || n <= 1) {
        for (long i = 0; i < n; i++) fn(ctx, i);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    pool_t p = {fn, ctx, n, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t th[256];
    int made = 0;
    for (int i = 0; i < nthreads; i++)
        if (pthread_create(&th[made], NULL, pool_worker, &p) == 0) made++;
    pool_worker(&p);
    for (int i = 0; i < made; i++) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------ Arpeggio fixed point (arpeggio_ref59.anneal) */
/*
 * One anneal on the lane-shared form: G (D x D, row-major, G[i*D + j] = G_ij), h (D), P copies, T (L passes), q.
 * Spins s[p*D + j] (copy-major). w_i = h_i - 2 sum_j G_ij m_j with m_j = sum_p s_pj. Pass k: p = k mod P, n = k div P,
 * A = 4 T_k, B_k = q + 2 T_k; lane j draws x = Philox(n >> 2, p D + j, r, inst)[n & 3] and flips iff
 * x A > (s_pj w_j + 2 G_jj + B_k) 2^32. 'arpeggio' (mode 0) scans lanes in index order and lets later lanes see
 * the updated fields; 'par' (mode 1) decides every lane of the pass on one snapshot, then applies the flips in index
 * order. A flip adds 4 s w_j + 8 G_jj to de and 4 s G_ij to every w_i. tests: 1 per pass plus 1 per flip not on
 * the last lane ('arpeggio'), 1 per pass ('par'). After each pass the state is kept if de < best (first minimum).
 * acc_k/acc_j (optional) receive the accepted (pass, lane) sequence; tests_k (optional, L) the tests per pass.
 */
typedef struct {
    int64_t best, flips, tests, n1, init, n_acc;
} arp_result;

/* Fields and start energy of a copy-major state: w_i = h_i - 2 sum_j G_ij m_j; returns -(h + w).m, which is
 * 2 E0 + 2 P tr(G) of the dDI expansion (2 E0 = -2 h.m + 2 m^T G m - 2 P tr G, and m^T G m = (h - w).m / 2). */
static int64_t fields_of(const int32_t *G, const int32_t *h, int D, int P, const int8_t *s, int64_t *m, int64_t *w)
{
    for (int j = 0; j < D; j++) {
        int64_t v = 0;
        for (int p = 0; p < P; p++) v += s[p * D + j];
        m[j] = v;
    }
    int64_t e = 0;
    for (int i = 0; i < D; i++) {
        int64_t acc = 0;
        for (int j = 0; j < D; j++) acc += (int64_t)G[(size_t)i * D + j] * m[j];
        w[i] = (int64_t)h[i] - 2 * acc;
        e -= ((int64_t)h[i] + w[i]) * m[i];
    }
    return e;
}

/* thr (P D, optional): learned start, spin +1 iff Philox(2^32 - 1, p D + j, r, inst)[0] >= 2^32 - thr_pj (thr = 2^31
 * is the random start). gate: take the learned start only if its energy is not above the random start's.
 * used (optional): 1 if the learned start was taken. */
static void arpeggio_one_ws(const int32_t *G, const int32_t *h, int D, int P, const int32_t *T, int L, int q,
                            uint32_t r, uint32_t inst, uint32_t k0, uint32_t k1, int start, int mode, int8_t *bs,
                            arp_result *res, int32_t *acc_k, int32_t *acc_j, long max_acc, int32_t *tests_k,
                            const uint32_t *thr, int gate, int8_t *used);

static void arpeggio_one(const int32_t *G, const int32_t *h, int D, int P, const int32_t *T, int L, int q, uint32_t r,
                         uint32_t inst, uint32_t k0, uint32_t k1, int start, int mode, int8_t *bs, arp_result *res,
                         int32_t *acc_k, int32_t *acc_j, long max_acc, int32_t *tests_k)
{
    arpeggio_one_ws(G, h, D, P, T, L, q, r, inst, k0, k1, start, mode, bs, res, acc_k, acc_j, max_acc, tests_k, NULL, 0,
                    NULL);
}

static void arpeggio_one_ws(const int32_t *G, const int32_t *h, int D, int P, const int32_t *T, int L, int q,
                            uint32_t r, uint32_t inst, uint32_t k0, uint32_t k1, int start, int mode, int8_t *bs,
                            arp_result *res, int32_t *acc_k, int32_t *acc_j, long max_acc, int32_t *tests_k,
                            const uint32_t *thr, int gate, int8_t *used)
{
    int8_t *s = (int8_t *)malloc((size_t)P * D);
    int64_t *w = (int64_t *)malloc(sizeof(int64_t) * D);
    int64_t *m = (int64_t *)malloc(sizeof(int64_t) * D);
    uint8_t *fl = (uint8_t *)malloc((size_t)D);
    if (thr) {
        int8_t *sr = (int8_t *)malloc((size_t)P * D);
        for (int p = 0; p < P; p++)
            for (int j = 0; j < D; j++) {
                uint32_t x = philox_word(MASK32, (uint32_t)(p * D + j), r, inst, k0, k1, 0);
                s[p * D + j] = (uint64_t)x + thr[p * D + j] >= 0x100000000ull ? 1 : -1;
                sr[p * D + j] = x >= 0x80000000u ? 1 : -1;
            }
        int take = 1;
        if (gate) {
            int64_t el = fields_of(G, h, D, P, s, m, w), er = fields_of(G, h, D, P, sr, m, w);
            take = el <= er;
        }
        if (!take) memcpy(s, sr, (size_t)P * D);
        if (used) *used = (int8_t)take;
        free(sr);
    } else if (start == 0) {
        for (int p = 0; p < P; p++)
            for (int j = 0; j < D; j++)
                s[p * D + j] = philox_word(MASK32, (uint32_t)(p * D + j), r, inst, k0, k1, 0) >= 0x80000000u ? 1 : -1;
    } else {
        uint64_t C = (uint64_t)bal_count(P);
        for (int j = 0; j < D; j++) {
            uint32_t x = philox_word(MASK32, (uint32_t)j, r, inst, k0, k1, 0);
            int pat = orc_bal_pat(P, (uint32_t)(((uint64_t)x * C) >> 32));
            for (int p = 0; p < P; p++) s[p * D + j] = ((pat >> p) & 1) ? 1 : -1;
        }
    }
    int64_t init = 0;
    for (int j = 0; j < D; j++) {
        int64_t v = 0;
        for (int p = 0; p < P; p++) v += s[p * D + j];
        m[j] = v;
        init += v != 0;
    }
    for (int i = 0; i < D; i++) {
        int64_t acc = 0;
        for (int j = 0; j < D; j++) acc += (int64_t)G[(size_t)i * D + j] * m[j];
        w[i] = (int64_t)h[i] - 2 * acc;
    }
    int64_t de = 0, best = (int64_t)1 << 62, flips = 0, tests = 0, n1 = 0, nacc = 0;
    for (int k = 0; k < L; k++) {
        int p = k % P;
        uint32_t n = (uint32_t)(k / P);
        int64_t A = 4 * (int64_t)T[k], Bk = (int64_t)q + 2 * (int64_t)T[k];
        int8_t *sp = s + (size_t)p * D;
        int64_t t = 1;
        if (mode == 0) {
            for (int j = 0; j < D; j++) {
                uint32_t x = philox_word(n >> 2, (uint32_t)(p * D + j), r, inst, k0, k1, (int)(n & 3));
                int64_t dj = G[(size_t)j * D + j];
                /* (s w + 2 G_jj + B) << 32 as Python ints: multiply, so negative values stay exact */
                if ((int64_t)x * A > (sp[j] * w[j] + 2 * dj + Bk) * (int64_t)4294967296LL) {
                    int64_t so = sp[j];
                    de += 4 * so * w[j] + 8 * dj;
                    for (int i = 0; i < D; i++) w[i] += 4 * so * (int64_t)G[(size_t)i * D + j];
                    sp[j] = (int8_t)(-so);
                    flips++;
                    if (acc_k && nacc < max_acc) {
                        acc_k[nacc] = k;
                        acc_j[nacc] = j;
                    }
                    nacc++;
                    if (j < D - 1) t++;
                }
            }
        } else {
            int any = 0;
            for (int j = 0; j < D; j++) {
                uint32_t x = philox_word(n >> 2, (uint32_t)(p * D + j), r, inst, k0, k1, (int)(n & 3));
                int64_t dj = G[(size_t)j * D + j];
                fl[j] = (int64_t)x * A > (sp[j] * w[j] + 2 * dj + Bk) * (int64_t)4294967296LL;
                any |= fl[j];
            }
            n1 += any;
            for (int j = 0; j < D; j++) {
                if (!fl[j]) continue;
                int64_t so = sp[j];
                int64_t dj = G[(size_t)j * D + j];
                de += 4 * so * w[j] + 8 * dj;
                for (int i = 0; i < D; i++) w[i] += 4 * so * (int64_t)G[(size_t)i * D + j];
                sp[j] = (int8_t)(-so);
                flips++;
                if (acc_k && nacc < max_acc) {
                    acc_k[nacc] = k;
                    acc_j[nacc] = j;
                }
                nacc++;
            }
            t = 1;
        }
        if (tests_k) tests_k[k] = (int32_t)t;
        tests += t;
        if (de < best) {
            best = de;
            memcpy(bs, s, (size_t)P * D);
        }
    }
    res->best = best;
    res->flips = flips;
    res->tests = tests;
    res->n1 = n1;
    res->init = init;
    res->n_acc = nacc;
    free(s);
    free(w);
    free(m);
    free(fl);
}

/* Single anneal with its full event record (the reference returns acc, tests, best, bs, init, n1).
 * out_best[6] = best, flips, tests, n1, init, n_acc. */
int orc_arpeggio_trace(const int32_t *G, const int32_t *h, int D, int P, const int32_t *T, int L, int q, uint32_t r,
                       uint32_t inst, uint64_t key, int start, int mode, int8_t *bs, int64_t *out_best,
                       int32_t *acc_k, int32_t *acc_j, long max_acc, int32_t *tests_k)
{
    if (D < 1 || L < 1 || (P != 2 && P != 4 && P != 6)) return -1;
    arp_result res;
    arpeggio_one(G, h, D, P, T, L, q, r, inst, (uint32_t)(key & MASK32), (uint32_t)(key >> 32), start, mode, bs, &res,
                 acc_k, acc_j, max_acc, tests_k);
    out_best[0] = res.best;
    out_best[1] = res.flips;
    out_best[2] = res.tests;
    out_best[3] = res.n1;
    out_best[4] = res.init;
    out_best[5] = res.n_acc;
    return 0;
}

typedef struct {
    const int32_t *G, *h, *T;
    const uint32_t *inst;
    int D, P, R, L, q, start, mode;
    uint32_t k0, k1;
    int8_t *out_s;
    int64_t *out_de, *out_st;
} arp_batch_ctx;

static void arp_batch_item(void *vctx, long t)
{
    arp_batch_ctx *c = (arp_batch_ctx *)vctx;
    long b = t / c->R;
    uint32_t r = (uint32_t)(t % c->R);
    arp_result res;
    arpeggio_one(c->G + (size_t)b * c->D * c->D, c->h + (size_t)b * c->D, c->D, c->P, c->T, c->L, c->q, r, c->inst[b],
                 c->k0, c->k1, c->start, c->mode, c->out_s + (size_t)t * c->P * c->D, &res, NULL, NULL, 0, NULL);
    c->out_de[t] = res.best;
    c->out_st[2 * t] = res.flips;
    c->out_st[2 * t + 1] = res.tests;
}

typedef struct {
    const int32_t *G, *h, *T;
    const uint32_t *inst, *thr;
    int D, P, R, L, q, gate;
    uint32_t k0, k1;
    int8_t *out_s, *used;
    int64_t *out_de, *out_st;
} arp_ws_ctx;

static void arp_ws_item(void *vctx, long t)
{
    arp_ws_ctx *c = (arp_ws_ctx *)vctx;
    long b = t / c->R;
    uint32_t r = (uint32_t)(t % c->R);
    arp_result res;
    arpeggio_one_ws(c->G + (size_t)b * c->D * c->D, c->h + (size_t)b * c->D, c->D, c->P, c->T, c->L, c->q, r,
                    c->inst[b], c->k0, c->k1, 0, 0, c->out_s + (size_t)t * c->P * c->D, &res, NULL, NULL, 0, NULL,
                    c->thr + (size_t)b * c->P * c->D, c->gate, c->used ? c->used + t : NULL);
    c->out_de[t] = res.best;
    c->out_st[2 * t] = res.flips;
    c->out_st[2 * t + 1] = res.tests;
}

/* First-accept anneal from a learned start: thr (B, P D) uint32 per-spin thresholds; gate 1 = energy-gated against
 * the random start. used (B, R) int8 (optional): 1 where the learned start was taken. */
int orc_arpeggio_ws(const int32_t *G, const int32_t *h, int B, int D, int P, int R, const int32_t *T, int L, int q,
                    const uint32_t *inst, uint64_t key, const uint32_t *thr, int gate, int8_t *out_s, int64_t *out_de,
                    int64_t *out_st, int8_t *used, int nthreads)
{
    if (D < 1 || L < 1 || R < 1 || B < 0 || (P != 2 && P != 4 && P != 6) || !thr) return -1;
    arp_ws_ctx c = {G, h, T, inst, thr, D, P, R, L, q, gate, (uint32_t)(key & MASK32), (uint32_t)(key >> 32), out_s,
                    used, out_de, out_st};
    run_pool(arp_ws_item, &c, (long)B * R, nthreads);
    return 0;
}

/* arpeggiofx.solve semantics (arpeggiofx.py:23-44): G (B, D, D) row-major, h (B, D), inst (B) uint32, key 64-bit
 * (k0 = low word, k1 = high word), start 0 'rand' / 1 'bal'. out_s (B, R, P D) copy-major, out_de (B, R),
 * out_st (B, R, 2) = (flips, tests). */
int orc_arpeggio_fx(const int32_t *G, const int32_t *h, int B, int D, int P, int R, const int32_t *T, int L, int q,
                    const uint32_t *inst, uint64_t key, int start, int mode, int8_t *out_s, int64_t *out_de,
                    int64_t *out_st, int nthreads)
{
    if (D < 1 || L < 1 || R < 1 || B < 0 || (P != 2 && P != 4 && P != 6)) return -1;
    arp_batch_ctx c = {G, h, T, inst, D, P, R, L, q, start, mode, (uint32_t)(key & MASK32), (uint32_t)(key >> 32),
                       out_s, out_de, out_st};
    run_pool(arp_batch_item, &c, (long)B * R, nthreads);
    return 0;
}

/* ------------------------------------------------------------------ Chord fixed point (chordfx.chord_fx) */
/*
 * J (B, N, N) int16 symmetric, h, h0 (B, N) int32, T, c (S) int32, q int. Start: spin i = +1 iff
 * Philox(2^32 - 1, i, r, inst)[0] >= 2^31, inst = inst0 + b. Step k: p = k mod P, n = k div P,
 * x = Philox(n div 4, i, r, inst)[n mod 4]; z = s u + q - c_k (prevflip ? -1 : 1); flip iff x 4T_k > (z + 2T_k) 2^32
 * and i is in copy p (spins [p N/P, (p+1) N/P)). All accepted flips of the step commit together (one snapshot).
 * u = h + 2 J s (kept exactly by integer updates), e = sum s (h - 2 h0 - u) after every step; the state is kept if
 * e < best (first minimum). mode 0: per-anneal (B, R, N) spins and (B, R) energies; mode 1: the lowest-energy anneal
 * per frame (lowest r on ties), (B, N) and (B,). stats (optional, 4, the counters of chord_kernel.cu:6-8): flips,
 * energy evaluations (steps with a flip, and k = 0), Philox blocks drawn in steps (one per spin of the active copy
 * every 4th update), steps x anneals.
 */
typedef struct {
    const int16_t *J;
    const int32_t *h, *h0, *T, *c;
    int N, S, q, P, R;
    uint32_t k0, k1;
    int64_t inst0;
    int8_t *s_out;
    int64_t *e_out;
    int64_t *stat; /* per item, 4 */
} chord_ctx;

static void chord_item(void *vctx, long t)
{
    chord_ctx *cx = (chord_ctx *)vctx;
    long b = t / cx->R;
    uint32_t r = (uint32_t)(t % cx->R);
    int N = cx->N, P = cx->P;
    const int16_t *J = cx->J + (size_t)b * N * N;
    const int32_t *h = cx->h + (size_t)b * N, *h0 = cx->h0 + (size_t)b * N;
    uint32_t inst = (uint32_t)(cx->inst0 + b);
    int64_t *u = (int64_t *)malloc(sizeof(int64_t) * N);
    int8_t *s = (int8_t *)malloc((size_t)N);
    int8_t *best_s = cx->s_out + (size_t)t * N;
    uint8_t *pf = (uint8_t *)calloc((size_t)N, 1);
    uint8_t *fl = (uint8_t *)malloc((size_t)N);
    int64_t nfl = 0, nev = 0, nrng = 0;
    for (int i = 0; i < N; i++) s[i] = philox_word(MASK32, (uint32_t)i, r, inst, cx->k0, cx->k1, 0) >= 0x80000000u ? 1 : -1;
    for (int i = 0; i < N; i++) {
        int64_t acc = 0;
        for (int j = 0; j < N; j++) acc += (int64_t)J[(size_t)i * N + j] * s[j];
        u[i] = (int64_t)h[i] + 2 * acc;
    }
    int64_t best_e = (int64_t)1 << 62;
    memcpy(best_s, s, (size_t)N);
    int per = N / P;
    for (int k = 0; k < cx->S; k++) {
        int p = k % P;
        int n = k / P;
        int64_t T = cx->T[k], c = cx->c[k];
        int any = 0;
        for (int i = 0; i < N; i++) {
            fl[i] = 0;
            if (i / per != p) continue;
            uint32_t x = philox_word((uint32_t)(n / 4), (uint32_t)i, r, inst, cx->k0, cx->k1, n % 4);
            if (n % 4 == 0) nrng++;
            int64_t z = s[i] * u[i] + cx->q - c * (pf[i] ? -1 : 1);
            fl[i] = (int64_t)x * (4 * T) > (z + 2 * T) * (int64_t)4294967296LL;
            any |= fl[i];
        }
        for (int j = 0; j < N; j++) {
            pf[j] = fl[j];
            if (!fl[j]) continue;
            int64_t d = -2 * (int64_t)s[j];
            for (int i = 0; i < N; i++) u[i] += 2 * d * (int64_t)J[(size_t)i * N + j];
            s[j] = (int8_t)(-s[j]);
            nfl++;
        }
        if (any || k == 0) nev++;
        int64_t e = 0;
        for (int i = 0; i < N; i++) e += s[i] * ((int64_t)h[i] - 2 * (int64_t)h0[i] - u[i]);
        if (e < best_e) {
            best_e = e;
            memcpy(best_s, s, (size_t)N);
        }
    }
    cx->e_out[t] = best_e;
    if (cx->stat) {
        cx->stat[4 * t] = nfl;
        cx->stat[4 * t + 1] = nev;
        cx->stat[4 * t + 2] = nrng;
        cx->stat[4 * t + 3] = cx->S;
    }
    free(u);
    free(s);
    free(pf);
    free(fl);
}

int orc_chord_fx(const int16_t *J, const int32_t *h, const int32_t *h0, int B, int N, const int32_t *T,
                 const int32_t *c, int S, int q, int P, int R, uint64_t seed, int64_t inst0, int mode, int8_t *out_s,
                 int64_t *out_e, int64_t *stats, int nthreads)
{
    if (N < 1 || P < 1 || N % P || R < 1 || S < 0 || B < 0) return -1;
    int8_t *s_all = mode == 1 ? (int8_t *)malloc((size_t)B * R * N) : out_s;
    int64_t *e_all = mode == 1 ? (int64_t *)malloc(sizeof(int64_t) * (size_t)B * R) : out_e;
    int64_t *st = stats ? (int64_t *)calloc((size_t)B * R * 4, sizeof(int64_t)) : NULL;
    chord_ctx cx = {J, h, h0, T, c, N, S, q, P, R, (uint32_t)(seed & MASK32), (uint32_t)(seed >> 32), inst0, s_all, e_all, st};
    run_pool(chord_item, &cx, (long)B * R, nthreads);
    if (stats) {
        for (int x = 0; x < 4; x++) stats[x] = 0;
        for (long t = 0; t < (long)B * R; t++)
            for (int x = 0; x < 4; x++) stats[x] += st[4 * t + x];
        free(st);
    }
    if (mode == 1) {
        for (long b = 0; b < B; b++) {
            int im = 0;
            for (int r = 1; r < R; r++)
                if (e_all[b * R + r] < e_all[b * R + im]) im = r;
            out_e[b] = e_all[b * R + im];
            memcpy(out_s + (size_t)b * N, s_all + ((size_t)b * R + im) * N, (size_t)N);
        }
        free(s_all);
        free(e_all);
    }
    return 0;
}

/* ------------------------------------------------------------------ float twins (FP32 and FP64) */
/*
 * Same trajectories as the integer anneals when run in float64 on dyadic inputs (J, h, T, q = integers / 2^12):
 * the fixed-point test x 4T > (z + 2T) 2^32 is the float test z < (x 2^-32 - 1/2) 4T. In FP32 the operations are
 * done in this order, without fused multiply-add:
 *   r = ((float)x * 2^-32 - 0.5f) * (4 T)        z = s * u + q (Arpeggio: s * w + 2 G_jj + q)
 * Arpeggio float: energies in E0 units (the scaled Ising energy -h0.s - s^T J s); a flip of lane j with old spin s
 * adds 2 s u_j = 2 s w_j + 4 G_jj to de (the integer path's de is 2 * 2^12 times this on dyadic inputs) and
 * 4 s G_ij to every field w_i.
 * Chord float: u = h + 2 J s, z = s u + q - c_k (prevflip ? -1 : 1), energy E0 = -h0.s - s^T J s tracked per flip as
 * dE = 2 s_j (u_j - h_j + h0_j) - 4 J_jj, best on strict decrease.
 */
#define FLOAT_TWINS(FT, SUF)                                                                                          \
    typedef struct {                                                                                                  \
        const FT *G, *h, *T;                                                                                          \
        const uint32_t *inst;                                                                                         \
        int D, P, R, L, start;                                                                                        \
        FT q;                                                                                                         \
        uint32_t k0, k1;                                                                                              \
        int8_t *out_s;                                                                                                \
        FT *out_de, *out_w;                                                                                           \
        int64_t *out_st;                                                                                              \
    } farp_ctx_##SUF;                                                                                                 \
    static void farp_item_##SUF(void *vctx, long t)                                                                   \
    {                                                                                                                 \
        farp_ctx_##SUF *c = (farp_ctx_##SUF *)vctx;                                                                   \
        long b = t / c->R;                                                                                            \
        uint32_t r = (uint32_t)(t % c->R), inst = c->inst[b];                                                         \
        int D = c->D, P = c->P;                                                                                       \
        const FT *G = c->G + (size_t)b * D * D, *h = c->h + (size_t)b * D;                                            \
        int8_t *s = (int8_t *)malloc((size_t)P * D);                                                                  \
        FT *w = (FT *)malloc(sizeof(FT) * D);                                                                         \
        int8_t *bs = c->out_s + (size_t)t * P * D;                                                                    \
        if (c->start == 0) {                                                                                          \
            for (int p = 0; p < P; p++)                                                                               \
                for (int j = 0; j < D; j++)                                                                           \
                    s[p * D + j] =                                                                                    \
                        philox_word(MASK32, (uint32_t)(p * D + j), r, inst, c->k0, c->k1, 0) >= 0x80000000u ? 1 : -1; \
        } else {                                                                                                      \
            for (int j = 0; j < D; j++) {                                                                             \
                uint32_t x = philox_word(MASK32, (uint32_t)j, r, inst, c->k0, c->k1, 0);                              \
                int pat = orc_bal_pat(P, (uint32_t)(((uint64_t)x * (uint64_t)bal_count(P)) >> 32));                   \
                for (int p = 0; p < P; p++) s[p * D + j] = ((pat >> p) & 1) ? 1 : -1;                                 \
            }                                                                                                         \
        }                                                                                                             \
        for (int i = 0; i < D; i++) {                                                                                 \
            FT acc = 0;                                                                                               \
            for (int j = 0; j < D; j++) {                                                                             \
                int m = 0;                                                                                            \
                for (int p = 0; p < P; p++) m += s[p * D + j];                                                        \
                acc = acc + G[(size_t)i * D + j] * (FT)m;                                                             \
            }                                                                                                         \
            w[i] = h[i] - (FT)2 * acc;                                                                                \
        }                                                                                                             \
        FT de = 0, best = (FT)INFINITY;                                                                               \
        int64_t flips = 0, tests = 0;                                                                                 \
        for (int k = 0; k < c->L; k++) {                                                                              \
            int p = k % P;                                                                                            \
            uint32_t n = (uint32_t)(k / P);                                                                           \
            FT T4 = (FT)4 * c->T[k];                                                                                  \
            int8_t *sp = s + (size_t)p * D;                                                                           \
            int64_t tt = 1;                                                                                           \
            for (int j = 0; j < D; j++) {                                                                             \
                uint32_t x = philox_word(n >> 2, (uint32_t)(p * D + j), r, inst, c->k0, c->k1, (int)(n & 3));         \
                FT rr = ((FT)x * (FT)2.3283064365386963e-10 - (FT)0.5) * T4;                                          \
                FT gjj = G[(size_t)j * D + j];                                                                        \
                FT z = (FT)sp[j] * w[j] + (FT)2 * gjj + c->q;                                                         \
                if (z < rr) {                                                                                         \
                    FT so = (FT)sp[j];                                                                                \
                    de = de + ((FT)2 * so * w[j] + (FT)4 * gjj);                                                      \
                    for (int i = 0; i < D; i++) w[i] = w[i] + (FT)4 * so * G[(size_t)i * D + j];                      \
                    sp[j] = (int8_t)(-sp[j]);                                                                         \
                    flips++;                                                                                          \
                    if (j < D - 1) tt++;                                                                              \
                }                                                                                                     \
            }                                                                                                         \
            tests += tt;                                                                                              \
            if (de < best) {                                                                                          \
                best = de;                                                                                            \
                memcpy(bs, s, (size_t)P * D);                                                                         \
            }                                                                                                         \
        }                                                                                                             \
        c->out_de[t] = best;                                                                                          \
        c->out_st[2 * t] = flips;                                                                                     \
        c->out_st[2 * t + 1] = tests;                                                                                 \
        if (c->out_w) memcpy(c->out_w + (size_t)t * D, w, sizeof(FT) * D);                                            \
        free(s);                                                                                                      \
        free(w);                                                                                                      \
    }                                                                                                                 \
    int orc_arpeggio_float_##SUF(const FT *G, const FT *h, int B, int D, int P, int R, const FT *T, int L, FT q,       \
                                 const uint32_t *inst, uint64_t key, int start, int8_t *out_s, FT *out_de,            \
                                 int64_t *out_st, FT *out_w, int nthreads)                                            \
    {                                                                                                                 \
        if (D < 1 || L < 1 || R < 1 || (P != 2 && P != 4 && P != 6)) return -1;                                       \
        farp_ctx_##SUF c = {G, h, T, inst, D, P, R, L, start, q, (uint32_t)(key & MASK32), (uint32_t)(key >> 32),     \
                            out_s, out_de, out_w, out_st};                                                            \
        run_pool(farp_item_##SUF, &c, (long)B * R, nthreads);                                                         \
        return 0;                                                                                                     \
    }                                                                                                                 \
    typedef struct {                                                                                                  \
        const FT *J, *h, *h0, *T, *c;                                                                                 \
        int N, S, P, R;                                                                                               \
        FT q;                                                                                                         \
        uint32_t k0, k1;                                                                                              \
        int64_t inst0;                                                                                                \
        int8_t *out_s;                                                                                                \
        FT *out_e, *out_u;                                                                                            \
        int64_t *out_fl;                                                                                              \
    } fch_ctx_##SUF;                                                                                                  \
    static void fch_item_##SUF(void *vctx, long t)                                                                    \
    {                                                                                                                 \
        fch_ctx_##SUF *cx = (fch_ctx_##SUF *)vctx;                                                                    \
        long b = t / cx->R;                                                                                           \
        uint32_t r = (uint32_t)(t % cx->R);                                                                           \
        int N = cx->N, P = cx->P, per = N / P;                                                                        \
        const FT *J = cx->J + (size_t)b * N * N, *h = cx->h + (size_t)b * N, *h0 = cx->h0 + (size_t)b * N;           \
        uint32_t inst = (uint32_t)(cx->inst0 + b);                                                                    \
        FT *u = (FT *)malloc(sizeof(FT) * N);                                                                         \
        int8_t *s = (int8_t *)malloc((size_t)N);                                                                      \
        uint8_t *pf = (uint8_t *)calloc((size_t)N, 1), *fl = (uint8_t *)malloc((size_t)N);                            \
        int8_t *best_s = cx->out_s + (size_t)t * N;                                                                   \
        for (int i = 0; i < N; i++)                                                                                   \
            s[i] = philox_word(MASK32, (uint32_t)i, r, inst, cx->k0, cx->k1, 0) >= 0x80000000u ? 1 : -1;              \
        for (int i = 0; i < N; i++) {                                                                                 \
            FT acc = 0;                                                                                               \
            for (int j = 0; j < N; j++) acc = acc + J[(size_t)i * N + j] * (FT)s[j];                                  \
            u[i] = h[i] + (FT)2 * acc;                                                                                \
        }                                                                                                             \
        FT e = 0;                                                                                                     \
        for (int i = 0; i < N; i++) e = e + (FT)s[i] * (h[i] - (FT)2 * h0[i] - u[i]);                                 \
        e = e * (FT)0.5;                                                                                              \
        FT best_e = (FT)INFINITY;                                                                                     \
        int64_t nfl = 0;                                                                                              \
        memcpy(best_s, s, (size_t)N);                                                                                 \
        for (int k = 0; k < cx->S; k++) {                                                                             \
            int p = k % P, n = k / P;                                                                                 \
            FT T4 = (FT)4 * cx->T[k], ck = cx->c[k];                                                                  \
            int any = 0;                                                                                              \
            for (int i = 0; i < N; i++) {                                                                             \
                fl[i] = 0;                                                                                            \
                if (i / per != p) continue;                                                                           \
                uint32_t x = philox_word((uint32_t)(n / 4), (uint32_t)i, r, inst, cx->k0, cx->k1, n % 4);             \
                FT rr = ((FT)x * (FT)2.3283064365386963e-10 - (FT)0.5) * T4;                                          \
                FT z = (FT)s[i] * u[i] + cx->q - ck * (pf[i] ? (FT)-1 : (FT)1);                                       \
                fl[i] = z < rr;                                                                                       \
                any |= fl[i];                                                                                         \
            }                                                                                                         \
            for (int j = 0; j < N; j++) {                                                                             \
                pf[j] = fl[j];                                                                                        \
                if (!fl[j]) continue;                                                                                 \
                FT sj = (FT)s[j];                                                                                     \
                e = e + ((FT)2 * sj * (u[j] - h[j] + h0[j]) - (FT)4 * J[(size_t)j * N + j]);                          \
                for (int i = 0; i < N; i++) u[i] = u[i] - (FT)4 * sj * J[(size_t)i * N + j];                          \
                s[j] = (int8_t)(-s[j]);                                                                               \
                nfl++;                                                                                                \
            }                                                                                                         \
            if ((any || k == 0) && e < best_e) {                                                                      \
                best_e = e;                                                                                           \
                memcpy(best_s, s, (size_t)N);                                                                         \
            }                                                                                                         \
        }                                                                                                             \
        cx->out_e[t] = best_e;                                                                                        \
        if (cx->out_u) memcpy(cx->out_u + (size_t)t * N, u, sizeof(FT) * N);                                          \
        if (cx->out_fl) cx->out_fl[t] = nfl;                                                                          \
        free(u);                                                                                                      \
        free(s);                                                                                                      \
        free(pf);                                                                                                     \
        free(fl);                                                                                                     \
    }                                                                                                                 \
    int orc_chord_float_##SUF(const FT *J, const FT *h, const FT *h0, int B, int N, const FT *T, const FT *c, int S,  \
                              FT q, int P, int R, uint64_t seed, int64_t inst0, int8_t *out_s, FT *out_e, FT *out_u,  \
                              int64_t *out_fl, int nthreads)                                                          \
    {                                                                                                                 \
        if (N < 1 || P < 1 || N % P || R < 1 || S < 0) return -1;                                                     \
        fch_ctx_##SUF cx = {J, h, h0, T, c, N, S, P, R, q, (uint32_t)(seed & MASK32), (uint32_t)(seed >> 32), inst0,  \
                            out_s, out_e, out_u, out_fl};                                                             \
        run_pool(fch_item_##SUF, &cx, (long)B * R, nthreads);                                                         \
        return 0;                                                                                                     \
    }

FLOAT_TWINS(float, f32)
FLOAT_TWINS(double, f64)

/* ------------------------------------------------------------------ lane-shared Chord (chord_ref.chord_ref) */
/*
 * G (D x D, symmetric, with diagonal), h (D), P copies, S steps with T_k, c_k, pinning q; anneal r of instance inst,
 * 64-bit seed. Streams: rng 1 = production xorshift32, one state per coordinate j seeded
 * fmix32(fmix32(fmix32(fmix32(lo ^ inst) ^ hi) ^ r) + (j + 1) 0x9E3779B9) (0 -> 0x9E3779B9), stepped once before the
 * start and once per step, start spin of copy p = bit 31 - p, step draw u = x >> 16 (RB = 16 fraction bits); rng 0 =
 * Philox, start Philox(2^32 - 1, p D + j, r, inst)[0] >= 2^31, step draw Philox(n >> 2, p D + j, r, inst)[n & 3]
 * with n = k div P (RB = 32). Step k: p = k mod P, A = 4 T_k, B_k = q - c_k + 2 T_k; lane j of copy p flips iff
 * u A > (s_pj w_j + 2 G_jj + B_k) 2^RB; every flip of the step commits from the same fields: sf = s_p on flipped
 * lanes, de += 4 sf.w + 8 sf^T G sf, w += 4 G sf. The state is kept when de < best (first minimum).
 */
static inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    return h ^ (h >> 16);
}

static inline uint32_t xs32(uint32_t x)
{
    x ^= x << 13;
    x ^= x >> 17;
    return x ^ (x << 5);
}

static void chord_lane_one(const int32_t *G, const int32_t *h, int D, int P, const int32_t *T, const int32_t *c, int S,
                           int q, uint32_t r, uint32_t inst, uint64_t seed, int rng, int8_t *bs, int64_t *best_out,
                           int64_t *flips_out)
{
    const uint32_t lo = (uint32_t)(seed & MASK32), hi = (uint32_t)(seed >> 32);
    int8_t *s = (int8_t *)malloc((size_t)P * D);
    uint32_t *x = (uint32_t *)malloc(sizeof(uint32_t) * D);
    int64_t *m = (int64_t *)malloc(sizeof(int64_t) * D), *w = (int64_t *)malloc(sizeof(int64_t) * D);
    int64_t *sf = (int64_t *)malloc(sizeof(int64_t) * D), *gs = (int64_t *)malloc(sizeof(int64_t) * D);
    const uint32_t base = fmix32(fmix32(fmix32(lo ^ inst) ^ hi) ^ r);
    for (int j = 0; j < D; j++) {
        if (rng) {
            uint32_t st = fmix32(base + (uint32_t)(j + 1) * 0x9E3779B9u);
            x[j] = xs32(st ? st : 0x9E3779B9u);
            for (int p = 0; p < P; p++) s[p * D + j] = ((x[j] >> (31 - p)) & 1u) ? 1 : -1;
        } else {
            for (int p = 0; p < P; p++)
                s[p * D + j] = philox_word(MASK32, (uint32_t)(p * D + j), r, inst, lo, hi, 0) >= 0x80000000u ? 1 : -1;
        }
    }
    fields_of(G, h, D, P, s, m, w);
    int64_t de = 0, best = (int64_t)1 << 62, flips = 0;
    memcpy(bs, s, (size_t)P * D);
    const int RB = rng ? 16 : 32;
    for (int k = 0; k < S; k++) {
        const int p = k % P, n = k / P;
        const int64_t A = 4 * (int64_t)T[k], Bk = (int64_t)q - c[k] + 2 * (int64_t)T[k];
        for (int j = 0; j < D; j++) {
            uint64_t u;
            if (rng) {
                x[j] = xs32(x[j]);
                u = x[j] >> 16;
            } else {
                u = philox_word((uint32_t)(n >> 2), (uint32_t)(p * D + j), r, inst, lo, hi, n & 3);
            }
            const int64_t zp = s[p * D + j] * w[j] + 2 * (int64_t)G[(size_t)j * D + j] + Bk;
            const __int128 lhs = (__int128)u * A, rhs = (__int128)zp * ((__int128)1 << RB);
            sf[j] = lhs > rhs ? s[p * D + j] : 0;
        }
        int64_t d1 = 0, d2 = 0;
        for (int i = 0; i < D; i++) {
            int64_t acc = 0;
            for (int j = 0; j < D; j++) acc += (int64_t)G[(size_t)i * D + j] * sf[j];
            gs[i] = acc;
            d1 += sf[i] * w[i];
            d2 += sf[i] * acc;
        }
        de += 4 * d1 + 8 * d2;
        for (int i = 0; i < D; i++) {
            w[i] += 4 * gs[i];
            if (sf[i]) {
                s[p * D + i] = (int8_t)-s[p * D + i];
                flips++;
            }
        }
        if (de < best) {
            best = de;
            memcpy(bs, s, (size_t)P * D);
        }
    }
    *best_out = best;
    *flips_out = flips;
    free(s);
    free(x);
    free(m);
    free(w);
    free(sf);
    free(gs);
}

typedef struct {
    const int32_t *G, *h, *T, *c;
    const uint32_t *inst;
    int D, P, R, S, q, rng;
    uint64_t seed;
    int8_t *out_s;
    int64_t *out_de, *out_fl;
} chord_lane_ctx;

static void chord_lane_item(void *vctx, long t)
{
    chord_lane_ctx *c = (chord_lane_ctx *)vctx;
    long b = t / c->R;
    uint32_t r = (uint32_t)(t % c->R);
    chord_lane_one(c->G + (size_t)b * c->D * c->D, c->h + (size_t)b * c->D, c->D, c->P, c->T, c->c, c->S, c->q, r,
                   c->inst[b], c->seed, c->rng, c->out_s + (size_t)t * c->P * c->D, c->out_de + t, c->out_fl + t);
}

/* chord_ref.chord_ref on (B, D, D) / (B, D) integers: out_s (B, R, P D) copy-major best spins, out_de (B, R) best
 * energy relative to the start, out_fl (B, R) flips. */
int orc_chord_lane(const int32_t *G, const int32_t *h, int B, int D, int P, int R, const int32_t *T, const int32_t *c,
                   int S, int q, const uint32_t *inst, uint64_t seed, int rng, int8_t *out_s, int64_t *out_de,
                   int64_t *out_fl, int nthreads)
{
    if (D < 1 || S < 0 || R < 1 || B < 0 || P < 1 || P > 32) return -1;
    chord_lane_ctx cx = {G, h, T, c, inst, D, P, R, S, q, rng, seed, out_s, out_de, out_fl};
    run_pool(chord_lane_item, &cx, (long)B * R, nthreads);
    return 0;
}
