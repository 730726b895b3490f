/*
 * Seeded synthetic CNF generators for the simplifier workloads.
 *
 * Literal encoding is the simplifier's: lit = 2*v + sign, sign 1 = negative,
 * variables numbered 1..nvars (ParaFROST src/cpu/const.h:72-78).
 * Every generator writes clauses as a flat literal array plus an offsets
 * array of length ncls+1 (CSR), literals in generation order (the simplifier
 * sorts them on upload, as ParaFROST's extract() does, src/cpu/simplify.cpp:95-112).
 *
 * Workloads:
 *   uniform k-SAT      : k distinct uniform variables per clause, fair signs.
 *   geometric widths   : width = wmin + (number of fair-coin heads before a tail), clamped to wmax.
 *   scale-free + gates : variable i (by a seeded permutation rank r) is drawn with
 *                        weight (r+1)^(-1/(beta-1)) which yields a power-law occurrence
 *                        distribution with exponent beta (scale-free random SAT model);
 *                        widths 40% binary, else 3 + geometric clamped to 12; a fraction
 *                        of the clauses are Tseitin AND/OR gate definitions with fan-in 2..4
 *                        and a dedicated output variable.
 * The PRNG is xoshiro256** seeded through splitmix64, so outputs are identical on every host.
 */
#include <stdint.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(const uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return result;
}

/* uniform integer in [0, n) by 64x64->128 multiply-shift */
static inline uint64_t rng_below(rng_t *r, uint64_t n)
{
    return (uint64_t)(((__uint128_t)rng_next(r) * n) >> 64);
}

static inline int rng_coin(rng_t *r) { return (int)(rng_next(r) >> 63); }

static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static inline int contains_var(const uint32_t *lits, int n, uint32_t v)
{
    for (int i = 0; i < n; i++) if ((lits[i] >> 1) == v) return 1;
    return 0;
}

static inline int geom_width(rng_t *r, int wmin, int wmax)
{
    int w = wmin;
    while (w < wmax && rng_coin(r)) w++;
    return w;
}

/* k-SAT with k distinct uniform variables per clause. Returns #literals. */
int64_t pfgen_uniform_ksat(uint32_t nvars, uint64_t ncls, int k, uint64_t seed,
                           uint32_t *lits, uint64_t *offs)
{
    if (nvars < (uint32_t)k || k < 1) return -1;
    rng_t r; rng_seed(&r, seed);
    uint64_t n = 0;
    offs[0] = 0;
    for (uint64_t c = 0; c < ncls; c++) {
        uint32_t *cl = lits + n;
        for (int j = 0; j < k; j++) {
            uint32_t v;
            do { v = 1 + (uint32_t)rng_below(&r, nvars); } while (contains_var(cl, j, v));
            cl[j] = 2 * v + (uint32_t)rng_coin(&r);
        }
        n += k;
        offs[c + 1] = n;
    }
    return (int64_t)n;
}

/* Mixed widths: wmin + geometric(1/2) clamped to wmax, distinct uniform vars. */
int64_t pfgen_geometric(uint32_t nvars, uint64_t ncls, int wmin, int wmax, uint64_t seed,
                        uint32_t *lits, uint64_t *offs)
{
    if (wmin < 1 || wmax < wmin || nvars < (uint32_t)wmax) return -1;
    rng_t r; rng_seed(&r, seed);
    uint64_t n = 0;
    offs[0] = 0;
    for (uint64_t c = 0; c < ncls; c++) {
        const int w = geom_width(&r, wmin, wmax);
        uint32_t *cl = lits + n;
        for (int j = 0; j < w; j++) {
            uint32_t v;
            do { v = 1 + (uint32_t)rng_below(&r, nvars); } while (contains_var(cl, j, v));
            cl[j] = 2 * v + (uint32_t)rng_coin(&r);
        }
        n += w;
        offs[c + 1] = n;
    }
    return (int64_t)n;
}

/* ---------------- scale-free CNF with Tseitin gate definitions ---------------- */

typedef struct {
    double *cdf;     /* cumulative weights by rank */
    uint32_t *perm;  /* rank -> variable */
    uint32_t n;
} plaw_t;

static int plaw_init(plaw_t *p, uint32_t nvars, double beta, rng_t *r)
{
    p->n = nvars;
    p->cdf = (double *)malloc(sizeof(double) * nvars);
    p->perm = (uint32_t *)malloc(sizeof(uint32_t) * nvars);
    if (!p->cdf || !p->perm) return -1;
    const double a = 1.0 / (beta - 1.0);
    double acc = 0.0;
    for (uint32_t i = 0; i < nvars; i++) {
        acc += pow((double)i + 1.0, -a);
        p->cdf[i] = acc;
        p->perm[i] = i + 1;
    }
    for (uint32_t i = nvars - 1; i > 0; i--) {
        uint32_t j = (uint32_t)rng_below(r, (uint64_t)i + 1);
        uint32_t t = p->perm[i]; p->perm[i] = p->perm[j]; p->perm[j] = t;
    }
    return 0;
}

static inline uint32_t plaw_draw(const plaw_t *p, rng_t *r)
{
    const double u = rng_unit(r) * p->cdf[p->n - 1];
    uint32_t lo = 0, hi = p->n - 1;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (p->cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    return p->perm[lo];
}

/* Scale-free instance. gate_frac of the clauses (rounded up to whole gates) are
 * AND/OR definitions; output variables are the top ids (nvars - ngates + 1 .. nvars)
 * which also take part in the power-law draw. Clause order is shuffled.
 * Buffers must hold ncls clauses of up to 12 literals (gate clauses are <= 5).
 * Returns #literals, or -1 on bad arguments / allocation failure. */
int64_t pfgen_scalefree(uint32_t nvars, uint64_t ncls, double beta, double gate_frac,
                        uint64_t seed, uint32_t *lits, uint64_t *offs)
{
    if (nvars < 64 || beta <= 1.0 || gate_frac < 0 || gate_frac > 0.9) return -1;
    rng_t r; rng_seed(&r, seed);
    plaw_t pl;
    if (plaw_init(&pl, nvars, beta, &r)) return -1;
    /* decide the gates first */
    const uint64_t gate_target = (uint64_t)ceil(gate_frac * (double)ncls);
    uint64_t ngates = 0, gcls = 0;
    uint8_t *fanin = (uint8_t *)malloc(gate_target + 1);
    while (gcls < gate_target) {
        int k = 2 + (int)rng_below(&r, 3);
        fanin[ngates++] = (uint8_t)k;
        gcls += (uint64_t)k + 1;
    }
    if (gcls > ncls || ngates >= nvars / 2) { free(fanin); free(pl.cdf); free(pl.perm); return -1; }
    uint32_t *tl = (uint32_t *)malloc(sizeof(uint32_t) * ncls * 12);
    uint64_t *to = (uint64_t *)malloc(sizeof(uint64_t) * (ncls + 1));
    uint64_t n = 0, c = 0;
    to[0] = 0;
    const uint32_t out0 = nvars - (uint32_t)ngates + 1;
    for (uint64_t g = 0; g < ngates; g++) {
        const int k = fanin[g];
        const uint32_t out = out0 + (uint32_t)g;
        const int isand = rng_coin(&r);
        uint32_t in[4];
        for (int j = 0; j < k; j++) {
            uint32_t v;
            int dup;
            do {
                v = plaw_draw(&pl, &r);
                dup = (v == out);
                for (int q = 0; q < j && !dup; q++) dup = ((in[q] >> 1) == v);
            } while (dup);
            in[j] = 2 * v + (uint32_t)rng_coin(&r);
        }
        /* AND: out -> in_j  and  (in_1 & .. & in_k) -> out
           OR : in_j -> out  and  out -> (in_1 | .. | in_k) */
        const uint32_t pout = 2 * out, nout = 2 * out + 1;
        for (int j = 0; j < k; j++) {
            tl[n++] = isand ? nout : pout;
            tl[n++] = isand ? in[j] : (in[j] ^ 1u);
            to[++c] = n;
        }
        tl[n++] = isand ? pout : nout;
        for (int j = 0; j < k; j++) tl[n++] = isand ? (in[j] ^ 1u) : in[j];
        to[++c] = n;
    }
    while (c < ncls) {
        const int w = rng_below(&r, 10) < 4 ? 2 : geom_width(&r, 3, 12);
        uint32_t *cl = tl + n;
        for (int j = 0; j < w; j++) {
            uint32_t v;
            do { v = plaw_draw(&pl, &r); } while (contains_var(cl, j, v));
            cl[j] = 2 * v + (uint32_t)rng_coin(&r);
        }
        n += w;
        to[++c] = n;
    }
    /* shuffle clause order */
    uint64_t *order = (uint64_t *)malloc(sizeof(uint64_t) * ncls);
    for (uint64_t i = 0; i < ncls; i++) order[i] = i;
    for (uint64_t i = ncls - 1; i > 0; i--) {
        uint64_t j = rng_below(&r, i + 1);
        uint64_t t = order[i]; order[i] = order[j]; order[j] = t;
    }
    uint64_t m = 0;
    offs[0] = 0;
    for (uint64_t i = 0; i < ncls; i++) {
        const uint64_t s = to[order[i]], e = to[order[i] + 1];
        memcpy(lits + m, tl + s, sizeof(uint32_t) * (e - s));
        m += e - s;
        offs[i + 1] = m;
    }
    free(order); free(tl); free(to); free(fanin); free(pl.cdf); free(pl.perm);
    return (int64_t)m;
}

/* DIMACS writer (p cnf header, one clause per line, 0-terminated). */
int pfgen_write_dimacs(const char *path, uint32_t nvars, uint64_t ncls,
                       const uint32_t *lits, const uint64_t *offs)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    static char buf[1 << 16];
    setvbuf(f, buf, _IOFBF, sizeof(buf));
    fprintf(f, "p cnf %u %llu\n", nvars, (unsigned long long)ncls);
    for (uint64_t c = 0; c < ncls; c++) {
        for (uint64_t k = offs[c]; k < offs[c + 1]; k++) {
            const uint32_t l = lits[k];
            fprintf(f, "%s%u ", (l & 1u) ? "-" : "", l >> 1);
        }
        fputs("0\n", f);
    }
    return fclose(f);
}

/* BMC-like requests: a random and-inverter graph (AIG) unrolled over time frames.
 * Nodes: nl latches, ni primary inputs, ng 2-input AND gates; gate j takes two earlier
 * nodes, each edge inverted with probability pinv; every latch's next state is a random
 * node literal. Frame 0 has a variable per latch; in frame t > 0 a latch is the literal
 * of its next-state node in frame t - 1. Every frame has a variable per input and gate:
 *   var(latch i, frame 0) = 1 + i
 *   var(input i, frame t) = 1 + nl + t (ni + ng) + i,  var(gate j, t) = ... + ni + j
 * A gate g = a & b is Tseitin-encoded as (~g | a), (~g | b), (g | ~a | ~b); a clause
 * whose literals collapse (a == b, a == ~b) is written the way it comes out, with a
 * repeated literal dropped and a tautology skipped. Only frames [t0, t1) are emitted,
 * so an instance grows by generating its new frames. The structure depends on the
 * seed only. Returns the literal count (offs gets ncls + 1 entries); *ncls_out and
 * *nvars_out receive the clause count and the variable count after t1 frames.
 * iface (nullable, nl entries) receives the variables of the latch literals at the
 * start of frame t1: the only earlier variables that frames t1, t1 + 1, .. refer to. */
typedef struct { uint32_t a, b; } aig_gate;

int64_t pfgen_aig_bmc(uint64_t seed, uint32_t ng, uint32_t nl, uint32_t ni, double pinv, uint32_t t0, uint32_t t1,
                      uint32_t *lits, uint64_t *offs, uint64_t *ncls_out, uint32_t *nvars_out, uint32_t *iface)
{
    if (t1 < t0 || ng == 0) return -1;
    rng_t r;
    rng_seed(&r, seed);
    const uint32_t nn = nl + ni + ng;   /* node literal = 2 node + sign */
    aig_gate *g = (aig_gate *)malloc(sizeof(aig_gate) * ng);
    uint32_t *next = (uint32_t *)malloc(sizeof(uint32_t) * (nl ? nl : 1));
    uint32_t *lat = (uint32_t *)malloc(sizeof(uint32_t) * (nl ? nl : 1));    /* latch literals of the frame */
    uint32_t *nlat = (uint32_t *)malloc(sizeof(uint32_t) * (nl ? nl : 1));
    if (!g || !next || !lat || !nlat) { free(g); free(next); free(lat); free(nlat); return -1; }
    for (uint32_t j = 0; j < ng; j++) {
        const uint32_t lim = nl + ni + j;
        const uint32_t x = (uint32_t)rng_below(&r, lim ? lim : 1), y = (uint32_t)rng_below(&r, lim ? lim : 1);
        g[j].a = 2 * x + (rng_unit(&r) < pinv);
        g[j].b = 2 * y + (rng_unit(&r) < pinv);
    }
    for (uint32_t i = 0; i < nl; i++) next[i] = 2 * (uint32_t)rng_below(&r, nn) + (rng_unit(&r) < pinv);
    for (uint32_t i = 0; i < nl; i++) lat[i] = 2 * (1 + i);
    const uint32_t per = ni + ng;
    uint64_t w = 0, c = 0;
    offs[0] = 0;
    for (uint32_t t = 0; t < t1; t++) {
        const uint32_t fb = 1 + nl + t * per;   /* first variable of frame t */
        /* literal of node-literal x in frame t */
#define NODE_LIT(x) (((x) >> 1) < nl ? (lat[(x) >> 1] ^ ((x) & 1u)) : (2 * (fb + ((x) >> 1) - nl) + ((x) & 1u)))
        if (t >= t0) {
            for (uint32_t j = 0; j < ng; j++) {
                const uint32_t gv = 2 * (fb + ni + j);
                const uint32_t a = NODE_LIT(g[j].a), b = NODE_LIT(g[j].b);
                /* (~g | a) */
                lits[w++] = gv ^ 1u; lits[w++] = a; offs[++c] = w;
                /* (~g | b) */
                if (b != a) { lits[w++] = gv ^ 1u; lits[w++] = b; offs[++c] = w; }
                /* (g | ~a | ~b) */
                if (a == (b ^ 1u)) continue;   /* tautology */
                lits[w++] = gv; lits[w++] = a ^ 1u;
                if (b != a) lits[w++] = b ^ 1u;
                offs[++c] = w;
            }
        }
        for (uint32_t i = 0; i < nl; i++) nlat[i] = NODE_LIT(next[i]);
        for (uint32_t i = 0; i < nl; i++) lat[i] = nlat[i];
#undef NODE_LIT
    }
    if (iface)
        for (uint32_t i = 0; i < nl; i++) iface[i] = lat[i] >> 1;
    free(g); free(next); free(lat); free(nlat);
    if (ncls_out) *ncls_out = c;
    if (nvars_out) *nvars_out = nl + t1 * per;
    return (int64_t)w;
}

/* SCLAUSE images ([word0, sig, size, lits..] per clause) to CSR: the host-side
 * write-back of a downloaded clause database. offs gets ncls + 1 entries (at most
 * nwords / 3 + 1). Returns the literal count, -1 on a malformed image. */
int64_t pfgen_images_to_csr(const uint32_t *words, uint64_t nwords, uint32_t *lits, uint64_t *offs, uint64_t *ncls_out)
{
    uint64_t w = 0, c = 0, n = 0;
    offs[0] = 0;
    while (w < nwords) {
        if (w + 3 > nwords) return -1;
        const uint32_t sz = words[w + 2];
        if (w + 3 + sz > nwords) return -1;
        memcpy(lits + n, words + w + 3, sizeof(uint32_t) * sz);
        n += sz;
        offs[++c] = n;
        w += 3ull + sz;
    }
    if (ncls_out) *ncls_out = c;
    return (int64_t)n;
}
