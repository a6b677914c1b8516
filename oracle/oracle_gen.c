/*
 * oracle_gen.c -- CPU restatement of the reference's graph generator
 * (generator/prng_engine.hpp, generator/rmat_generator.h,
 * generator/edge_list_utils.h) plus the build's own seeded relabelling and
 * shuffle.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ sitmo */

static inline uint64_t rotl64(uint64_t x, unsigned r) {
    return (x << r) | (x >> (64u - r));
}

/* prng_engine.hpp:305-341.  Twenty rounds of a Threefish-256-style mix on the
 * 256-bit counter; a key word is injected every fourth round.  Even rounds pair
 * words (0,1),(2,3); odd rounds pair (0,3),(2,1).  The rotation amounts of the
 * eight-round cycle and the key-schedule constant are those of the published
 * sitmo v5 engine. */
void oracle_sitmo_block(const uint64_t kin[4], const uint64_t s[4], uint64_t out[4]) {
    static const unsigned rot[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                                       {25, 33}, {46, 12}, {58, 22}, {32, 32}};
    uint64_t k[5];
    uint64_t b[4];
    for (int i = 0; i < 4; ++i) {
        k[i] = kin[i];
        b[i] = s[i];
    }
    k[4] = 0x1BD11BDAA9FC1A22ULL ^ k[0] ^ k[1] ^ k[2] ^ k[3];
    for (int r = 0; r < 20; ++r) {
        /* pair (x0,x1) and (z0,z1) */
        int x1 = (r & 1) ? 3 : 1;
        int z1 = (r & 1) ? 1 : 3;
        if ((r & 3) == 0) {
            int j = r >> 2;
            b[x1] += k[(j + 1) % 5];
            b[z1] += k[(j + 3) % 5] + (uint64_t)j;
            b[0] += b[x1] + k[j % 5];
            b[2] += b[z1] + k[(j + 2) % 5];
        } else {
            b[0] += b[x1];
            b[2] += b[z1];
        }
        b[x1] = rotl64(b[x1], rot[r & 7][0]);
        b[z1] = rotl64(b[z1], rot[r & 7][1]);
        b[x1] ^= b[0];
        b[z1] ^= b[2];
    }
    for (int i = 0; i < 4; ++i) out[i] = b[i] + k[i];
    out[3] += 5;
}

/* Output n of an engine seeded with seed(uint32): block n/8 of counter
 * {n/8,0,0,0}; even words are the low halves, odd words the high halves
 * (prng_engine.hpp:190-211). */
uint32_t oracle_sitmo_word(uint32_t seed, uint64_t n) {
    uint64_t k[4] = {seed, 0, 0, 0};
    uint64_t s[4] = {n >> 3, 0, 0, 0};
    uint64_t o[4];
    oracle_sitmo_block(k, s, o);
    unsigned w = (unsigned)(n & 7);
    uint64_t x = o[w >> 1];
    return (w & 1) ? (uint32_t)(x >> 32) : (uint32_t)(x & 0xFFFFFFFFu);
}

typedef struct {
    uint64_t k[4];
    uint64_t pos;      /* next output index */
    uint64_t blk;      /* cached block index, UINT64_MAX = none */
    uint64_t o[4];
} sitmo_stream;

static void stream_init(sitmo_stream *st, uint32_t seed) {
    memset(st, 0, sizeof(*st));
    st->k[0] = seed;
    st->blk = UINT64_MAX;
}

static uint32_t stream_next(sitmo_stream *st) {
    uint64_t b = st->pos >> 3;
    if (b != st->blk) {
        uint64_t s[4] = {b, 0, 0, 0};
        oracle_sitmo_block(st->k, s, st->o);
        st->blk = b;
    }
    unsigned w = (unsigned)(st->pos & 7);
    st->pos++;
    uint64_t x = st->o[w >> 1];
    return (w & 1) ? (uint32_t)(x >> 32) : (uint32_t)(x & 0xFFFFFFFFu);
}

/* std::generate_canonical<double, 53> over a 32-bit engine takes k = 2 draws:
 * (u0 + u1 * 2^32) / 2^64 rounded in double, clamped below 1
 * (rmat_generator.h:19-59 and libstdc++ bits/random.tcc). */
static double stream_canonical(sitmo_stream *st) {
    double u0 = (double)stream_next(st);
    double u1 = (double)stream_next(st);
    double sum = u0;
    sum += u1 * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* rmat_generator.h:91-140 next_edge */
static void rmat_next(sitmo_stream *st, int scale, double a, double b, double c,
                      double d, int64_t *src, int64_t *dst) {
    double A = a, B = b, C = c, D = d;
    int64_t i = 0, j = 0;
    int64_t bit = (int64_t)1 << (scale - 1);
    for (;;) {
        const double r = stream_canonical(st);
        if (r > A) {
            if (r <= A + B)
                j |= bit;
            else if (r <= A + B + C)
                i |= bit;
            else {
                j |= bit;
                i |= bit;
            }
        }
        if (bit == 1) break;
        A *= (9.5 + stream_canonical(st)) / 10;
        B *= (9.5 + stream_canonical(st)) / 10;
        C *= (9.5 + stream_canonical(st)) / 10;
        D *= (9.5 + stream_canonical(st)) / 10;
        {
            const double norm = 1.0 / (A + B + C + D);
            A *= norm;
            B *= norm;
            C *= norm;
        }
        D = 1.0 - (A + B + C);
        bit >>= 1;
    }
    *src = i;
    *dst = j;
}

static int floor_log2(int64_t nv) {
    int s = 0;
    while (nv >>= 1) ++s;
    return s;
}

/* rmat_generator.h:143-193 rmat_fill: edge i draws from stream position
 * i*(5*(SCALE-1)+1)*2; afterwards self-loops are re-rolled serially from the
 * continuing stream. */
int oracle_rmat_raw(int64_t nv, int64_t num_edges, double a, double b, double c,
                    double d, uint32_t seed, int64_t *src, int64_t *dst) {
    int scale = floor_log2(nv);
    if (scale < 1 || num_edges < 0) return -1;
    sitmo_stream st;
    stream_init(&st, seed);
    for (int64_t i = 0; i < num_edges; ++i) rmat_next(&st, scale, a, b, c, d, &src[i], &dst[i]);
    for (int64_t i = 0; i < num_edges; ++i)
        while (src[i] == dst[i]) rmat_next(&st, scale, a, b, c, d, &src[i], &dst[i]);
    return 0;
}

/* ------------------------------------------------- flip / sort / dedup */

typedef struct {
    int64_t s, d;
} pair64;

static int cmp_pair(const void *x, const void *y) {
    const pair64 *p = (const pair64 *)x, *q = (const pair64 *)y;
    if (p->s != q->s) return p->s < q->s ? -1 : 1;
    if (p->d != q->d) return p->d < q->d ? -1 : 1;
    return 0;
}

/* edge_list_utils.h:13-57: orient src < dst, sort by (src, dst), unique. */
int64_t oracle_flip_sort_dedup(int64_t n, int64_t *src, int64_t *dst) {
    if (n <= 0) return 0;
    pair64 *p = (pair64 *)malloc(sizeof(pair64) * (size_t)n);
    if (!p) return -1;
    for (int64_t i = 0; i < n; ++i) {
        int64_t s = src[i], d = dst[i];
        if (s > d) {
            int64_t t = s;
            s = d;
            d = t;
        }
        p[i].s = s;
        p[i].d = d;
    }
    qsort(p, (size_t)n, sizeof(pair64), cmp_pair);
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (m > 0 && p[m - 1].s == p[i].s && p[m - 1].d == p[i].d) continue;
        p[m++] = p[i];
    }
    for (int64_t i = 0; i < m; ++i) {
        src[i] = p[i].s;
        dst[i] = p[i].d;
    }
    free(p);
    return m;
}

/* ---------------------------------------- build-defined relabel / shuffle */

static inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Four rounds of (odd multiply, add, xor-shift-right) on s-bit integers is a
 * bijection of [0, 2^s); cycle-walking restricts it to [0, nv). */
int64_t oracle_vperm(int64_t x, int64_t nv, uint64_t seed) {
    if (nv <= 1) return x;
    int s = 0;
    while (((int64_t)1 << s) < nv) ++s;
    uint64_t mask = (s >= 64) ? ~0ULL : (((uint64_t)1 << s) - 1);
    unsigned h = (unsigned)((s + 1) / 2);
    uint64_t y = (uint64_t)x;
    do {
        for (uint64_t r = 0; r < 4; ++r) {
            uint64_t k = splitmix64(seed * 4 + r + 0x5851F42D4C957F2DULL);
            y = (y * (k | 1)) & mask;
            y = (y + (k >> 32)) & mask;
            y ^= y >> h;
        }
    } while (y >= (uint64_t)nv);
    return (int64_t)y;
}

/* splitmix64 is a bijection of uint64, so keys of distinct j never tie. */
uint64_t oracle_shuffle_key(uint64_t j, uint64_t seed) {
    uint64_t sk = splitmix64(seed ^ 0xD1B54A32D192ED03ULL);
    return splitmix64(j ^ sk);
}

typedef struct {
    uint64_t key;
    int64_t s, d;
} keyed_edge;

static int cmp_keyed(const void *x, const void *y) {
    const keyed_edge *p = (const keyed_edge *)x, *q = (const keyed_edge *)y;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return 0;
}

static int shuffle_edges(int64_t m, int64_t *src, int64_t *dst, uint64_t seed) {
    if (m <= 0) return 0;
    keyed_edge *e = (keyed_edge *)malloc(sizeof(keyed_edge) * (size_t)m);
    if (!e) return -1;
    for (int64_t j = 0; j < m; ++j) {
        e[j].key = oracle_shuffle_key((uint64_t)j, seed);
        e[j].s = src[j];
        e[j].d = dst[j];
    }
    qsort(e, (size_t)m, sizeof(keyed_edge), cmp_keyed);
    for (int64_t j = 0; j < m; ++j) {
        src[j] = e[j].s;
        dst[j] = e[j].d;
    }
    free(e);
    return 0;
}

int oracle_make_rmat(int64_t nv, int64_t num_edges, double a, double b, double c,
                     double d, uint32_t seed, int64_t *src, int64_t *dst,
                     int64_t *m_out) {
    if (oracle_rmat_raw(nv, num_edges, a, b, c, d, seed, src, dst)) return -1;
    int64_t m = oracle_flip_sort_dedup(num_edges, src, dst);
    if (m < 0) return -1;
    for (int64_t j = 0; j < m; ++j) {
        src[j] = oracle_vperm(src[j], nv, seed);
        dst[j] = oracle_vperm(dst[j], nv, seed);
    }
    if (shuffle_edges(m, src, dst, seed)) return -1;
    *m_out = m;
    return 0;
}

static inline int64_t mulhi_range(uint64_t x, int64_t nv) {
    return (int64_t)(((unsigned __int128)x * (unsigned __int128)(uint64_t)nv) >> 64);
}

int oracle_make_er(int64_t nv, int64_t num_edges, uint32_t seed, int64_t *src,
                   int64_t *dst, int64_t *m_out) {
    if (nv < 2 || num_edges < 0) return -1;
    sitmo_stream st;
    stream_init(&st, seed);
    int64_t n = 0;
    for (int64_t i = 0; i < num_edges; ++i) {
        uint64_t w0 = stream_next(&st), w1 = stream_next(&st);
        uint64_t w2 = stream_next(&st), w3 = stream_next(&st);
        int64_t s = mulhi_range(w0 | (w1 << 32), nv);
        int64_t t = mulhi_range(w2 | (w3 << 32), nv);
        if (s == t) continue;
        src[n] = s;
        dst[n] = t;
        ++n;
    }
    int64_t m = oracle_flip_sort_dedup(n, src, dst);
    if (m < 0) return -1;
    if (shuffle_edges(m, src, dst, seed)) return -1;
    *m_out = m;
    return 0;
}

int oracle_make_lattice(int64_t rows, int64_t cols, uint32_t seed, int64_t *src,
                        int64_t *dst, int64_t *m_out) {
    if (rows < 1 || cols < 1) return -1;
    int64_t m = 0;
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t c = 0; c < cols; ++c) {
            int64_t v = r * cols + c;
            if (c + 1 < cols) {
                src[m] = v;
                dst[m] = v + 1;
                ++m;
            }
            if (r + 1 < rows) {
                src[m] = v;
                dst[m] = v + cols;
                ++m;
            }
        }
    if (shuffle_edges(m, src, dst, seed)) return -1;
    *m_out = m;
    return 0;
}

/* edge_list_utils.h:171-198 dump_bin: one text header line, then int64 pairs. */
int oracle_write_el64(const char *path, int64_t nv, int64_t m, const int64_t *src,
                      const int64_t *dst) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    fprintf(fp,
            " --format el64 --num_edges %lld --num_vertices %lld --is_undirected"
            " --is_deduped --is_permuted\n",
            (long long)m, (long long)nv);
    for (int64_t i = 0; i < m; ++i) {
        int64_t e[2] = {src[i], dst[i]};
        if (fwrite(e, sizeof(int64_t), 2, fp) != 2) {
            fclose(fp);
            return -1;
        }
    }
    return fclose(fp) == 0 ? 0 : -1;
}
