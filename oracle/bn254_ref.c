/*
 * oracle/bn254_ref.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, written for this project, of zkMetal's reference CPU
 * Pippenger for BN254 G1 (Sources/NeonFieldOps/bn254_msm.c).  It is the parity
 * checker for the HIP MSM and the timed "CPU baseline".  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Algorithm followed (same steps, own code):
 *   - Fp: 4x64-bit little-endian Montgomery limbs, CIOS multiply with a
 *     128-bit accumulator (bn254_msm.c:38-154), add/sub/double with one
 *     conditional correction (:269-330), Fermat inverse a^(p-2) (:337-354).
 *   - Jacobian points (X,Y,Z), identity Z == 0 stored as (1,1,0)
 *     (:362-370); doubling for a = 0 (:373-418); full addition with the
 *     H == 0 exception rule (:421-469); mixed addition where the affine point
 *     (0,0) is the identity (:472-533).
 *   - batch_to_affine with one inversion, identity -> (0,0) (:544-599).
 *   - unsigned c-bit windows over all 256 scalar bits, no reduction mod r
 *     (:605-615), window width from optimal_window_bits (:621-631).
 *   - per-window worker: bucket accumulation with mixed adds, batch to
 *     affine, running-sum reduction (:647-698), one thread per window
 *     (pthreads here, GCD dispatch_apply there) and Horner (:1145-1187).
 *
 * Also here, for tests: GLV decomposition (Sources/zkMetal/MSM/GLV.swift:37-92),
 * signed-digit recoding (Sources/Shaders/msm/msm_kernels.metal:399-439), and
 * scalar arithmetic mod r for the closed form sum(s_i*k_i)*G.
 *
 * Built a second time with -DORACLE_GRUMPKIN (liboracle_grumpkin.so): the
 * reference's Grumpkin CPU Pippenger (grumpkin_ops.c:528-560) is the same
 * algorithm over the field r with group order p; only the constants differ.
 * The GLV routines are BN254-only and are left out of that build.
 *
 * The last block (BN254 builds only) is a G2 Pippenger over Fp2 = Fp[u] /
 * (u^2 + 1) with the reference's G2 point formulas (BN254Pairing.swift:
 * 45-70, 314-359), the G2 parity checker at sizes the Python oracle cannot
 * reach and the timed G2 CPU baseline.
 *
 * Parity pins: see oracle/bn254.py's header (EIP-196 vectors, closed form,
 * brute force, algebraic invariants) and tests/test_oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

typedef unsigned __int128 u128;

#ifndef ORACLE_GRUMPKIN
/* p, -p^-1 mod 2^64, R mod p (bn254_msm.c:24-32) */
static const uint64_t Q[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL,
                              0xb85045b68181585dULL, 0x30644e72e131a029ULL};
static const uint64_t QINV = 0x87d20782e4866389ULL;
static const uint64_t MONE[4] = {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL,
                                 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL};
static const uint64_t RR[4] = {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL,
                               0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL};
/* group order r (bn254_msm.c:704-707, GLV.swift:8-11) */
static const uint64_t RORD[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL,
                                 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
#else
/* Grumpkin (y^2 = x^3 - 17 over BN254's r): the same Pippenger, point
 * formulas and window rule as grumpkin_pippenger_msm
 * (Sources/NeonFieldOps/grumpkin_ops.c:129-186, 304-377, 434-560) with the
 * base field r: r, -r^-1 mod 2^64, R mod r (grumpkin_ops.c:17-25), R^2 mod r. */
static const uint64_t Q[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL,
                              0xb85045b68181585dULL, 0x30644e72e131a029ULL};
static const uint64_t QINV = 0xc2e1f593efffffffULL;
static const uint64_t MONE[4] = {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL,
                                 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL};
static const uint64_t RR[4] = {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL,
                               0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL};
/* group order: BN254's base-field prime p */
static const uint64_t RORD[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL,
                                 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
#endif

/* ---------------- Fp ---------------- */

/* t (5 words, value < 2p) -> t mod p in out; keeps t when t < p */
static void f_final_sub(const uint64_t t[5], uint64_t out[4]) {
    uint64_t d[4];
    uint64_t borrow = 0;
#pragma GCC unroll 4
    for (int j = 0; j < 4; j++) {
        u128 s = (u128)t[j] - Q[j] - borrow;
        d[j] = (uint64_t)s;
        borrow = (uint64_t)(s >> 127) & 1;
    }
    const int keep = !t[4] && borrow;
#pragma GCC unroll 4
    for (int j = 0; j < 4; j++) out[j] = keep ? t[j] : d[j];
}

/* Montgomery product a*b/2^256 mod p, CIOS with the row loops fully unrolled
 * (bn254_msm.c:38-154 unrolls the same four rounds by hand).  Round i adds
 * a*b_i to the running value and then m*p with m = t_0 * (-p^-1) mod 2^64,
 * shifting one word down; the running value stays below 2p (p < 2^254), so
 * one spare word t[4] holds every carry. */
static void f_mul(const uint64_t a[4], const uint64_t b[4],
                                                        uint64_t out[4]) {
    uint64_t t[5] = {0, 0, 0, 0, 0};
#pragma GCC unroll 4
    for (int i = 0; i < 4; i++) {
        const uint64_t bi = b[i];
        u128 acc = (u128)a[0] * bi + t[0];
        t[0] = (uint64_t)acc;
#pragma GCC unroll 3
        for (int j = 1; j < 4; j++) {
            acc = (u128)a[j] * bi + t[j] + (uint64_t)(acc >> 64);
            t[j] = (uint64_t)acc;
        }
        const uint64_t hi = t[4] + (uint64_t)(acc >> 64);
        const uint64_t m = t[0] * QINV;
        acc = (u128)m * Q[0] + t[0];
#pragma GCC unroll 3
        for (int j = 1; j < 4; j++) {
            acc = (u128)m * Q[j] + t[j] + (uint64_t)(acc >> 64);
            t[j - 1] = (uint64_t)acc;
        }
        acc = (u128)hi + (uint64_t)(acc >> 64);
        t[3] = (uint64_t)acc;
        t[4] = (uint64_t)(acc >> 64);
    }
    f_final_sub(t, out);
}

/* Montgomery square a^2/2^256 mod p with its own product schedule (the
 * reference has a dedicated squaring too, bn254_msm.c:156-267): the six
 * cross products a_i a_j (i < j) once, the 512-bit sum doubled by a one-bit
 * shift, the four squares a_i^2 added on the diagonal (10 word products
 * instead of 16), then four word-by-word Montgomery reductions of the
 * 8-word value. */
static void f_sqr(const uint64_t a[4], uint64_t out[4]) {
    uint64_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    /* cross products, one row per i: w[i+j] += a_i a_j for j > i */
#pragma GCC unroll 3
    for (int i = 0; i < 3; i++) {
        uint64_t c = 0;
#pragma GCC unroll 3
        for (int j = i + 1; j < 4; j++) {
            const u128 x = (u128)a[i] * a[j] + w[i + j] + c;
            w[i + j] = (uint64_t)x;
            c = (uint64_t)(x >> 64);
        }
        w[i + 4] = c;
    }
    /* double */
    w[7] = w[6] >> 63;
#pragma GCC unroll 5
    for (int k = 6; k > 1; k--) w[k] = (w[k] << 1) | (w[k - 1] >> 63);
    w[1] <<= 1;
    /* diagonal */
    uint64_t c = 0;
#pragma GCC unroll 4
    for (int i = 0; i < 4; i++) {
        const u128 sq = (u128)a[i] * a[i];
        u128 x = (u128)w[2 * i] + (uint64_t)sq + c;
        w[2 * i] = (uint64_t)x;
        x = (u128)w[2 * i + 1] + (uint64_t)(sq >> 64) + (uint64_t)(x >> 64);
        w[2 * i + 1] = (uint64_t)x;
        c = (uint64_t)(x >> 64);
    }
    (void)c;  /* 0: a^2 < p^2 < 2^508 */
    /* reduce: w[i..i+3] += m_i p (the low word becomes zero); the round's
     * carry goes into w[i+4], whose own overflow bit enters the next round's
     * last word */
    uint64_t extra = 0;
#pragma GCC unroll 4
    for (int i = 0; i < 4; i++) {
        const uint64_t m = w[i] * QINV;
        u128 x = (u128)m * Q[0] + w[i];
#pragma GCC unroll 3
        for (int j = 1; j < 4; j++) {
            x = (u128)m * Q[j] + w[i + j] + (uint64_t)(x >> 64);
            w[i + j] = (uint64_t)x;
        }
        x = (u128)w[i + 4] + (uint64_t)(x >> 64) + extra;
        w[i + 4] = (uint64_t)x;
        extra = (uint64_t)(x >> 64);
    }
    uint64_t t[5] = {w[4], w[5], w[6], w[7], extra};
    f_final_sub(t, out);
}

static int f_is_zero(const uint64_t a[4]) { return (a[0] | a[1] | a[2] | a[3]) == 0; }

static void f_add(const uint64_t a[4], const uint64_t b[4], uint64_t out[4]) {
    uint64_t s[4], d[4];
    uint64_t c = 0;
    for (int j = 0; j < 4; j++) {
        u128 x = (u128)a[j] + b[j] + c;
        s[j] = (uint64_t)x;
        c = (uint64_t)(x >> 64);
    }
    uint64_t borrow = 0;
    for (int j = 0; j < 4; j++) {
        u128 x = (u128)s[j] - Q[j] - borrow;
        d[j] = (uint64_t)x;
        borrow = (uint64_t)(x >> 127) & 1;
    }
    if (c || !borrow) memcpy(out, d, 32);
    else memcpy(out, s, 32);
}

static void f_sub(const uint64_t a[4], const uint64_t b[4], uint64_t out[4]) {
    uint64_t d[4];
    uint64_t borrow = 0;
    for (int j = 0; j < 4; j++) {
        u128 x = (u128)a[j] - b[j] - borrow;
        d[j] = (uint64_t)x;
        borrow = (uint64_t)(x >> 127) & 1;
    }
    if (borrow) {
        uint64_t c = 0;
        for (int j = 0; j < 4; j++) {
            u128 x = (u128)d[j] + Q[j] + c;
            d[j] = (uint64_t)x;
            c = (uint64_t)(x >> 64);
        }
    }
    memcpy(out, d, 32);
}

static void f_dbl(const uint64_t a[4], uint64_t out[4]) { f_add(a, a, out); }

/* Fermat inverse: a^(p-2) by square-and-multiply over the exponent bits */
static void f_inv(const uint64_t a[4], uint64_t out[4]) {
    uint64_t e[4] = {Q[0] - 2, Q[1], Q[2], Q[3]};
    uint64_t acc[4], base[4];
    memcpy(acc, MONE, 32);
    memcpy(base, a, 32);
    for (int w = 0; w < 4; w++)
        for (int b = 0; b < 64; b++) {
            if ((e[w] >> b) & 1) f_mul(acc, base, acc);
            f_sqr(base, base);
        }
    memcpy(out, acc, 32);
}

/* ---------------- Jacobian points: 12 limbs (X, Y, Z) ---------------- */

static void j_set_id(uint64_t p[12]) {
    memcpy(p, MONE, 32);
    memcpy(p + 4, MONE, 32);
    memset(p + 8, 0, 32);
}
static int j_is_id(const uint64_t p[12]) { return f_is_zero(p + 8); }
static int a_is_id(const uint64_t q[8]) {
    uint64_t acc = 0;
    for (int i = 0; i < 8; i++) acc |= q[i];
    return acc == 0;
}

/* doubling on y^2 = x^3 + 3 (a = 0) */
static void j_dbl(const uint64_t p[12], uint64_t r[12]) {
    if (j_is_id(p)) { memcpy(r, p, 96); return; }
    uint64_t xx[4], yy[4], yyyy[4], d[4], e[4], f[4], t[4], u[4], out[12];
    f_sqr(p, xx);
    f_sqr(p + 4, yy);
    f_sqr(yy, yyyy);
    f_add(p, yy, t);           /* d = 2((x+yy)^2 - xx - yyyy) */
    f_sqr(t, t);
    f_sub(t, xx, t);
    f_sub(t, yyyy, t);
    f_dbl(t, d);
    f_dbl(xx, t);              /* e = 3 xx */
    f_add(t, xx, e);
    f_sqr(e, f);               /* f = e^2 */
    f_dbl(d, t);               /* X3 = f - 2d */
    f_sub(f, t, out);
    f_sub(d, out, t);          /* Y3 = e (d - X3) - 8 yyyy */
    f_mul(e, t, u);
    f_dbl(yyyy, t);
    f_dbl(t, t);
    f_dbl(t, t);
    f_sub(u, t, out + 4);
    f_add(p + 4, p + 8, t);    /* Z3 = (y+z)^2 - yy - zz */
    f_sqr(t, t);
    f_sub(t, yy, t);
    f_sqr(p + 8, u);
    f_sub(t, u, out + 8);
    memcpy(r, out, 96);
}

static void j_add(const uint64_t p[12], const uint64_t q[12], uint64_t r[12]) {
    if (j_is_id(p)) { memcpy(r, q, 96); return; }
    if (j_is_id(q)) { memcpy(r, p, 96); return; }
    uint64_t z1z1[4], z2z2[4], u1[4], u2[4], s1[4], s2[4], h[4], rr[4], ii[4], jj[4],
        v[4], t[4], out[12];
    f_sqr(p + 8, z1z1);
    f_sqr(q + 8, z2z2);
    f_mul(p, z2z2, u1);
    f_mul(q, z1z1, u2);
    f_mul(q + 8, z2z2, t);
    f_mul(p + 4, t, s1);
    f_mul(p + 8, z1z1, t);
    f_mul(q + 4, t, s2);
    f_sub(u2, u1, h);
    f_sub(s2, s1, t);
    f_dbl(t, rr);
    if (f_is_zero(h)) {
        if (f_is_zero(rr)) { j_dbl(p, r); return; }
        j_set_id(r);
        return;
    }
    f_dbl(h, t);
    f_sqr(t, ii);
    f_mul(h, ii, jj);
    f_mul(u1, ii, v);
    f_sqr(rr, out);            /* X3 = rr^2 - jj - 2v */
    f_sub(out, jj, out);
    f_dbl(v, t);
    f_sub(out, t, out);
    f_sub(v, out, t);          /* Y3 = rr (v - X3) - 2 s1 jj */
    f_mul(rr, t, out + 4);
    f_mul(s1, jj, t);
    f_dbl(t, t);
    f_sub(out + 4, t, out + 4);
    f_add(p + 8, q + 8, t);    /* Z3 = ((z1+z2)^2 - z1z1 - z2z2) h */
    f_sqr(t, t);
    f_sub(t, z1z1, t);
    f_sub(t, z2z2, t);
    f_mul(t, h, out + 8);
    memcpy(r, out, 96);
}

/* Jacobian + affine; affine (0,0) is the identity */
static void j_add_aff(const uint64_t p[12], const uint64_t q[8], uint64_t r[12]) {
    if (a_is_id(q)) { memcpy(r, p, 96); return; }
    if (j_is_id(p)) {
        memcpy(r, q, 64);
        memcpy(r + 8, MONE, 32);
        return;
    }
    uint64_t z1z1[4], u2[4], s2[4], h[4], rr[4], ii[4], jj[4], v[4], t[4], hh[4], out[12];
    f_sqr(p + 8, z1z1);
    f_mul(q, z1z1, u2);
    f_mul(p + 8, z1z1, t);
    f_mul(q + 4, t, s2);
    f_sub(u2, p, h);
    f_sub(s2, p + 4, t);
    f_dbl(t, rr);
    if (f_is_zero(h)) {
        if (f_is_zero(rr)) { j_dbl(p, r); return; }
        j_set_id(r);
        return;
    }
    f_dbl(h, t);
    f_sqr(t, ii);
    f_mul(h, ii, jj);
    f_mul(p, ii, v);
    f_sqr(rr, out);
    f_sub(out, jj, out);
    f_dbl(v, t);
    f_sub(out, t, out);
    f_sub(v, out, t);
    f_mul(rr, t, out + 4);
    f_mul(p + 4, jj, t);
    f_dbl(t, t);
    f_sub(out + 4, t, out + 4);
    f_add(p + 8, h, t);        /* Z3 = (z1 + h)^2 - z1z1 - hh */
    f_sqr(t, t);
    f_sub(t, z1z1, t);
    f_sqr(h, hh);
    f_sub(t, hh, out + 8);
    memcpy(r, out, 96);
}

/* n Jacobian points -> affine with one inversion; identity -> (0,0) */
static void batch_affine(const uint64_t *proj, uint64_t *aff, int n) {
    if (n <= 0) return;
    uint64_t *pre = (uint64_t *)malloc((size_t)n * 32);
    int first = -1;
    for (int i = 0; i < n; i++) {
        const uint64_t *z = proj + (size_t)i * 12 + 8;
        if (j_is_id(proj + (size_t)i * 12)) {
            if (i == 0) memcpy(pre, MONE, 32);
            else memcpy(pre + (size_t)i * 4, pre + (size_t)(i - 1) * 4, 32);
        } else if (first < 0) {
            first = i;
            memcpy(pre + (size_t)i * 4, z, 32);
        } else {
            f_mul(pre + (size_t)(i - 1) * 4, z, pre + (size_t)i * 4);
        }
    }
    if (first < 0) {
        memset(aff, 0, (size_t)n * 64);
        free(pre);
        return;
    }
    uint64_t inv[4];
    f_inv(pre + (size_t)(n - 1) * 4, inv);
    for (int i = n - 1; i >= 0; i--) {
        const uint64_t *pt = proj + (size_t)i * 12;
        uint64_t *o = aff + (size_t)i * 8;
        if (j_is_id(pt)) { memset(o, 0, 64); continue; }
        uint64_t zi[4], zi2[4], zi3[4];
        if (i > first) {
            f_mul(inv, pre + (size_t)(i - 1) * 4, zi);
            f_mul(inv, pt + 8, inv);
        } else {
            memcpy(zi, inv, 32);
        }
        f_sqr(zi, zi2);
        f_mul(zi2, zi, zi3);
        f_mul(pt, zi2, o);
        f_mul(pt + 4, zi3, o + 4);
    }
    free(pre);
}

/* ---------------- Pippenger ---------------- */

static uint32_t window_digit(const uint32_t *s, int w, int c) {
    int bit = w * c, word = bit >> 5, sh = bit & 31;
    uint64_t v = s[word];
    if (word + 1 < 8) v |= (uint64_t)s[word + 1] << 32;
    return (uint32_t)((v >> sh) & ((1u << c) - 1u));
}

int oracle_window_bits(int n) {
    static const int lim[8] = {4, 32, 256, 2048, 8192, 32768, 131072, 524288};
    static const int bits[8] = {3, 5, 8, 10, 11, 13, 14, 15};
    for (int i = 0; i < 8; i++)
        if (n <= lim[i]) return bits[i];
    return 16;
}

typedef struct {
    const uint64_t *pts;
    const uint32_t *sc;
    int n, c, nwin, stride, first;
    uint64_t *sums;            /* nwin x 12 */
} job_t;

static void one_window(const uint64_t *pts, const uint32_t *sc, int n, int c, int w,
                       uint64_t sum[12]) {
    if (c < 1 || c > 24) return;
    int nb = (1 << c) - 1;
    uint64_t *bk = (uint64_t *)malloc((size_t)(nb + 1) * 96);
    for (int b = 0; b <= nb; b++) j_set_id(bk + (size_t)b * 12);
    for (int i = 0; i < n; i++) {
        uint32_t d = window_digit(sc + (size_t)i * 8, w, c);
        if (d) j_add_aff(bk + (size_t)d * 12, pts + (size_t)i * 8, bk + (size_t)d * 12);
    }
    uint64_t *aff = (uint64_t *)malloc((size_t)nb * 64);
    batch_affine(bk + 12, aff, nb);
    uint64_t run[12], acc[12];
    j_set_id(run);
    j_set_id(acc);
    for (int j = nb - 1; j >= 0; j--) {
        if (!a_is_id(aff + (size_t)j * 8)) j_add_aff(run, aff + (size_t)j * 8, run);
        j_add(acc, run, acc);
    }
    memcpy(sum, acc, 96);
    free(bk);
    free(aff);
}

static void *window_thread(void *arg) {
    job_t *jb = (job_t *)arg;
    for (int w = jb->first; w < jb->nwin; w += jb->stride)
        one_window(jb->pts, jb->sc, jb->n, jb->c, w, jb->sums + (size_t)w * 12);
    return NULL;
}

/* MSM over n affine Montgomery points (8 u64 each) and n standard-form
 * 256-bit scalars (8 u32 each).  window_bits <= 0 selects
 * oracle_window_bits(n); threads <= 0 runs one thread per window.
 * result: Jacobian Montgomery (X, Y, Z), identity = (1, 1, 0). */
void oracle_pippenger_msm(const uint64_t *points, const uint32_t *scalars, int n,
                          int window_bits, int threads, uint64_t result[12]) {
    if (n <= 0) { j_set_id(result); return; }
    int c = window_bits > 0 ? window_bits : oracle_window_bits(n);
    int nwin = (256 + c - 1) / c;
    uint64_t *sums = (uint64_t *)malloc((size_t)nwin * 96);
    int nt = threads > 0 ? threads : nwin;
    if (nt > nwin) nt = nwin;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    job_t *jobs = (job_t *)malloc(sizeof(job_t) * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        job_t j = {points, scalars, n, c, nwin, nt, t, sums};
        jobs[t] = j;
    }
    for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, window_thread, &jobs[t]);
    window_thread(&jobs[0]);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
    uint64_t acc[12];
    memcpy(acc, sums + (size_t)(nwin - 1) * 12, 96);
    for (int w = nwin - 2; w >= 0; w--) {
        for (int s = 0; s < c; s++) j_dbl(acc, acc);
        j_add(acc, sums + (size_t)w * 12, acc);
    }
    memcpy(result, acc, 96);
    free(sums);
    free(th);
    free(jobs);
}

/* Jacobian -> affine (bn254_msm.c:1501-1514): identity -> (0,0) */
void oracle_jacobian_to_affine(const uint64_t p[12], uint64_t aff[8]) {
    if (j_is_id(p)) { memset(aff, 0, 64); return; }
    uint64_t zi[4], zi2[4], zi3[4];
    f_inv(p + 8, zi);
    f_sqr(zi, zi2);
    f_mul(zi2, zi, zi3);
    f_mul(p, zi2, aff);
    f_mul(p + 4, zi3, aff + 4);
}

/* ---------------- helpers used by tests ---------------- */

void oracle_fp_mul(const uint64_t a[4], const uint64_t b[4], uint64_t r[4]) { f_mul(a, b, r); }
void oracle_fp_sqr(const uint64_t a[4], uint64_t r[4]) { f_sqr(a, r); }
void oracle_fp_add(const uint64_t a[4], const uint64_t b[4], uint64_t r[4]) { f_add(a, b, r); }
void oracle_fp_sub(const uint64_t a[4], const uint64_t b[4], uint64_t r[4]) { f_sub(a, b, r); }
void oracle_fp_inv(const uint64_t a[4], uint64_t r[4]) { f_inv(a, r); }
void oracle_to_mont(const uint64_t a[4], uint64_t r[4]) { f_mul(a, RR, r); }
void oracle_from_mont(const uint64_t a[4], uint64_t r[4]) {
    static const uint64_t one[4] = {1, 0, 0, 0};
    f_mul(a, one, r);
}
void oracle_jac_add(const uint64_t p[12], const uint64_t q[12], uint64_t r[12]) { j_add(p, q, r); }
void oracle_jac_dbl(const uint64_t p[12], uint64_t r[12]) { j_dbl(p, r); }
void oracle_jac_add_affine(const uint64_t p[12], const uint64_t q[8], uint64_t r[12]) {
    j_add_aff(p, q, r);
}

/* k*P for an affine P by MSB-first double-and-add; result Jacobian */
void oracle_scalar_mul(const uint64_t aff[8], const uint32_t k[8], uint64_t out[12]) {
    uint64_t acc[12];
    j_set_id(acc);
    for (int bit = 255; bit >= 0; bit--) {
        j_dbl(acc, acc);
        if ((k[bit >> 5] >> (bit & 31)) & 1) j_add_aff(acc, aff, acc);
    }
    memcpy(out, acc, 96);
}

/* Scalars mod r, standard (non-Montgomery) 256-bit integers as 4 u64. */
static int geq4(const uint64_t a[4], const uint64_t b[4]) {
    for (int i = 3; i >= 0; i--) {
        if (a[i] != b[i]) return a[i] > b[i];
    }
    return 1;
}
static uint64_t sub4(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) {
        u128 x = (u128)a[i] - b[i] - br;
        r[i] = (uint64_t)x;
        br = (uint64_t)(x >> 127) & 1;
    }
    return br;
}
static uint64_t add4(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) {
        u128 x = (u128)a[i] + b[i] + c;
        r[i] = (uint64_t)x;
        c = (uint64_t)(x >> 64);
    }
    return c;
}
static void load_s(const uint32_t *s, uint64_t v[4]) {
    for (int i = 0; i < 4; i++) v[i] = (uint64_t)s[2 * i] | ((uint64_t)s[2 * i + 1] << 32);
}
static void store_s(const uint64_t v[4], uint32_t *s) {
    for (int i = 0; i < 4; i++) {
        s[2 * i] = (uint32_t)v[i];
        s[2 * i + 1] = (uint32_t)(v[i] >> 32);
    }
}
/* sum_i s_i * k_i mod r, for the closed form of MSMs over bases k_i*G: the
 * 512-bit products summed exactly in 9 words (n < 2^31 adds at most 31
 * bits), then one reduction of the sum, bit by bit from the top. */
void oracle_scalar_dot(const uint32_t *s, const uint32_t *k, int n, uint32_t out[8]) {
    uint64_t sum[9] = {0};
    for (int i = 0; i < n; i++) {
        uint64_t a[4], b[4];
        load_s(s + (size_t)i * 8, a);
        load_s(k + (size_t)i * 8, b);
        for (int x = 0; x < 4; x++) {
            uint64_t carry = 0;
            for (int y = 0; y < 4; y++) {
                u128 t = (u128)a[x] * b[y] + sum[x + y] + carry;
                sum[x + y] = (uint64_t)t;
                carry = (uint64_t)(t >> 64);
            }
            for (int z = x + 4; carry && z < 9; z++) {
                u128 t = (u128)sum[z] + carry;
                sum[z] = (uint64_t)t;
                carry = (uint64_t)(t >> 64);
            }
        }
    }
    uint64_t acc[4] = {0, 0, 0, 0};
    for (int bit = 9 * 64 - 1; bit >= 0; bit--) {
        uint64_t t[4], one[4] = {(sum[bit >> 6] >> (bit & 63)) & 1, 0, 0, 0};
        uint64_t c = add4(t, acc, acc);
        if (c || geq4(t, RORD)) sub4(t, t, RORD);
        c = add4(acc, t, one);
        if (c || geq4(acc, RORD)) sub4(acc, acc, RORD);
    }
    store_s(acc, out);
}

#ifndef ORACLE_GRUMPKIN
/* GLV split (GLV.swift:37-92): k mod r -> (k1, k2, neg1, neg2) with
 * k == (neg1 ? -k1 : k1) + lambda * (neg2 ? -k2 : k2)  (mod r). */
static const uint64_t G1C[3] = {0x7a7bd9d4391eb18dULL, 0x4ccef014a773d2cfULL, 0x2ULL};
static const uint64_t G2C[2] = {0xd91d232ec7e0b3d7ULL, 0x2ULL};
static const uint64_t A1C = 0x89d3256894d213e3ULL;
static const uint64_t A2C[2] = {0x0be4e1541221250bULL, 0x6f4d8248eeb859fdULL};
static const uint64_t MB1C[2] = {0x8211bbeb7d4f1128ULL, 0x6f4d8248eeb859fcULL};
static const uint64_t B2C = 0x89d3256894d213e3ULL;

/* schoolbook product of na x nb words into out (na+nb words) */
static void mulw(const uint64_t *a, int na, const uint64_t *b, int nb, uint64_t *out) {
    memset(out, 0, sizeof(uint64_t) * (size_t)(na + nb));
    for (int i = 0; i < na; i++) {
        uint64_t carry = 0;
        for (int j = 0; j < nb; j++) {
            u128 x = (u128)a[i] * b[j] + out[i + j] + carry;
            out[i + j] = (uint64_t)x;
            carry = (uint64_t)(x >> 64);
        }
        out[i + nb] = carry;
    }
}

void oracle_glv_decompose(const uint32_t kin[8], uint32_t k1o[8], uint32_t k2o[8], int *neg1,
                          int *neg2) {
    uint64_t k[4];
    load_s(kin, k);
    while (geq4(k, RORD)) sub4(k, k, RORD);
    uint64_t p7[7], p6[6];
    mulw(k, 4, G1C, 3, p7);
    mulw(k, 4, G2C, 2, p6);
    uint64_t c1[2] = {p7[4], p7[5]};  /* floor(k*g1 / 2^256) */
    uint64_t c2 = p6[4];              /* floor(k*g2 / 2^256) */
    uint64_t t3[3], t4[4];
    /* k1 = k - c2*a1 - c1*a2  (mod r) */
    uint64_t c2a1[4] = {0, 0, 0, 0};
    mulw(&c2, 1, &A1C, 1, t3);
    c2a1[0] = t3[0];
    c2a1[1] = t3[1];
    uint64_t k1[4];
    if (sub4(k1, k, c2a1)) add4(k1, k1, RORD);
    mulw(c1, 2, A2C, 2, t4);
    if (sub4(k1, k1, t4)) add4(k1, k1, RORD);
    /* k2 = c2*(-b1) - c1*b2 as a signed integer */
    uint64_t x[4] = {0, 0, 0, 0}, y[4] = {0, 0, 0, 0}, k2[4];
    mulw(&c2, 1, MB1C, 2, t3);
    memcpy(x, t3, 24);
    mulw(c1, 2, &B2C, 1, t3);
    memcpy(y, t3, 24);
    int n2 = 0;
    if (sub4(k2, x, y)) {
        uint64_t z[4] = {0, 0, 0, 0};
        sub4(k2, z, k2);
        n2 = 1;
    }
    uint64_t half[4] = {(RORD[0] >> 1) | (RORD[1] << 63), (RORD[1] >> 1) | (RORD[2] << 63),
                        (RORD[2] >> 1) | (RORD[3] << 63), RORD[3] >> 1};
    int n1 = 0;
    if (geq4(k1, half)) {
        sub4(k1, RORD, k1);
        n1 = 1;
    }
    store_s(k1, k1o);
    store_s(k2, k2o);
    *neg1 = n1;
    *neg2 = n2;
}

#endif /* ORACLE_GRUMPKIN */

/* Signed c-bit digits (msm_kernels.metal:399-439): nwin windows, digit d in
 * (-2^(c-1), 2^(c-1)], returned as signed ints; *carry_out is the carry
 * left after the last window (zero whenever k < 2^(c*nwin - 1)). */
void oracle_signed_digits(const uint32_t k[8], int c, int nwin, int32_t *digits,
                          int *carry_out) {
    uint32_t mask = (1u << c) - 1u, half = 1u << (c - 1), full = 1u << c, carry = 0;
    for (int w = 0; w < nwin; w++) {
        int bit = w * c, word = bit >> 5, sh = bit & 31;
        uint32_t v = 0;
        if (word < 8) {
            uint64_t x = k[word];
            if (word + 1 < 8) x |= (uint64_t)k[word + 1] << 32;
            v = (uint32_t)(x >> sh) & mask;
        }
        uint32_t d = v + carry;
        carry = 0;
        if (d > half) {
            digits[w] = -(int32_t)(full - d);
            carry = 1;
        } else {
            digits[w] = (int32_t)d;
        }
    }
    *carry_out = (int)carry;
}

#ifndef ORACLE_GRUMPKIN
/* ---------------- BN254 G2 ----------------
 *
 * Fp2 = Fp[u] / (u^2 + 1), an element c0 + c1 u as 8 limbs (c0 | c1), both
 * in the Fp Montgomery form above: Karatsuba multiply, (a0+a1)(a0-a1) square
 * and norm inverse as BN254Pairing.swift:45-70.  Points on the twist
 * y^2 = x^3 + 3/(9+u) in Jacobian (X, Y, Z) over Fp2, 24 limbs, identity
 * Z == 0 stored as (1, 1, 0): doubling and full addition as g2Double /
 * g2Add (BN254Pairing.swift:314-359), the same a = 0 formulas as the G1
 * code above, with the mixed addition derived from them (Z2 = 1).  Affine
 * bases are 16 limbs (x.c0, x.c1, y.c0, y.c1), all-zero = the identity.
 *
 * The reference's CPU G2 MSM (BN254G2MSMEngine.swift:542-560,
 * Groth16Prover.swift:374-402) is bit-serial Straus, about 127 n additions
 * for n bases; the MSM here is the G1 Pippenger above (unsigned c-bit
 * windows over 256 bits, buckets by mixed addition, one inversion per
 * window, running sums, Horner) over G2, which is what a CPU baseline at
 * n = 2^22 can run.  Scalars are not reduced mod r: the bases lie in the
 * subgroup of order r, so s P = (s mod r) P. */

static int f2_is_zero(const uint64_t a[8]) { return f_is_zero(a) && f_is_zero(a + 4); }
static void f2_add(const uint64_t a[8], const uint64_t b[8], uint64_t r[8]) {
    f_add(a, b, r);
    f_add(a + 4, b + 4, r + 4);
}
static void f2_sub(const uint64_t a[8], const uint64_t b[8], uint64_t r[8]) {
    f_sub(a, b, r);
    f_sub(a + 4, b + 4, r + 4);
}
static void f2_dbl(const uint64_t a[8], uint64_t r[8]) { f2_add(a, a, r); }
static void f2_mul(const uint64_t a[8], const uint64_t b[8], uint64_t r[8]) {
    uint64_t t0[4], t1[4], sa[4], sb[4], m[4];
    f_mul(a, b, t0);
    f_mul(a + 4, b + 4, t1);
    f_add(a, a + 4, sa);
    f_add(b, b + 4, sb);
    f_mul(sa, sb, m);
    f_sub(t0, t1, r);
    f_add(t0, t1, t0);
    f_sub(m, t0, r + 4);
}
static void f2_sqr(const uint64_t a[8], uint64_t r[8]) {
    uint64_t t[4], s[4], d[4];
    f_mul(a, a + 4, t);
    f_add(a, a + 4, s);
    f_sub(a, a + 4, d);
    f_mul(s, d, r);
    f_dbl(t, r + 4);
}
static void f2_inv(const uint64_t a[8], uint64_t r[8]) {
    static const uint64_t zero[4] = {0, 0, 0, 0};
    uint64_t n0[4], n1[4], ni[4], t[4];
    f_sqr(a, n0);
    f_sqr(a + 4, n1);
    f_add(n0, n1, n0);
    f_inv(n0, ni);
    f_mul(a, ni, r);
    f_mul(a + 4, ni, t);
    f_sub(zero, t, r + 4);
}

static void g2_set_id(uint64_t p[24]) {
    memset(p, 0, 192);
    memcpy(p, MONE, 32);
    memcpy(p + 8, MONE, 32);
}
static int g2_is_id(const uint64_t p[24]) { return f2_is_zero(p + 16); }
static int g2a_is_id(const uint64_t q[16]) {
    uint64_t acc = 0;
    for (int i = 0; i < 16; i++) acc |= q[i];
    return acc == 0;
}

static void g2_dbl(const uint64_t p[24], uint64_t r[24]) {
    if (g2_is_id(p)) { memcpy(r, p, 192); return; }
    uint64_t a[8], b[8], c[8], d[8], e[8], f[8], t[8], u[8], out[24];
    f2_sqr(p, a);
    f2_sqr(p + 8, b);
    f2_sqr(b, c);
    f2_add(p, b, t);           /* d = 2((x+b)^2 - a - c) */
    f2_sqr(t, t);
    f2_sub(t, a, t);
    f2_sub(t, c, t);
    f2_dbl(t, d);
    f2_dbl(a, t);              /* e = 3a */
    f2_add(t, a, e);
    f2_sqr(e, f);
    f2_dbl(d, t);              /* X3 = f - 2d */
    f2_sub(f, t, out);
    f2_sub(d, out, t);         /* Y3 = e (d - X3) - 8c */
    f2_mul(e, t, u);
    f2_dbl(c, t);
    f2_dbl(t, t);
    f2_dbl(t, t);
    f2_sub(u, t, out + 8);
    f2_add(p + 8, p + 16, t);  /* Z3 = (y+z)^2 - b - z^2 */
    f2_sqr(t, t);
    f2_sub(t, b, t);
    f2_sqr(p + 16, u);
    f2_sub(t, u, out + 16);
    memcpy(r, out, 192);
}

static void g2_add(const uint64_t p[24], const uint64_t q[24], uint64_t r[24]) {
    if (g2_is_id(p)) { memcpy(r, q, 192); return; }
    if (g2_is_id(q)) { memcpy(r, p, 192); return; }
    uint64_t z1z1[8], z2z2[8], u1[8], u2[8], s1[8], s2[8], h[8], rr[8], ii[8], jj[8],
        v[8], t[8], out[24];
    f2_sqr(p + 16, z1z1);
    f2_sqr(q + 16, z2z2);
    f2_mul(p, z2z2, u1);
    f2_mul(q, z1z1, u2);
    f2_mul(q + 16, z2z2, t);
    f2_mul(p + 8, t, s1);
    f2_mul(p + 16, z1z1, t);
    f2_mul(q + 8, t, s2);
    f2_sub(u2, u1, h);
    f2_sub(s2, s1, t);
    f2_dbl(t, rr);
    if (f2_is_zero(h)) {
        if (f2_is_zero(rr)) { g2_dbl(p, r); return; }
        g2_set_id(r);
        return;
    }
    f2_dbl(h, t);
    f2_sqr(t, ii);
    f2_mul(h, ii, jj);
    f2_mul(u1, ii, v);
    f2_sqr(rr, out);           /* X3 = rr^2 - jj - 2v */
    f2_sub(out, jj, out);
    f2_dbl(v, t);
    f2_sub(out, t, out);
    f2_sub(v, out, t);         /* Y3 = rr (v - X3) - 2 s1 jj */
    f2_mul(rr, t, out + 8);
    f2_mul(s1, jj, t);
    f2_dbl(t, t);
    f2_sub(out + 8, t, out + 8);
    f2_add(p + 16, q + 16, t); /* Z3 = ((z1+z2)^2 - z1z1 - z2z2) h */
    f2_sqr(t, t);
    f2_sub(t, z1z1, t);
    f2_sub(t, z2z2, t);
    f2_mul(t, h, out + 16);
    memcpy(r, out, 192);
}

/* Jacobian + affine: g2_add with Z2 = 1 (z2z2 = 1, u1 = X1, s1 = Y1) */
static void g2_add_aff(const uint64_t p[24], const uint64_t q[16], uint64_t r[24]) {
    if (g2a_is_id(q)) { memcpy(r, p, 192); return; }
    if (g2_is_id(p)) {
        memset(r, 0, 192);
        memcpy(r, q, 128);
        memcpy(r + 16, MONE, 32);
        return;
    }
    uint64_t z1z1[8], u2[8], s2[8], h[8], rr[8], ii[8], jj[8], v[8], t[8], hh[8], out[24];
    f2_sqr(p + 16, z1z1);
    f2_mul(q, z1z1, u2);
    f2_mul(p + 16, z1z1, t);
    f2_mul(q + 8, t, s2);
    f2_sub(u2, p, h);
    f2_sub(s2, p + 8, t);
    f2_dbl(t, rr);
    if (f2_is_zero(h)) {
        if (f2_is_zero(rr)) { g2_dbl(p, r); return; }
        g2_set_id(r);
        return;
    }
    f2_dbl(h, t);
    f2_sqr(t, ii);
    f2_mul(h, ii, jj);
    f2_mul(p, ii, v);
    f2_sqr(rr, out);
    f2_sub(out, jj, out);
    f2_dbl(v, t);
    f2_sub(out, t, out);
    f2_sub(v, out, t);
    f2_mul(rr, t, out + 8);
    f2_mul(p + 8, jj, t);
    f2_dbl(t, t);
    f2_sub(out + 8, t, out + 8);
    f2_add(p + 16, h, t);      /* Z3 = (z1 + h)^2 - z1z1 - h^2 = 2 z1 h */
    f2_sqr(t, t);
    f2_sub(t, z1z1, t);
    f2_sqr(h, hh);
    f2_sub(t, hh, out + 16);
    memcpy(r, out, 192);
}

/* n Jacobian G2 points -> affine with one Fp2 inversion; identity -> 0 */
static void g2_batch_affine(const uint64_t *proj, uint64_t *aff, int n) {
    if (n <= 0) return;
    uint64_t *pre = (uint64_t *)malloc((size_t)n * 64);
    uint64_t one2[8] = {0};
    memcpy(one2, MONE, 32);
    int first = -1;
    for (int i = 0; i < n; i++) {
        const uint64_t *pt = proj + (size_t)i * 24;
        if (g2_is_id(pt)) {
            if (i == 0) memcpy(pre, one2, 64);
            else memcpy(pre + (size_t)i * 8, pre + (size_t)(i - 1) * 8, 64);
        } else if (first < 0) {
            first = i;
            memcpy(pre + (size_t)i * 8, pt + 16, 64);
        } else {
            f2_mul(pre + (size_t)(i - 1) * 8, pt + 16, pre + (size_t)i * 8);
        }
    }
    if (first < 0) {
        memset(aff, 0, (size_t)n * 128);
        free(pre);
        return;
    }
    uint64_t inv[8];
    f2_inv(pre + (size_t)(n - 1) * 8, inv);
    for (int i = n - 1; i >= 0; i--) {
        const uint64_t *pt = proj + (size_t)i * 24;
        uint64_t *o = aff + (size_t)i * 16;
        if (g2_is_id(pt)) { memset(o, 0, 128); continue; }
        uint64_t zi[8], zi2[8], zi3[8];
        if (i > first) {
            f2_mul(inv, pre + (size_t)(i - 1) * 8, zi);
            f2_mul(inv, pt + 16, inv);
        } else {
            memcpy(zi, inv, 64);
        }
        f2_sqr(zi, zi2);
        f2_mul(zi2, zi, zi3);
        f2_mul(pt, zi2, o);
        f2_mul(pt + 8, zi3, o + 8);
    }
    free(pre);
}

static void g2_one_window(const uint64_t *pts, const uint32_t *sc, int n, int c, int w,
                          uint64_t sum[24]) {
    if (c < 1 || c > 24) return;
    int nb = (1 << c) - 1;
    uint64_t *bk = (uint64_t *)malloc((size_t)(nb + 1) * 192);
    for (int b = 0; b <= nb; b++) g2_set_id(bk + (size_t)b * 24);
    for (int i = 0; i < n; i++) {
        uint32_t d = window_digit(sc + (size_t)i * 8, w, c);
        if (d) g2_add_aff(bk + (size_t)d * 24, pts + (size_t)i * 16, bk + (size_t)d * 24);
    }
    uint64_t *aff = (uint64_t *)malloc((size_t)nb * 128);
    g2_batch_affine(bk + 24, aff, nb);
    uint64_t run[24], acc[24];
    g2_set_id(run);
    g2_set_id(acc);
    for (int j = nb - 1; j >= 0; j--) {
        if (!g2a_is_id(aff + (size_t)j * 16)) g2_add_aff(run, aff + (size_t)j * 16, run);
        g2_add(acc, run, acc);
    }
    memcpy(sum, acc, 192);
    free(bk);
    free(aff);
}

static void *g2_window_thread(void *arg) {
    job_t *jb = (job_t *)arg;
    for (int w = jb->first; w < jb->nwin; w += jb->stride)
        g2_one_window(jb->pts, jb->sc, jb->n, jb->c, w, jb->sums + (size_t)w * 24);
    return NULL;
}

/* G2 MSM over n affine bases (16 u64 each, layout above) and n standard-form
 * 256-bit scalars (8 u32 each); window_bits <= 0 selects
 * oracle_window_bits(n), threads <= 0 one thread per window.  result:
 * Jacobian (X, Y, Z), each c0 | c1 in Fp Montgomery form. */
void oracle_g2_pippenger_msm(const uint64_t *points, const uint32_t *scalars, int n,
                             int window_bits, int threads, uint64_t result[24]) {
    if (n <= 0) { g2_set_id(result); return; }
    int c = window_bits > 0 ? window_bits : oracle_window_bits(n);
    int nwin = (256 + c - 1) / c;
    uint64_t *sums = (uint64_t *)malloc((size_t)nwin * 192);
    int nt = threads > 0 ? threads : nwin;
    if (nt > nwin) nt = nwin;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    job_t *jobs = (job_t *)malloc(sizeof(job_t) * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        job_t j = {points, scalars, n, c, nwin, nt, t, sums};
        jobs[t] = j;
    }
    for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, g2_window_thread, &jobs[t]);
    g2_window_thread(&jobs[0]);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
    uint64_t acc[24];
    memcpy(acc, sums + (size_t)(nwin - 1) * 24, 192);
    for (int w = nwin - 2; w >= 0; w--) {
        for (int s = 0; s < c; s++) g2_dbl(acc, acc);
        g2_add(acc, sums + (size_t)w * 24, acc);
    }
    memcpy(result, acc, 192);
    free(sums);
    free(th);
    free(jobs);
}

void oracle_g2_jacobian_to_affine(const uint64_t p[24], uint64_t aff[16]) {
    if (g2_is_id(p)) { memset(aff, 0, 128); return; }
    uint64_t zi[8], zi2[8], zi3[8];
    f2_inv(p + 16, zi);
    f2_sqr(zi, zi2);
    f2_mul(zi2, zi, zi3);
    f2_mul(p, zi2, aff);
    f2_mul(p + 8, zi3, aff + 8);
}

void oracle_g2_jac_add(const uint64_t p[24], const uint64_t q[24], uint64_t r[24]) { g2_add(p, q, r); }
void oracle_g2_jac_dbl(const uint64_t p[24], uint64_t r[24]) { g2_dbl(p, r); }
void oracle_g2_jac_add_affine(const uint64_t p[24], const uint64_t q[16], uint64_t r[24]) {
    g2_add_aff(p, q, r);
}
#endif /* ORACLE_GRUMPKIN */
