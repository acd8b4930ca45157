/* Writes log_table.h: the 128-entry reduction table of the device log.
 *
 * The algorithm (Szabolcs Nagy, ARM optimized-routines 2018, glibc >= 2.28)
 * writes x = 2^k z with z in [0x1.6p-1, 0x1.6p0), split into 128 subintervals
 * of equal width in each binade (80 of width 2^-8 below 1, 48 of width 2^-7
 * above), and uses log(x) = k ln2 + log(c) + log1p(z/c - 1) with c near the
 * centre of z's subinterval.  Entry i holds invc = 1/c (a double) and
 * logc = log(c) rounded to double; c itself is 1/invc exactly, carried as
 * chi + clo = RN(c) + RN(c - RN(c)) where no fused multiply-add is available.
 * The published selection procedure tries the 2^29 doubles either side of
 * RN(1/centre) as invc and keeps those for which
 *   (1) 0x1.8p9 + logc is exact, i.e. logc is a multiple of 2^-43, so that
 *       k ln2hi + logc has no rounding error;
 *   (2) z - chi - clo rounds with an error below 2^-66 in r = (z - c)/c for
 *       every z of the subinterval (z - chi is exact; the subtraction of clo
 *       rounds at the ulp of the result, largest at the subinterval's ends);
 *   (3) the rounding error of logc = RN(log c) is below 2^-66;
 * and of those takes the one whose rounding error in (3) is smallest.
 *
 * The search evaluates -log(invc) * 2^43 for every candidate in 128-bit
 * fixed point (80 fractional bits): blocks of 2^16 consecutive candidates
 * start from a quad-precision logarithm and step by second-order differences
 * of the local Taylor expansion, whose neglected cubic term is below 2^-60
 * units.  Candidates within 2^-23 units of the 2^-43 grid (criteria 1 and 3)
 * are then checked in quad precision against all three criteria.  About
 * 2^37 candidates in all; half a minute on 8 cores.
 *
 * Build and run: gcc -O2 -fopenmp gen_log_table.c -lquadmath -o gen && ./gen > log_table.h
 * (./gen FIRST END prints only the table rows of entries FIRST..END-1.)
 */
#include <omp.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

static uint64_t bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static double from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

#define FRAC 80
#define HALF_RANGE (1L << 29)
#define BLOCK (1L << 16)
#define MAX_CAND 4096

/* x * 2^FRAC as a signed 128-bit integer, rounded to nearest */
static i128 to_fixed(__float128 x) {
    __float128 s = x * 0x1p80Q;
    __float128 fl = floorq(s);
    __float128 hi = floorq(fl * 0x1p-64Q);
    __float128 lo = fl - hi * 0x1p64Q;
    i128 v = ((i128)(int64_t)hi << 64) + (i128)(uint64_t)lo;
    if (s - fl >= 0.5Q) v += 1;
    return v;
}

/* criterion (2): the largest rounding error of (z - chi) - clo over z in
 * [zlo, zhi], as it enters r = (z - chi - clo) * invc */
static __float128 split_error(double zlo, double zhi, double chi, double clo, double invc) {
    __float128 worst = 0;
    const double ends[2] = {zlo, zhi};
    for (int e = 0; e < 2; ++e) {
        const double t = ends[e] - chi;   /* exact: z and chi lie within a factor of 2 */
        const double res = t - clo;
        const __float128 err = fabsq((__float128)res - ((__float128)t - (__float128)clo));
        if (err > worst) worst = err;
    }
    return worst * (__float128)invc;
}

int main(int argc, char **argv) {
    static long cand[MAX_CAND];
    const int whole = argc < 3;
    const int first = whole ? 0 : atoi(argv[1]), end = whole ? 128 : atoi(argv[2]);
    if (first < 0 || end > 128 || first >= end) { fprintf(stderr, "entries must lie in 0..128\n"); return 2; }
    if (whole) {
        printf("// Generated by gen_log_table.c: entry i = {1/c, log(c)} for the log's subinterval i.\n");
        printf("#pragma once\n#include <cstdint>\nnamespace amcfd {\n");
        printf("static constexpr uint64_t kLogTable[256] = {\n");
    }
    for (int i = first; i < end; ++i) {
        const double zlo = i < 80 ? 0.6875 + i * 0x1p-8 : 1.0 + (i - 80) * 0x1p-7;
        const double width = i < 80 ? 0x1p-8 : 0x1p-7;
        const double zhi = zlo + width - (i < 80 ? 0x1p-53 : 0x1p-52);   /* last double of the subinterval */
        const double centre = zlo + 0.5 * width;
        const uint64_t b0 = bits(1.0 / centre);
        const u128 half_unit = (u128)1 << (FRAC - 1);
        const u128 frac_mask = ((u128)1 << FRAC) - 1;
        const u128 near = (u128)1 << (FRAC - 23);   /* 2^-66 in units of 2^-43 */
        int ncand = 0;
        const long nblocks = 2 * HALF_RANGE / BLOCK;
#pragma omp parallel for schedule(static)
        for (long blk = 0; blk < nblocks; ++blk) {
            const long k0 = -HALF_RANGE + blk * BLOCK;
            const double invs = from_bits(b0 + k0);
            const double u = from_bits(b0 + k0 + 1) - invs;   /* one ulp, constant inside the binade */
            const __float128 e = (__float128)u / (__float128)invs;
            /* F(m) = -log(invs + m u) 2^43 = F0 + F1 m + F2 m^2 + O(m^3) */
            const __float128 F0 = -logq((__float128)invs) * 0x1p43Q;
            const __float128 F1 = -e * 0x1p43Q;
            const __float128 F2 = e * e * 0.5Q * 0x1p43Q;
            i128 f = to_fixed(F0);
            i128 d1 = to_fixed(F1 + F2);
            const i128 d2 = to_fixed(2 * F2);
            for (long m = 0; m < BLOCK; ++m) {
                const u128 fr = (u128)f & frac_mask;
                const u128 dist = fr < half_unit ? fr : ((u128)1 << FRAC) - fr;
                if (dist < near) {
#pragma omp critical
                    {
                        if (ncand < MAX_CAND) cand[ncand] = k0 + m;
                        ++ncand;
                    }
                }
                f += d1;
                d1 += d2;
            }
        }
        if (ncand > MAX_CAND) { fprintf(stderr, "entry %d: too many candidates\n", i); return 1; }
        long best = 0;
        __float128 best_err = 1;
        int nkept = 0;
        for (int c = 0; c < ncand; ++c) {
            const double invc = from_bits(b0 + cand[c]);
            const __float128 L = -logq((__float128)invc);
            const double logc = (double)L;
            const double on_grid = logc * 0x1p43;
            if (on_grid != (double)(int64_t)on_grid) continue;                      /* (1) */
            const __float128 err = fabsq(L - (__float128)logc);
            if (!(err < 0x1p-66Q)) continue;                                          /* (3) */
            const __float128 cq = 1 / (__float128)invc;
            const double chi = (double)cq;
            const double clo = (double)(cq - (__float128)chi);
            if (!(split_error(zlo, zhi, chi, clo, invc) < 0x1p-66Q)) continue;      /* (2) */
            ++nkept;
            if (err < best_err || (err == best_err && cand[c] < best)) { best_err = err; best = cand[c]; }
        }
        if (nkept == 0) { fprintf(stderr, "entry %d: no candidate meets the criteria\n", i); return 1; }
        const double invc = from_bits(b0 + best);
        const double logc = (double)(-logq((__float128)invc));
        fprintf(stderr, "%3d invc %a logc %a (%d of %d candidates kept)\n", i, invc, logc, nkept, ncand);
        printf("    0x%016llxull, 0x%016llxull,\n", (unsigned long long)bits(invc), (unsigned long long)bits(logc));
    }
    if (whole) printf("};\n}  // namespace amcfd\n");
    return 0;
}
