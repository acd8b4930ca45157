// Compares amcfd::div_pos (three dependent operations) and its confirmation
// amcfd::div_pos_exact with IEEE division over seeded random operands.
// MODE 0: random quotients over a wide exponent range; MODE 1: dividends
// placed within a few ulps of b times a rounding midpoint of the quotient,
// where a one-correction quotient is most likely to round the wrong way;
// MODE 2: zero dividends of both signs; MODE 3: the correctly rounded
// quotient moved by one ulp either way, which the confirmation must flag
// every time.  Prints how many quotients differ from x / b, how many the
// confirmation flags, and how many differing quotients it misses (must be 0).
// Usage: div_pos_check MODE N [SEED]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../am-cfd_amd/csrc/div_pos.h"

static unsigned long long s;
static unsigned long long next() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return s;
}
static double uni(double lo, double hi) { return lo + (hi - lo) * ((next() >> 11) * 0x1p-53); }

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const int mode = std::atoi(argv[1]);
    const long n = std::atol(argv[2]);
    s = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 88172645463325252ull;
    long wrong = 0, flagged = 0, missed = 0;
    for (long i = 0; i < n; ++i) {
        const double b = std::ldexp(uni(1.0, 2.0), (int)(next() % 60) - 30);
        double x;
        if (mode == 0) {
            x = std::ldexp(uni(-2.0, 2.0), (int)(next() % 80) - 40);
        } else if (mode == 1) {
            const double q = std::ldexp(uni(1.0, 2.0), (int)(next() % 40) - 20);
            const double h = std::nextafter(q, INFINITY) - q;
            x = b * (q + 0.5 * h);
            for (int k = (int)(next() % 9) - 4; k != 0; k += (k > 0 ? -1 : 1))
                x = std::nextafter(x, k > 0 ? INFINITY : -INFINITY);
            if (next() & 1) x = -x;
        } else if (mode == 2) {
            x = (next() & 1) ? -0.0 : 0.0;
        } else {
            x = std::ldexp(uni(-2.0, 2.0), (int)(next() % 80) - 40);
        }
        const double y = 1.0 / b, ref = x / b;
        double q = amcfd::div_pos(x, b, y);
        if (mode == 3) q = std::nextafter(ref, (next() & 1) ? INFINITY : -INFINITY);
        const bool bad = __builtin_bit_cast(long long, q) != __builtin_bit_cast(long long, ref);
        const bool flag = !amcfd::div_pos_exact(x, b, y, q);
        if (bad && wrong < 5) std::printf("x=%a b=%a q=%a ref=%a\n", x, b, q, ref);
        wrong += bad;
        flagged += flag;
        missed += bad && !flag;
    }
    std::printf("wrong %ld flagged %ld missed %ld of %ld\n", wrong, flagged, missed, n);
    return 0;
}
