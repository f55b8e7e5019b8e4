/*
 * ref_philox_kat.cpp -- CPU ORACLE support (test infrastructure).
 *
 * Driver around the reference's vendored Random123 philox.h
 * (3P/_outside_libraries/Random123/philox.h, used by reference
 * 3P/_internal/shared.cuh:67-82 as r123::Philox4x32_R<10>).  Built by
 * oracle/Makefile straight from the reference tree into oracle/_ref/; it
 * reads lines "key0 key1 c0 c1 c2 c3" (unsigned decimal) on stdin and
 * prints the four output words in hex, one line per input.  Used only to
 * write tests/golden/philox_kat.txt.
 */
#include <stdio.h>
#include <Random123/philox.h>

int main() {
    typedef r123::Philox4x32_R<10> P;
    P rng;
    unsigned k0, k1, c0, c1, c2, c3;
    while (scanf("%u %u %u %u %u %u", &k0, &k1, &c0, &c1, &c2, &c3) == 6) {
        P::key_type key = {{k0, k1}};
        P::ctr_type ctr = {{c0, c1, c2, c3}};
        P::ctr_type r = rng(ctr, key);
        printf("%u %u %u %u %u %u %08x %08x %08x %08x\n", k0, k1, c0, c1, c2, c3,
               (unsigned)r.v[0], (unsigned)r.v[1], (unsigned)r.v[2], (unsigned)r.v[3]);
    }
    return 0;
}
