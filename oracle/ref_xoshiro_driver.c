/* Driver for the reference's own RNG header (include/random_xorshift.h),
 * compiled in place from the reference tree by oracle/Makefile (target ref).
 * Usage: xoshiro_ref <seed> <skip> <count>  -> prints <count> uint64 draws
 * (decimal, one per line) after skipping <skip> draws, using exactly the
 * reference's xoshiro256ss_init(seed) and xoshiro256ss(state). */
#include <stdio.h>
#include <stdlib.h>
#include <inttypes.h>
#include "random_xorshift.h"

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s seed skip count\n", argv[0]);
        return 2;
    }
    uint64_t seed = strtoull(argv[1], NULL, 10);
    uint64_t skip = strtoull(argv[2], NULL, 10);
    uint64_t count = strtoull(argv[3], NULL, 10);
    struct xoshiro256ss_state st = xoshiro256ss_init(seed);
    for (uint64_t i = 0; i < skip; i++) xoshiro256ss(&st);
    for (uint64_t i = 0; i < count; i++) printf("%" PRIu64 "\n", xoshiro256ss(&st));
    return 0;
}
