/* Thin extern "C" entry point over the reference Graph500 generator, compiled
 * from the reference's own sources where they lie (see oracle/Makefile).
 * Used only to write and check golden fixtures. */
#include <stdint.h>
#include "graph_generator.h"
#include "utils.h"

int ref_generate(int log_numverts, int64_t start, int64_t end, uint64_t s1, uint64_t s2,
                 int64_t *edges)
{
    uint_fast32_t seed[5];
    make_mrg_seed(s1, s2, seed);
    generate_kronecker_range(seed, log_numverts, start, end, (packed_edge *)edges);
    return 0;
}
