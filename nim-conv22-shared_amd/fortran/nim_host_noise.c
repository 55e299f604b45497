/* Seeded initial-condition perturbation used by the Fortran host
 * (nimh_initial_state). The reference starts from rest (general.f90:16-31);
 * the perturbation is keyed by (seed, field, cell) through the splitmix64
 * finaliser so that every cell's value is independent of traversal order.
 * Fortran has no unsigned 64-bit arithmetic, hence this C helper. */
#include <stdint.h>

static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double nimh_noise(int64_t seed, int32_t field, int64_t idx)
{
    uint64_t key = (uint64_t)seed * 0xD1342543DE82EF95ull;
    key ^= ((uint64_t)(uint32_t)field << 48);
    key += (uint64_t)idx;
    uint64_t z = mix64(mix64(key));
    return (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}
