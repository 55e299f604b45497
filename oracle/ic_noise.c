/* Seeded initial-condition noise shared by the oracle and the reference driver.
 *
 * TEST INFRASTRUCTURE: part of the CPU oracle. The product computes the same
 * sequence in nim-conv22-shared_amd/fortran/nim_host_noise.c; a CPU test
 * checks that both agree value for value.
 *
 * The reference has no random initial condition (general.f90:16-31 returns 0
 * for every field); the benchmark cases add this seeded perturbation so that
 * the solver is exercised away from the trivial state. The generator is the
 * splitmix64 finaliser applied to a key made of (seed, field, cell index), so
 * every cell's value is independent of the traversal order and of the mesh
 * decomposition.
 */
#include <stdint.h>

static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform value in [-1, 1) for (seed, field, idx); idx is the 0-based cell
 * number in the reference's natural order (theta fastest, then r, then z). */
double oracle_ic_noise(int64_t seed, int32_t field, int64_t idx)
{
    uint64_t key = (uint64_t)seed * 0xD1342543DE82EF95ull;
    key ^= ((uint64_t)(uint32_t)field << 48);
    key += (uint64_t)idx;
    uint64_t z = mix64(mix64(key));
    return (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}
