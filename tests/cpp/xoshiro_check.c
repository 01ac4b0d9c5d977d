/* Independent C statement of the seeded input generator of
 * theses_code_amd/rng.py: splitmix64 seeding of 1024 xoshiro256** lanes,
 * outputs interleaved lane by lane.  Prints the first n 64-bit outputs of
 * the stream of `seed`, one per line, in decimal. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define LANES 1024

static uint64_t splitmix64_next(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t xoshiro_next(uint64_t* s) {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    uint64_t seed = strtoull(argv[1], NULL, 0);
    long n = strtol(argv[2], NULL, 10);
    static uint64_t st[LANES][4];
    uint64_t x = seed;
    for (int i = 0; i < LANES; ++i)
        for (int k = 0; k < 4; ++k) st[i][k] = splitmix64_next(&x);
    for (long v = 0; v < n; ++v) printf("%llu\n", (unsigned long long)xoshiro_next(st[v % LANES]));
    return 0;
}
