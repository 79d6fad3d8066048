// Emits Philox4x32-10 words and uniforms from rocRAND's host-callable engine
// (rocrand_philox4x32_10.h, MIT) for the (seed, subsequence, offset) triples read
// from stdin: one "seed subsequence offset count" per line.  Output: one line
// per triple with `count` u32 words followed by their rocRAND uniforms.
#include <cstdio>
#include <rocrand/rocrand_kernel.h>

int main()
{
    unsigned long long seed, sub, off;
    int count;
    while (std::scanf("%llu %llu %llu %d", &seed, &sub, &off, &count) == 4) {
        rocrand_state_philox4x32_10 st;
        rocrand_init(seed, sub, off, &st);
        unsigned int w[64];
        for (int i = 0; i < count; ++i) w[i] = rocrand(&st);
        for (int i = 0; i < count; ++i) std::printf("%u ", w[i]);
        for (int i = 0; i < count; ++i)
            std::printf("%.9g ", (double)rocrand_device::detail::uniform_distribution(w[i]));
        std::printf("\n");
    }
    return 0;
}
