// Emits Philox4x32-10 words from PyTorch's at::philox_engine (ATen/core/
// PhiloxRNGEngine.h, BSD) for (seed, subsequence, offset) triples on stdin.
// The engine's offset counts 128-bit blocks, so offset = 4*block + word is
// realised as block skip plus `word` discarded draws.
#include <cstdio>
#include <ATen/core/PhiloxRNGEngine.h>

int main()
{
    unsigned long long seed, sub, off;
    int count;
    while (std::scanf("%llu %llu %llu %d", &seed, &sub, &off, &count) == 4) {
        at::philox_engine eng(seed, sub, off / 4);
        for (unsigned long long d = 0; d < off % 4; ++d) (void)eng();
        for (int i = 0; i < count; ++i) std::printf("%u ", (unsigned)eng());
        std::printf("\n");
    }
    return 0;
}
