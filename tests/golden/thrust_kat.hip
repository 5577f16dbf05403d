// Known answers of the Thrust RNG the reference renders with (thrust::minstd_rand +
// uniform_real_distribution<float>, uniform_int_distribution<int>, uniform_real_distribution<int>;
// core/src/gpu.cu:770-771, core/src/misc/gpu_randomizer.cuh:11-36, core/src/misc/bvh_node.cuh:70).
// The reference builds against CUDA 12.6 Thrust, which is not in this image; rocThrust
// (THRUST_VERSION 200805, /opt/rocm/include/thrust) implements the same engine and distributions.
// Host code only. Prints JSON on stdout.
#include <thrust/random.h>
#include <cstdio>
#include <cstring>

int main() {
    std::printf("{\"thrust_version\": %d,\n", THRUST_VERSION);
    const unsigned seeds[4] = {1u, 2580u, 2580u + 65535u, 1984u};
    std::printf("\"minstd_uniform_float\": {");
    for (int s = 0; s < 4; ++s) {
        thrust::minstd_rand rng(seeds[s]);
        thrust::uniform_real_distribution<float> d(0.0f, 1.0f);
        std::printf("%s\"%u\": [", s ? ", " : "", seeds[s]);
        for (int i = 0; i < 64; ++i) {
            float u = d(rng);
            unsigned b;
            std::memcpy(&b, &u, 4);
            std::printf("%s%u", i ? ", " : "", b);
        }
        std::printf("]");
    }
    std::printf("},\n\"minstd_raw_10000th_default\": ");
    {
        thrust::minstd_rand rng;
        unsigned v = 0;
        for (int i = 0; i < 10000; ++i) v = rng();
        std::printf("%u,\n", v);
    }
    std::printf("\"uniform_int_0_2_seed1984\": [");
    {
        thrust::minstd_rand rng(1984);
        thrust::uniform_int_distribution<int> d(0, 2);
        for (int i = 0; i < 64; ++i) std::printf("%s%d", i ? ", " : "", d(rng));
    }
    std::printf("],\n\"uniform_real_int_0_3_seed7\": [");
    {
        thrust::minstd_rand rng(7);
        thrust::uniform_real_distribution<int> d(0, 3);
        for (int i = 0; i < 16; ++i) std::printf("%s%d", i ? ", " : "", d(rng));
    }
    std::printf("]}\n");
    return 0;
}
