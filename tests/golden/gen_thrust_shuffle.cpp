// Writes golden vectors of thrust::shuffle(thrust::seq, ..., default_random_engine(seed))
// as ROCm's rocThrust implements it, for the `b -s` tie-break of balance
// (balance.cu:404-411 shuffles each cover's inputs with the engine seeded by
// the cover's index in its batch, then stable-sorts them by level).
// Build and run (host code only):
//   hipcc -O1 -std=c++17 -o /tmp/gen_thrust_shuffle tests/golden/gen_thrust_shuffle.cpp
//   /tmp/gen_thrust_shuffle > tests/golden/thrust_shuffle.txt
// Output: one line per case, "len seed p0 p1 ...", where position i of the
// shuffled sequence holds input element p_i.
#include <thrust/execution_policy.h>
#include <thrust/random.h>
#include <thrust/shuffle.h>
#include <cstdio>
#include <vector>

int main() {
    const int lens[] = {1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 1000, 1024, 1088};
    const unsigned seeds[] = {0, 1, 2, 3, 7, 100, 4095, 32767};
    for (int len : lens)
        for (unsigned seed : seeds) {
            std::vector<int> v(len);
            for (int i = 0; i < len; i++) v[i] = i;
            thrust::default_random_engine rng(seed);
            thrust::shuffle(thrust::seq, v.begin(), v.end(), rng);
            printf("%d %u", len, seed);
            for (int x : v) printf(" %d", x);
            printf("\n");
        }
    return 0;
}
