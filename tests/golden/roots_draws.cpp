// Writes the first COUNT draws of std::uniform_int_distribution<int64_t>(0, N-1)
// driven by a default-constructed std::knuth_b -- the generator pair the
// reference's root selection uses (cpu_2d/main.cpp:351-358).  Output: one
// draw per line.  Built and run by make_golden.py with the system g++.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    const int64_t n = std::atoll(argv[1]);
    const int count = std::atoi(argv[2]);
    std::knuth_b generator;
    std::uniform_int_distribution<int64_t> distribution(0, n - 1);
    for (int i = 0; i < count; ++i) std::printf("%lld\n", (long long)distribution(generator));
    return 0;
}
