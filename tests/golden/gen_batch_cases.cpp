// gen_batch_cases.cpp -- writes batch_cases.csv, the cylinder batch table drawn with the C++
// standard library's std::mt19937_64(20261202): per case, in draw order, size index = rng() % 4,
// Mach = 0.2 + 0.6 U, angle of attack = -5 + 10 U degrees, U = (rng() >> 11) * 2^-53.
//   g++ -std=c++17 -O2 -o /tmp/gen_batch_cases gen_batch_cases.cpp && /tmp/gen_batch_cases > batch_cases.csv
#include <cstdio>
#include <random>

int main()
{
    const int sizes[4][2] = {{128, 64}, {256, 128}, {384, 192}, {512, 256}};
    std::mt19937_64 rng(20261202ULL);
    std::printf("case,size,nth,nr,mach,alpha_deg\n");
    for (int k = 0; k < 96; ++k) {
        const unsigned long long s = rng() % 4;
        const double u1 = double(rng() >> 11) * 0x1.0p-53;
        const double mach = 0.2 + 0.6 * u1;
        const double u2 = double(rng() >> 11) * 0x1.0p-53;
        const double alpha = -5.0 + 10.0 * u2;
        std::printf("%d,%llu,%d,%d,%.17g,%.17g\n", k, s, sizes[s][0], sizes[s][1], mach, alpha);
    }
    return 0;
}
