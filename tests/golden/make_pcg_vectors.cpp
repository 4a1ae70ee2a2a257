// Writes known-answer vectors for PCG32 (XSH-RR 64/32, settable stream) from
// Melissa O'Neill's published C++ implementation (Apache-2.0 / MIT), as
// vendored inside pyarrow. The header is only used here, at fixture-generation
// time; the committed JSON is the fixture.
//
// Build + run (see make_pcg_vectors.sh):
//   g++ -O1 -std=c++17 -I<pyarrow>/include/arrow/vendored make_pcg_vectors.cpp
//   ./a.out > pcg32_vectors.json
#include <pcg/pcg_random.hpp>

#include <cstdint>
#include <cstdio>

using arrow_vendored::pcg32;

struct Case {
    uint64_t seed, stream, skip;
    int count;
};

int main() {
    const Case cases[] = {
        {42u, 54u, 0u, 8},
        {0x4E525253ull, 0u, 0u, 8},
        {0x4E525253ull ^ 1ull, 7u, 0u, 8},
        {0x4E525253ull, 3u, 1000003u, 6},
        {123456789ull, 987654321ull, (1ull << 20) + 17ull, 6},
        {0ull, 0ull, (1ull << 40) + 5ull, 4},
    };
    std::printf("{\n  \"generator\": \"pcg32 setseq_xsh_rr_64_32 (pcg_random.hpp)\",\n  \"cases\": [\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int c = 0; c < n; ++c) {
        pcg32 rng(cases[c].seed, cases[c].stream);
        rng.advance(cases[c].skip);
        std::printf("    {\"seed\": %llu, \"stream\": %llu, \"skip\": %llu, \"out\": [",
                    (unsigned long long)cases[c].seed, (unsigned long long)cases[c].stream,
                    (unsigned long long)cases[c].skip);
        for (int i = 0; i < cases[c].count; ++i) {
            std::printf("%s%u", i ? ", " : "", (unsigned)rng());
        }
        std::printf("]}%s\n", c + 1 < n ? "," : "");
    }
    std::printf("  ]\n}\n");
    return 0;
}
