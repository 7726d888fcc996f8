// Host-only checks of the drop-in's value types (no GPU call): Candidate's
// (distance, id) order and dynamic_bitset's Boost-compatible behaviour.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "panns/Searching.h"

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            return 1;                                              \
        }                                                          \
    } while (0)

int main()
{
    using PANNS::Candidate;
    std::vector<Candidate> v = {{7, 0.5f, false}, {3, 0.5f, true}, {9, 0.25f, false}, {1, 1.0f, false}};
    std::sort(v.begin(), v.end());
    CHECK(v[0].id_ == 9 && v[1].id_ == 3 && v[2].id_ == 7 && v[3].id_ == 1);
    CHECK(v[1].is_checked_ && !v[2].is_checked_);

    PANNS::dynamic_bitset b(130);
    CHECK(b.size() == 130 && b.num_blocks() == 3 && b.none());
    b[0] = true;
    b[64] = true;
    b[129] = true;
    CHECK(b[0] && b[64] && b[129] && !b[1] && !b[128]);
    CHECK(b.count() == 3);
    CHECK(b.blocks()[0] == 1ull && b.blocks()[1] == 1ull && b.blocks()[2] == 2ull);
    b[64] = b[0];
    b[0] = false;
    CHECK(!b[0] && b[64] && b.count() == 2);
    b.reset();
    CHECK(b.none() && b.size() == 130);
    PANNS::dynamic_bitset c;
    c.resize(70, true);
    CHECK(c.count() == 70 && c.blocks()[1] == (1ull << 6) - 1);
    printf("ok\n");
    return 0;
}
