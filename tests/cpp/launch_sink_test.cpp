// CPU unit test of the launch recording in transientx_amd/csrc/launch_sink.h,
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_capi.py.  It records launches of host functions with kernel-like
// signatures into a LaunchSink, the way the graph pipelines do, and checks:
//   * every argument is stored at the width of the parameter (an int passed
//     for a long is widened), aligned, and read back through params();
//   * a copied KNode (the beam keeps one of its transpose node) owns its
//     argument bytes: params() of the copy point into the copy, and
//     rewriting its first argument (a new block source) leaves the original
//     unchanged;
//   * nodes survive the sink's vector growing (the graph holds up to
//     2 + 3 * rows nodes).
// No GPU call is made: the sink path never launches.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../transientx_amd/csrc/launch_sink.h"

using namespace txgpu;

struct Args24 {
    const float *p;
    long n;
    int k;
    float t;
};

static void k_a(const uint8_t *, long, int, uint8_t, float *, long, const long *) {}
static void k_b(Args24) {}
static void k_c(long *, long) {}

static int fails = 0;
#define CHECK(c)                                                                                   \
    do {                                                                                           \
        if (!(c)) {                                                                                \
            std::fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c);                       \
            fails++;                                                                               \
        }                                                                                          \
    } while (0)

template <typename T>
static T at(KNode &n, size_t i)
{
    T v;
    std::vector<void *> p = n.params();
    std::memcpy(&v, p[i], sizeof(T));
    return v;
}

int main()
{
    LaunchSink sink;
    uint8_t blk[16] = {0};
    float ring[4] = {0};
    long clock = 7;
    const int ns_int = 26123;  // an int for the long parameter
    const int zp_int = 128;    // an int for the uint8_t parameter
    const int pitch_int = 300; // an int for the long parameter
    CHECK(kemit(&sink, nullptr, k_a, dim3(3, 2), dim3(256), 0, blk, ns_int, 4096, zp_int, ring, pitch_int,
                &clock) == hipSuccess);
    Args24 a{ring, 1L << 40, -3, 0.5f};
    CHECK(kemit(&sink, nullptr, k_b, dim3(1), dim3(64), 128, a) == hipSuccess);
    for (int i = 0; i < 100; i++) // the vector reallocates many times
        CHECK(kemit(&sink, nullptr, k_c, dim3(1), dim3(1), 0, &clock, i) == hipSuccess);
    CHECK(sink.nodes.size() == 102);

    KNode &n0 = sink.nodes[0];
    CHECK(n0.offs.size() == 7 && n0.sizes.size() == 7);
    const size_t want[7] = {sizeof(const uint8_t *), sizeof(long), sizeof(int), sizeof(uint8_t), sizeof(float *),
                            sizeof(long), sizeof(const long *)};
    for (int i = 0; i < 7; i++) {
        CHECK(n0.sizes[i] == want[i]);
        CHECK(n0.offs[i] % want[i] == 0);
        CHECK(n0.offs[i] + n0.sizes[i] <= n0.blob.size());
    }
    CHECK(at<const uint8_t *>(n0, 0) == blk);
    CHECK(at<long>(n0, 1) == 26123L);
    CHECK(at<int>(n0, 2) == 4096);
    CHECK(at<uint8_t>(n0, 3) == 128);
    CHECK(at<float *>(n0, 4) == ring);
    CHECK(at<long>(n0, 5) == 300L);
    CHECK(at<const long *>(n0, 6) == &clock);
    CHECK(n0.grid.x == 3 && n0.grid.y == 2 && n0.block.x == 256);

    KNode &n1 = sink.nodes[1];
    CHECK(n1.sizes.size() == 1 && n1.sizes[0] == sizeof(Args24) && n1.shm == 128);
    Args24 b = at<Args24>(n1, 0);
    CHECK(b.p == ring && b.n == (1L << 40) && b.k == -3 && b.t == 0.5f);
    for (int i = 0; i < 100; i++) CHECK(at<long>(sink.nodes[2 + i], 1) == i);

    // the beam's copy of its transpose node, and a new block source in it
    KNode trans = sink.nodes[0];
    std::vector<void *> pc = trans.params();
    for (size_t i = 0; i < pc.size(); i++) {
        CHECK((unsigned char *)pc[i] >= trans.blob.data());
        CHECK((unsigned char *)pc[i] + trans.sizes[i] <= trans.blob.data() + trans.blob.size());
    }
    uint8_t other[8];
    const uint8_t *src = other;
    std::memcpy(trans.blob.data() + trans.offs[0], &src, sizeof(src));
    CHECK(at<const uint8_t *>(trans, 0) == other);
    CHECK(at<const uint8_t *>(sink.nodes[0], 0) == blk);
    sink.nodes.clear(); // the copy outlives the sink
    CHECK(at<long>(trans, 1) == 26123L);

    if (fails) return 1;
    std::printf("launch_sink ok\n");
    return 0;
}
