// Host test of lbx::CopyPool (linbox_amd/csrc/copy_pool.h): every byte of each
// copy lands, nothing outside the destination range is written, for worker
// counts 0..15, sizes around the split threshold and the 4 KiB part
// boundaries, unaligned source and destination offsets, and many copies in a
// row on one pool.  Built with -fsanitize=thread and with address/undefined.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "copy_pool.h"

int main()
{
    const size_t sizes[] = {0, 1, 4095, 4097, (1u << 20) - 1, 1u << 20, (1u << 20) + 1, 3 * 4096 + 17,
                            (5u << 20) + 4093};
    const int workers[] = {0, 1, 3, 7, 15};
    const size_t guard = 64;
    const size_t offs[][2] = {{0, 0}, {3, 5}};
    std::vector<uint8_t> src((5u << 20) + 4093 + 2 * guard + 8);
    for (size_t i = 0; i < src.size(); ++i) src[i] = (uint8_t)(i * 2654435761u >> 13);
    long bad = 0, copies = 0;
    for (int w : workers) {
        lbx::CopyPool pool(w);
        for (int rep = 0; rep < 2; ++rep)
            for (size_t n : sizes)
                for (const auto &o : offs) {
                    std::vector<uint8_t> dst(n + 2 * guard + 8, 0xA5);
                    uint8_t *d = dst.data() + guard + o[1];
                    const uint8_t *s = src.data() + guard + o[0] + rep;
                    pool.copy(d, s, n);
                    ++copies;
                    if (n && std::memcmp(d, s, n) != 0) {
                        std::printf("mismatch: workers %d size %zu offsets %zu %zu\n", w, n, o[0], o[1]);
                        ++bad;
                    }
                    for (size_t i = 0; i < dst.size(); ++i) {
                        const bool inside = dst.data() + i >= d && dst.data() + i < d + n;
                        if (!inside && dst[i] != 0xA5) {
                            std::printf("write outside: workers %d size %zu at %zu\n", w, n, i);
                            ++bad;
                            break;
                        }
                    }
                }
    }
    std::printf("%ld copies, %ld failures\n", copies, bad);
    return bad ? 1 : 0;
}
