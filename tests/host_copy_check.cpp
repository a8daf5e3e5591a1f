// Checks sodecl_amd::parallel_copy (sodecl_amd/csrc/host_copy.hpp): every
// byte copied for sizes around the slice boundaries and odd remainders,
// from one thread and from several threads at once.
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../sodecl_amd/csrc/host_copy.hpp"

static int check(size_t bytes, unsigned seed)
{
    std::vector<unsigned char> src(bytes), dst(bytes + 64, 0xAB);
    for (size_t i = 0; i < bytes; ++i) src[i] = (unsigned char)(i * 131u + seed);
    sodecl_amd::parallel_copy(dst.data(), src.data(), bytes);
    for (size_t i = 0; i < bytes; ++i)
        if (dst[i] != src[i]) {
            std::printf("mismatch at %zu of %zu\n", i, bytes);
            return 1;
        }
    for (size_t i = bytes; i < bytes + 64; ++i)
        if (dst[i] != 0xAB) {
            std::printf("write past the end at %zu of %zu\n", i, bytes);
            return 1;
        }
    return 0;
}

int main()
{
    const size_t mib = (size_t)1 << 20;
    const size_t sl = sodecl_amd::kCopySliceMin;
    const size_t sizes[] = {0, 1, 63, sl - 1, sl, sl + 3, 2 * sl - 1, 2 * sl, 4 * sl + 3, 16 * mib, 16 * mib + 7,
                            33 * mib + 5};
    int bad = 0;
    for (size_t b : sizes) bad += check(b, 1);
    // several callers at once: jobs run one after another
    std::vector<std::thread> callers;
    std::vector<int> res(4, 0);
    for (int c = 0; c < 4; ++c)
        callers.emplace_back([&, c] {
            for (int r = 0; r < 8; ++r) res[c] += check(5 * mib + 8 * c + r, 7 + c);
        });
    for (auto& t : callers) t.join();
    for (int r : res) bad += r;
    std::printf("workers %d, %s\n", sodecl_amd::HostWorkers::get().size(), bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
