// Randomised check of rtm_hip::Arena (rtm-acceleration_amd/csrc/arena.hpp)
// against a byte map: live blocks never overlap, stay inside the block and
// are 256-byte aligned; after everything is released one allocation of the
// whole block succeeds again (coalescing is complete); an allocation fails
// only when no free run of the rounded size exists (first fit finds any).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "arena.hpp"

int main(int argc, char** argv) {
    const unsigned seed = argc > 1 ? (unsigned)std::atoi(argv[1]) : 1u;
    const size_t size = 1 << 20;
    std::vector<char> base(size);
    std::vector<int> owner(size / 256, -1);       // per 256-byte granule
    rtm_hip::Arena a;
    a.reset(base.data(), size);
    std::mt19937 rng(seed);
    struct Live { char* p; size_t n; int id; };
    std::vector<Live> live;
    int next_id = 0;
    for (int step = 0; step < 20000; ++step) {
        const bool do_alloc = live.empty() || rng() % 100 < 55;
        if (do_alloc) {
            size_t n = 1 + rng() % (rng() % 4 == 0 ? 65536 : 4096);
            char* p = static_cast<char*>(a.alloc(n));
            const size_t g = (n + 255) / 256;
            // is there any free run of g granules?
            size_t run = 0;
            bool any = false;
            for (size_t k = 0; k < owner.size() && !any; ++k) {
                run = owner[k] < 0 ? run + 1 : 0;
                any = run >= g;
            }
            if (!p) {
                if (any) { std::printf("FAIL alloc refused with a free run (step %d)\n", step); return 1; }
                continue;
            }
            if (!any) { std::printf("FAIL alloc granted without a free run (step %d)\n", step); return 1; }
            const size_t off = (size_t)(p - base.data());
            if (off % 256 || off + g * 256 > size) { std::printf("FAIL placement (step %d)\n", step); return 1; }
            for (size_t k = off / 256; k < off / 256 + g; ++k) {
                if (owner[k] >= 0) { std::printf("FAIL overlap (step %d)\n", step); return 1; }
                owner[k] = next_id;
            }
            if (!a.contains(p)) { std::printf("FAIL contains (step %d)\n", step); return 1; }
            live.push_back({p, n, next_id++});
        } else {
            const size_t i = rng() % live.size();
            const Live l = live[i];
            live.erase(live.begin() + (long)i);
            a.release(l.p);
            const size_t off = (size_t)(l.p - base.data());
            for (size_t k = off / 256; k < off / 256 + (l.n + 255) / 256; ++k) owner[k] = -1;
        }
    }
    for (const Live& l : live) a.release(l.p);
    if (a.alloc(size) != base.data()) { std::printf("FAIL not coalesced\n"); return 1; }
    std::printf("ok\n");
    return 0;
}
