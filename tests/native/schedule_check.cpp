// Prints, as one JSON object per line, the step arguments that job_args
// (csrc/batch_args.hpp) gives every step of a batched job, for the cases
// named on the command line as nt:store:fmt.  Buffers are fake addresses, so
// each pointer prints as the index of what it points at; the test checks
// the schedule's dataflow (which field each step reads and writes, which
// snapshot slot and scale buffers) on the host, and which steps may run a
// lean build (step_is_lean).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "batch_args.hpp"

using rtm_hip::JobDesc;
using rtm_hip::StepArgs;

static constexpr long long kInt = 1000;   // interior cells
static constexpr int kCap = 100;          // per-wave maxima per parity

static long long idx(const void* p, uintptr_t base, long long stride) {
    if (!p) return -1;
    return (long long)(((uintptr_t)p - base) / (uintptr_t)stride);
}

int main(int argc, char** argv) {
    for (int c = 1; c < argc; ++c) {
        int nt = 0, store = 0, fmt = 0;
        if (std::sscanf(argv[c], "%d:%d:%d", &nt, &store, &fmt) != 3) return 2;
        JobDesc J;
        std::memset(&J, 0, sizeof J);
        J.nt = nt;
        J.store = store;
        J.snap_fmt = fmt;
        J.n_int = kInt;
        J.parts_cap = kCap;
        J.bwd_base = nt & 1;
        // fields 0-1, the dry-step output 2, hold buffers 3-4, the zero field 5 (fld() below)
        for (int k = 0; k < 3; ++k) J.fld[k] = (float*)(uintptr_t)(0x1000000u * (k + 1));
        for (int k = 0; k < 2; ++k) J.hold[k] = (float*)(uintptr_t)(0x1000000u * (k + 4));
        J.zero = (const float*)(uintptr_t)(0x1000000u * 6);
        J.pair_first = rtm_hip::pair_first(nt, store, fmt != 0);
        J.snap32 = (float*)(uintptr_t)0x10000000u;
        J.snap16 = (uint16_t*)(uintptr_t)0x20000000u;
        J.wave_max = (float*)(uintptr_t)0x30000000u;
        J.slice_max = (float*)(uintptr_t)0x40000000u;
        J.part_counts = (int*)(uintptr_t)0x50000000u;
        J.snap_max = (float*)(uintptr_t)0x60000000u;
        const bool s16 = fmt != 0;
        J.steps = 2 * nt + (s16 ? 1 : 0);
        for (int total = 0; total < J.steps; ++total) {
            if (s16 && total == nt) {
                std::printf("{\"case\":\"%s\",\"total\":%d,\"kind\":\"book\"}\n", argv[c], total);
                continue;
            }
            const StepArgs a = s16 ? rtm_hip::job_args<true>(J, total, 0) : rtm_hip::job_args<false>(J, total, 0);
            const StepArgs ad = s16 ? rtm_hip::job_args<true>(J, total, 1) : rtm_hip::job_args<false>(J, total, 1);
            auto fld = [&](const float* p) { return p ? (long long)((uintptr_t)p / 0x1000000u) - 1 : -1LL; };
            std::printf(
                "{\"case\":\"%s\",\"total\":%d,\"kind\":\"%s\",\"it\":%d,\"prev\":%lld,\"cur\":%lld,\"next\":%lld,"
                "\"zp\":%d,\"zc\":%d,\"snap_out\":%lld,\"snap_in\":%lld,\"wmax\":%lld,\"wcount\":%lld,"
                "\"red_in\":%lld,\"red_count\":%lld,\"red_out\":%lld,\"red_out_count\":%lld,"
                "\"qparts\":%lld,\"qcount\":%lld,\"qout\":%lld,\"qmax\":%lld,"
                "\"s16in\":%lld,\"s16max\":%lld,\"pending\":%lld,\"lean\":%d,"
                "\"hold\":%lld,\"snap_in2\":%lld,\"s16in2\":%lld,\"s16max2\":%lld,\"dry_next\":%lld,"
                "\"dry_prev\":%lld,\"dry_cur\":%lld}\n",
                argv[c], total, total < nt ? "fwd" : "bwd", a.it, fld(a.p_prev), fld(a.p_cur), fld(a.p_next),
                a.zero_prev, a.zero_cur, idx(a.snap_out, 0x10000000u, kInt * 4), idx(a.snap_in, 0x10000000u, kInt * 4),
                idx(a.wave_max, 0x30000000u, kCap * 4), idx(a.wave_count, 0x50000000u, 4),
                idx(a.red_in, 0x30000000u, kCap * 4), idx(a.red_count, 0x50000000u, 4),
                idx(a.red_out, 0x40000000u, rtm_hip::kReduceWaves * 4), idx(a.red_out_count, 0x50000000u, 4),
                idx(a.quant_parts, 0x40000000u, rtm_hip::kReduceWaves * 4), idx(a.quant_count, 0x50000000u, 4),
                idx(a.quant_out, 0x20000000u, kInt * 2), idx(a.quant_max_out, 0x60000000u, 4),
                idx(a.snap16_in, 0x20000000u, kInt * 2), idx(a.snap16_max, 0x60000000u, 4), fld(a.pending_field),
                (int)rtm_hip::step_is_lean(total < nt ? rtm_hip::Dir::Forward : rtm_hip::Dir::Backward, a, s16),
                fld(a.hold_field), idx(a.snap_in2, 0x10000000u, kInt * 4), idx(a.snap16_in2, 0x20000000u, kInt * 2),
                idx(a.snap16_max2, 0x60000000u, 4), fld(ad.p_next), fld(ad.p_prev), fld(ad.p_cur));
        }
    }
    return 0;
}
