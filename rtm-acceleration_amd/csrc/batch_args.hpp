// batch_args.hpp -- the step arguments of a batched job at a given step,
// computed from its descriptor on the device (tick-argument table) or on
// the host (a job that runs alone).
#pragma once

#include <cstddef>

#include "batch_kernels.hpp"

namespace rtm_hip {

// Two field buffers: step s of a phase writes p^{n+1} over its own p^{n-1},
// in place (a step reads p^{n-1} only at its own output cells, in the lane
// that writes p^{n+1} there, before that write), so the buffer of step s is
// fld[(base + s) & 1] and steps s and s - 2 share it.
__host__ __device__ inline float* rotate(const JobDesc& J, int r) {
    return (r & 1) ? J.fld[1] : J.fld[0];
}

// The field written by backward step s of job J: the rotation's buffer, or
// on a Hold step of paired imaging a hold buffer (pair_step).  The step two
// after a Hold step, whose p^{n-1} is held, still writes buffer
// (bwd_base + s) & 1: the one that does not hold its p^n, last written by
// step s - 4.  bwd_base = nt & 1 keeps the forward phase's last field
// (buffer (nt - 1) & 1) intact through backward step 0, which images it
// when it is a pending 16-bit snapshot.
__host__ __device__ inline float* bwd_out(const JobDesc& J, int s) {
    const PairStep p = pair_step(J.nt - 1 - s, J.store, J.pair_first);
    return p.role == PairRole::Hold ? J.hold[p.hold] : rotate(J, J.bwd_base + s);
}

// Step arguments of job J after `total` of its steps (forward steps
// 0..nt-1, the bookkeeping tick of a 16-bit job, then backward steps
// 0..nt-1); written to the tick-argument table by k_tick_args, or built on
// the host for a job that runs alone (plain step-kernel launches).
template <bool S16>
__host__ __device__ inline StepArgs job_args(const JobDesc& J, int total, int dry) {
    StepArgs a{};
    constexpr bool wide = !S16;
    const int fwd = total < J.nt;
    const int s = fwd ? total : total - J.nt - (wide ? 0 : 1);
    // a phase's first steps read p^{n-1} (steps 0, 1) and p^n (step 0) as
    // zero: from the model's zero field, so the march tile loads them as any
    // other field (the flags serve the other mappings)
    a.zero_prev = s <= 1;
    a.zero_cur = s == 0;
    if (fwd) {
        a.p_prev = a.zero_prev ? J.zero : rotate(J, s);
        a.p_cur = a.zero_cur ? J.zero : rotate(J, s + 1);
        a.p_next = rotate(J, s);
    } else {
        a.p_prev = a.zero_prev ? J.zero : bwd_out(J, s - 2);
        a.p_cur = a.zero_cur ? J.zero : bwd_out(J, s - 1);
        a.p_next = bwd_out(J, s);
    }
    // a dry step (the autotuner's timing launches on live jobs) must not
    // overwrite the job's p^{n-1} in place, nor a held field: it updates the
    // scratch field in place instead, which touches as many arrays as a real
    // step (its values are never used)
    if (dry) a.p_prev = a.p_next = J.fld[2];
    a.coef = J.coef;
    a.vdt2 = J.vdt2;
    a.wx = J.wx;
    a.wz = J.wz;
    a.nxe = J.nxe; a.nze = J.nze; a.ldz = J.ldz; a.nb = J.nb; a.nx = J.nx; a.nz = J.nz;
    a.ntz = J.ntz; a.ntx = J.ntx; a.ntiles = J.ntiles;
    a.ntz_magic = J.ntz_magic;
    const int it = fwd ? s : J.nt - 1 - s;
    a.it = it;
    a.tile_ptr = fwd ? J.tile_ptr[0] : J.tile_ptr[1];
    a.entries = fwd ? J.entries[0] : J.entries[1];
    a.wavelet = J.wavelet;
    a.amps = J.amps;
    a.amp_count = J.amp_count;
    a.image = J.image;
    a.illum = J.illum;
    a.dry = dry;
    a.snap_fmt = J.snap_fmt;
    // Every field is assigned once, unconditionally: stores of different
    // fields in different branches can be merged into one store at a
    // computed offset, which would keep the whole struct in scratch memory.
    const int st = J.store;
    const bool stored = (it % st) == 0;
    const long long slot = it / st;
    const bool half16 = !wide && fwd;
    const bool red = half16 && it >= 1 && ((it - 1) % st) == 0;          // slices of step it-1's maxima
    const bool quant = half16 && it >= 2 && ((it - 2) % st) == 0;        // codes of step it-2
    const int p0 = (int)(slot & 1), p1 = (int)(((it - 1) / st) & 1), p2 = (int)(((it - 2) / st) & 1);
    // the last forward field, imaged at backward step 0 (its slices were
    // reduced by the bookkeeping tick)
    const bool pending = !wide && !fwd && stored && s == 0;
    // backward imaging by pairing role: none on a Hold step; a Pair step also
    // images the held stored step, slot + 1
    const PairRole role = fwd ? PairRole::None : pair_step(it, st, J.pair_first).role;
    const bool imaging = !fwd && stored && role != PairRole::Hold;
    const bool paired = role == PairRole::Pair;
    a.hold_field = paired ? J.hold[pair_step(it, st, J.pair_first).hold] : nullptr;
    a.snap_in2 = (wide && paired) ? J.snap32 + (slot + 1) * J.n_int : nullptr;
    a.snap16_in2 = (!wide && paired) ? J.snap16 + (slot + 1) * J.n_int : nullptr;
    a.snap16_max2 = (!wide && paired) ? J.snap_max + slot + 1 : nullptr;
    a.snap_out = (wide && fwd && stored) ? J.snap32 + slot * J.n_int : nullptr;
    a.snap_in = (wide && imaging) ? J.snap32 + slot * J.n_int : nullptr;
    a.wave_max = (half16 && stored) ? J.wave_max + (ptrdiff_t)p0 * J.parts_cap : nullptr;
    a.wave_count = (half16 && stored) ? J.part_counts + p0 : nullptr;
    a.red_in = red ? J.wave_max + (ptrdiff_t)p1 * J.parts_cap : nullptr;
    a.red_count = red ? J.part_counts + p1 : nullptr;
    a.red_out = red ? J.slice_max + p1 * kReduceWaves : nullptr;
    a.red_out_count = red ? J.part_counts + 2 + p1 : nullptr;
    a.quant_parts = quant ? J.slice_max + p2 * kReduceWaves : (pending ? J.slice_max + p0 * kReduceWaves : nullptr);
    a.quant_count = quant ? J.part_counts + 2 + p2 : (pending ? J.part_counts + 2 + p0 : nullptr);
    a.quant_out = quant ? J.snap16 + (it - 2) / st * J.n_int : nullptr;
    a.quant_max_out = quant ? J.snap_max + (it - 2) / st : nullptr;
    a.parts_cap = J.parts_cap;
    a.snap16_max = (!wide && imaging && !pending) ? J.snap_max + slot : nullptr;
    a.pending_field = pending ? rotate(J, J.nt - 1) : nullptr;
    a.snap16_in = (!wide && imaging && !pending) ? J.snap16 + slot * J.n_int : nullptr;
    return a;
}

}  // namespace rtm_hip
