// march_tile.hpp -- device code shared by the per-shot and the batched step
// kernels: the march-mapping tile (one wave64 computes one tile of
// (64 - 2H) depth cells x TX columns; see step_kernels.hip) and its helpers.
// Included only by HIP translation units.
#pragma once

#include "step_kernels.hpp"

namespace rtm_hip {
namespace device {

// Every buffer the step touches is device global memory.  Pointers that reach
// a kernel through memory (the batched kernel reads its arguments from a
// per-job table) would otherwise be accessed with flat instructions, which
// also count against the LDS/scalar wait counter; the views below make every
// access a global one, as for pointers passed as kernel arguments.
template <typename T>
using global_t = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ global_t<T>* gbl(T* p) {
    return (global_t<T>*)p;
}

// Values that no wave of the launch writes (tile tables, and the scales and
// counts of earlier launches) are read through the constant address space,
// so a wave-uniform read becomes a scalar load: it waits on its own counter,
// and using the value does not wait for the vector loads still in flight
// (the vector counter retires in order).
template <typename T>
using const_t = __attribute__((address_space(4))) const T;
template <typename T>
__device__ __forceinline__ T sld(const T* p) {
    return *(const_t<T>*)p;
}

// The kernel arguments on a tile's path to its first field load and first
// arithmetic, requested in one batch at kernel entry: an empty asm that takes
// them as scalar operands makes the compiler load them all before it, with
// one wait.  Left to itself, it requests them in two or three rounds (around
// the tile-index arithmetic and the early exit, and some after the field
// loads), each round a scalar round trip on the wave's critical path
// (backward 2301 x 751 step 7.66 -> 7.46 us, forward 5.89 -> 5.78 us).  More
// operands than these make it reload some of them after the wait; the
// entry count and the time step, first needed by the entry loads after the
// field loads, come with the entry bounds there.  (The stencil coefficients
// of shifts 1-4 are among them; an order-10 or order-12 stencil loads its
// last ones later.)
__device__ __forceinline__ void request_args(const StepArgs& a) {
    asm volatile("" ::"s"(a.p_cur), "s"(a.p_prev), "s"(a.vdt2), "s"(a.wz), "s"(a.wx), "s"(a.p_next),
                 "s"(a.ldz), "s"(a.nze), "s"(a.nxe), "s"(a.ntz_magic), "s"(a.ntz), "s"(a.ntiles));
    asm volatile("" ::"s"(a.tile_ptr), "s"(a.entries), "s"(a.amps), "s"(a.coef.c0), "s"(a.coef.cx[1]),
                 "s"(a.coef.cx[2]), "s"(a.coef.cx[3]), "s"(a.coef.cx[4]), "s"(a.coef.cz[1]), "s"(a.coef.cz[2]),
                 "s"(a.coef.cz[3]), "s"(a.coef.cz[4]));
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Give XCD k the
// k-th contiguous block of logical tiles so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
    const int q = ntiles >> 3, r = ntiles & 7;
    const int xcd = b & 7, loc = b >> 3;
    return (xcd < r) ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

// Value held by lane+1 / lane-1 of the wave (DPP wave_shl:1 / wave_shr:1, a
// VALU move instead of an LDS-crossbar permute).  Lanes with no source read 0;
// they are never output lanes for shifts up to H.
// bound_ctrl: a lane with no source reads 0 without an old value to merge, so
// a shift is one DPP move with no initialising move before it.
__device__ __forceinline__ float from_lane_above(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_lane_below(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

// max over the wave's 64 lanes, in every lane.  DPP moves (no LDS round
// trips): a running max along each row of 16 lanes (row_shr 1, 2, 4, 8; a lane
// with no source keeps its own value), then row_bcast:15 and row_bcast:31
// carry the row maxima up to lane 63, which is read back.  max is exact, so
// the order does not change the result.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_max(float v) {
    const int o = __float_as_int(v);
    return fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(o, o, kCtrl, kRowMask, 0xf, false)));
}
__device__ __forceinline__ float wave_max(float m) {
    m = dpp_max<0x111, 0xf>(m);      // row_shr:1
    m = dpp_max<0x112, 0xf>(m);      // row_shr:2
    m = dpp_max<0x114, 0xf>(m);      // row_shr:4
    m = dpp_max<0x118, 0xf>(m);      // row_shr:8
    m = dpp_max<0x142, 0xa>(m);      // row_bcast:15 into rows 1 and 3
    m = dpp_max<0x143, 0xc>(m);      // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), 63));
}

// max(0, parts[0..n)) over a workgroup of kThreads threads, in thread 0.  Each
// thread issues its loads four at a time (independent loads in flight, one
// round trip per four values instead of one per value); max is exact, so the
// order of the reduction does not change the result.
template <int kThreads>
__device__ __forceinline__ float block_max(const float* parts, int n) {
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
    int i = (int)threadIdx.x;
    for (; i + 3 * kThreads < n; i += 4 * kThreads) {
        const float a = gbl(parts)[i], b = gbl(parts)[i + kThreads];
        const float c = gbl(parts)[i + 2 * kThreads], d = gbl(parts)[i + 3 * kThreads];
        m0 = fmaxf(m0, a);
        m1 = fmaxf(m1, b);
        m2 = fmaxf(m2, c);
        m3 = fmaxf(m3, d);
    }
    for (; i < n; i += kThreads) m0 = fmaxf(m0, gbl(parts)[i]);
    float m = wave_max(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)));
    __shared__ float wm[kThreads / kTileZ];
    if ((threadIdx.x & (kTileZ - 1)) == 0) wm[threadIdx.x / kTileZ] = m;
    __syncthreads();
    m = threadIdx.x < kThreads / kTileZ ? wm[threadIdx.x] : 0.0f;
    return wave_max(m);
}

// v * 2^k rounded to FP16 or BF16 (round to nearest even).
__device__ __forceinline__ uint16_t encode_snapshot(float v, int k, int fmt) {
    const float s = ldexpf(v, k);
    if (fmt == (int)SnapFmt::BF16) {
        unsigned b = __float_as_uint(s);
        b += 0x7fffu + ((b >> 16) & 1u);
        return (uint16_t)(b >> 16);
    }
    const _Float16 h = (_Float16)s;
    uint16_t bits;
    __builtin_memcpy(&bits, &h, sizeof bits);
    return bits;
}

// The first 64 injection entries of a tile, one per lane, with their
// amplitude at this step (source: wavelet[it]; receiver, backward: its trace
// sample), loaded at the start of the tile together with its other operands.
// The epilogue then reads them with v_readlane (no memory round trip after
// the stencil); entries past the first 64 are read from memory there.
struct TileEntries {
    int base;          // the tile's first entry
    int cell, code;    // entry base + lane (cell -1 past the tile's entries)
    float amp;
};

// tile_ptr[t] (t is the wave's tile, the same in every lane) with a scalar
// load: no wait on the vector loads in flight, and the entry loops run on
// scalar registers, with scalar loads for their rare entries past 64.
__device__ __forceinline__ int entry_bound(const StepArgs& a, int t) {
    return sld(a.tile_ptr + __builtin_amdgcn_readfirstlane(t));
}

template <Dir D>
__device__ __forceinline__ TileEntries load_entries(const StepArgs& a, int e0, int e1, int lane) {
    TileEntries te;
    te.base = e0;
    te.cell = -1;
    te.code = 0;
    te.amp = 0.0f;
    // No load here depends on another: the forward amplitude is the wavelet
    // sample for every entry (only source entries, code < 0, use it).
    if constexpr (D == Dir::Forward) te.amp = e0 < e1 ? sld(a.wavelet + a.it) : 0.0f;
    if (e0 + lane < e1) {
        te.cell = gbl(a.entries)[e0 + lane].cell;
        te.code = gbl(a.entries)[e0 + lane].code;
        if constexpr (D == Dir::Backward) te.amp = gbl(a.amps)[(ptrdiff_t)a.it * a.amp_count + e0 + lane];
    }
    return te;
}

// v, computed here with the whole wave active: an empty asm that reads and
// writes the register keeps the compiler from sinking the arithmetic into
// the masked store block after it (there, each column would wait again for
// every earlier load and store of the wave, the previous column's store
// included).
__device__ __forceinline__ float settle(float v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Waits, once and before any store of the tile, for the entry registers:
// a first use inside the per-cell entry loops would wait there instead, on
// a counter that by then also counts the tile's earlier stores.
__device__ __forceinline__ void entries_ready(const TileEntries& te) {
    asm volatile("" ::"v"(te.cell), "v"(te.code), "v"(te.amp));
}

__device__ __forceinline__ int entry_cell(const StepArgs& a, const TileEntries& te, int e) {
    const int k = e - te.base;
    return k < kTileZ ? __builtin_amdgcn_readlane(te.cell, k) : sld(&a.entries[e].cell);
}

// Source / receiver injection and trace recording of one output cell, in
// entry order (the tile's entries are sorted by cell, then code): entries
// [lo, hi) of the tile are the ones that may hit this cell; cell is -1 in a
// lane that outputs no cell.
//
// Must be called with every lane of the wave active (never inside a
// lane-dependent branch): the entries are read across lanes with readlane,
// and a register's value in lanes that are off in a divergent region is not
// kept by the compiler.  The loop itself has no lane-dependent branch either:
// the add is a select, and only the recording store is masked.
template <Dir D>
__device__ __forceinline__ float apply_entries(const StepArgs& a, const TileEntries& te, int cell, float out,
                                               float v2, int lo, int hi) {
    for (int e = lo; e < hi; ++e) {
        const int k = e - te.base;
        int en_cell, en_code;
        float amp;
        if (k < kTileZ) {
            en_cell = __builtin_amdgcn_readlane(te.cell, k);
            en_code = __builtin_amdgcn_readlane(te.code, k);
            amp = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(te.amp), k));
        } else {
            // entries past the wave's 64 lanes: scalar loads (e is uniform),
            // which wait on their own counter; a vector load here would make
            // the loop wait for every store before it
            en_cell = sld(&a.entries[e].cell);
            en_code = sld(&a.entries[e].code);
            if constexpr (D == Dir::Backward) amp = sld(a.amps + (ptrdiff_t)a.it * a.amp_count + e);
            else amp = en_code < 0 ? sld(a.wavelet + a.it) : 0.0f;
        }
        const bool hit = en_cell == cell;
        if (D == Dir::Forward && en_code >= 0) {          // en_code is wave-uniform
            if (hit) gbl(a.rec_out)[(ptrdiff_t)en_code * a.rec_stride + a.it] = out;
            continue;
        }
        const float inj = out + v2 * amp;
        out = hit ? inj : out;
    }
    return out;
}

// What the snapshot/imaging epilogue does on this launch.  Every kernel is
// built twice: for FP32 snapshots (S16 false) and for 16-bit ones (S16
// true), so neither carries the other's code.
struct SnapCtx {
    bool imaging;      // backward stored step
    bool store32;      // forward stored step, FP32 snapshot
    bool store_max;    // forward stored step, 16-bit snapshot (per-wave max)
    bool quantise;     // forward, two steps after a stored step (16-bit)
    int k_in;          // backward decode exponent
    int k_in2;         // paired imaging: the held stored step's decode exponent
    int k_q;           // forward encode exponent
    float q_max;       // the quantised snapshot's max
    float part;        // quant_parts[lane], loaded before the tile's loads
};

// max of parts[0 .. *count) (at most kReduceWaves values), in every lane,
// from v = parts[lane]: all kReduceWaves slots exist, so the load does not
// wait for the count (both come from the previous launch, a memory round
// trip each).
__device__ __forceinline__ float parts_max_of(float v, const int* count, int lane) {
    const int n = min(sld(count), kReduceWaves);
    return wave_max(lane < n ? v : 0.0f);
}
__device__ __forceinline__ float parts_max(const float* parts, const int* count, int lane) {
    return parts_max_of(gbl(parts)[lane], count, lane);
}

// The epilogue's flags (no memory access).
template <bool S16>
__device__ __forceinline__ SnapCtx snap_ctx(const StepArgs& a, Dir d) {
    SnapCtx c;
    if constexpr (S16) {
        c.imaging = d == Dir::Backward && (a.snap16_in != nullptr || a.pending_field != nullptr);
        c.store32 = false;
        c.store_max = d == Dir::Forward && a.wave_max != nullptr;
        c.quantise = d == Dir::Forward && a.quant_out != nullptr;
    } else {
        c.imaging = d == Dir::Backward && a.snap_in != nullptr;
        c.store32 = d == Dir::Forward && a.snap_out != nullptr;
        c.store_max = c.quantise = false;
    }
    c.k_in = c.k_in2 = c.k_q = 0;
    c.q_max = c.part = 0.0f;
    return c;
}

// Issues the load of this lane's slice max for the 16-bit scale of the
// launch (quantise, or the pending snapshot's imaging) before the tile's own
// loads: the scale's chain then overlaps them instead of following them.
template <bool S16>
__device__ __forceinline__ void snap_prefetch(const StepArgs& a, SnapCtx& c, int lane) {
    if constexpr (S16)
        if (c.quantise || (c.imaging && a.pending_field)) c.part = gbl(a.quant_parts)[lane];
}

// The 16-bit scales of the epilogue.  Called after the tile's loads are
// issued: the reduction waits for its own load, and so for all before it.
template <bool S16>
__device__ __forceinline__ void snap_scales(const StepArgs& a, SnapCtx& c, int lane) {
    if constexpr (S16) {
        if (c.imaging)
            c.k_in = snapshot_scale_exponent(a.pending_field ? parts_max_of(c.part, a.quant_count, lane)
                                                             : sld(a.snap16_max));
        if (c.imaging && a.hold_field) c.k_in2 = snapshot_scale_exponent(sld(a.snap16_max2));
        if (c.quantise) {
            c.q_max = parts_max_of(c.part, a.quant_count, lane);
            c.k_q = snapshot_scale_exponent(c.q_max);
        }
    }
}

// The stored bits of the snapshot value for imaging at interior index j /
// extended index i: an FP32 value, or for 16-bit snapshots the aligned
// 32-bit word holding the code (the 16-bit stores are allocated with one
// spare word at the end) or, for a snapshot still pending in its field, the
// field's FP32 value.  One 32-bit load from a uniformly selected address (two
// loads in uniform branches would need a wait between them, as they share a
// destination), and no use of the value here: the decode is left to
// snapshot_value, after every load of the tile is issued (a use here would
// wait for this load, and with it for every load before it).
template <bool S16>
__device__ __forceinline__ uint32_t snapshot_bits(const StepArgs& a, ptrdiff_t j, ptrdiff_t i) {
    if constexpr (!S16) return __float_as_uint(__builtin_nontemporal_load(gbl(a.snap_in) + j));
    const uintptr_t addr = a.pending_field ? (uintptr_t)(a.pending_field + i) : (uintptr_t)(a.snap16_in + j);
    return __builtin_nontemporal_load(gbl((const uint32_t*)(addr & ~(uintptr_t)3)));
}

// 16-bit code -> the stored float (exact): the hardware conversion for FP16,
// a shift for BF16, then the exact scale 2^-k.
__device__ __forceinline__ float decode_code(uint32_t code, int k, int fmt) {
    float f;
    if (fmt == (int)SnapFmt::BF16) {
        f = __uint_as_float(code << 16);
    } else {
        const uint16_t c16 = (uint16_t)code;
        _Float16 h;
        __builtin_memcpy(&h, &c16, sizeof h);
        f = (float)h;
    }
    return ldexpf(f, -k);
}

// Snapshot value f from its stored bits (snapshot_bits, interior index j):
// FP32 as is, a decoded 16-bit code (the half of the word that holds element
// j), or a pending field value encoded and decoded here, which gives the
// value the store would hold.
template <bool S16>
__device__ __forceinline__ float snapshot_value(const StepArgs& a, const SnapCtx& c, uint32_t bits, ptrdiff_t j) {
    if constexpr (!S16) return __uint_as_float(bits);
    uint32_t code;
    if (a.pending_field) code = encode_snapshot(__uint_as_float(bits), c.k_in, a.snap_fmt);
    else code = ((uintptr_t)(a.snap16_in + j) & 2) ? bits >> 16 : bits & 0xffffu;
    return decode_code(code, c.k_in, a.snap_fmt);
}

// ---------------------------------------------------------------------------
// Pieces shared by the march and stream mappings.
// ---------------------------------------------------------------------------

// Laplacian (coefficients cf) at the lane's cell of column c of the p^n window q (q[c-H..c+H]
// hold columns x-H..x+H): c0 p + sum_k (cx_k (x+k + x-k) + cz_k (z+k + z-k)),
// summed in the reference's order (src/rtm/forward.cpp:38-44).  Depth
// neighbours come from other lanes: chained one-lane DPP shifts, or
// ds_bpermute (__shfl_up/__shfl_down) by k lanes.
template <int H, bool DPP>
__device__ __forceinline__ float laplacian(const Coeffs& cf, const float* q, int c) {
    const float pc = q[c];
    float lap = cf.c0 * pc;
    float zp = pc, zm = pc;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        if constexpr (DPP) {
            zp = from_lane_above(zp);                           // p^n at z + k
            zm = from_lane_below(zm);                           // p^n at z - k
        } else {
            zp = __shfl_down(pc, (unsigned)k, kTileZ);
            zm = __shfl_up(pc, (unsigned)k, kTileZ);
        }
        const float tx_k = cf.cx[k] * (q[c + k] + q[c - k]);
        const float tz_k = cf.cz[k] * (zp + zm);
        lap = lap + (tx_k + tz_k);
    }
    return lap;
}

// Per-column operands of one output column x: p^{n-1}, (v dt)^2, the x
// taper and, for a backward imaging step, the snapshot value and the image
// and illumination being accumulated.
struct ColumnOps {
    float pv, vv, wx, iv, sv;
    uint32_t fb;                     // snapshot bits (snapshot_bits)
    uint32_t fb2;                    // paired imaging: the held stored step's snapshot bits
    float hv;                        // paired imaging: its field, kept in the hold buffer
};

// The image and illumination of one cell on an imaging step: o.iv / o.sv
// plus, on a paired step, the held stored step's f2 b2 / f2^2 first, then this
// step's f b / f^2 (the reference's order, src/rtm/imaging.cpp:25-40, applied
// in descending time).
template <bool S16>
__device__ __forceinline__ void image_cell(const StepArgs& a, int k_in2, const ColumnOps& o, float fv, float out,
                                           float& img, float& ill) {
    float iI = o.iv, sS = o.sv;
    if (a.hold_field) {                                       // wave-uniform
        float f2;
        if constexpr (S16) f2 = decode_code(o.fb2, k_in2, a.snap_fmt);
        else f2 = __uint_as_float(o.fb2);
        iI = iI + f2 * o.hv;
        sS = sS + f2 * f2;
    }
    img = iI + fv * out;
    ill = sS + fv * fv;
}

template <int H, Dir D, bool S16>
__device__ __forceinline__ ColumnOps column_ops(const StepArgs& a, const SnapCtx& sc, int x, int z, bool z_out,
                                                bool z_int, int jz) {
    ColumnOps o;
    const bool ok = z_out && x < a.nxe - H;
    const ptrdiff_t i = (ptrdiff_t)x * a.ldz + z;
    o.pv = (ok && !a.zero_prev) ? gbl(a.p_prev)[i] : 0.0f;
    o.vv = ok ? gbl(a.vdt2)[i] : 0.0f;
    o.wx = x < a.nxe ? gbl(a.wx)[x] : 0.0f;
    o.iv = o.sv = o.hv = 0.0f;
    o.fb = o.fb2 = 0;
    if constexpr (D == Dir::Backward) {
        const int jx = x - a.nb;
        const bool in = sc.imaging && ok && z_int && jx >= 0 && jx < a.nx;
        const ptrdiff_t j = (ptrdiff_t)jx * a.nz + jz;
        o.fb = in ? snapshot_bits<S16>(a, j, i) : 0u;
        o.iv = in ? gbl(a.image)[j] : 0.0f;
        o.sv = in ? gbl(a.illum)[j] : 0.0f;
        const bool pair = in && a.hold_field;
        if constexpr (S16) o.fb2 = pair ? (uint32_t)gbl(a.snap16_in2)[j] : 0u;
        else o.fb2 = pair ? __float_as_uint(__builtin_nontemporal_load(gbl(a.snap_in2) + j)) : 0u;
        o.hv = pair ? gbl(a.hold_field)[i] : 0.0f;
    }
    return o;
}

__device__ __forceinline__ float load_cur(const StepArgs& a, int x, int z, bool z_load) {
    return (z_load && x < a.nxe && !a.zero_cur) ? gbl(a.p_cur)[(ptrdiff_t)x * a.ldz + z] : 0.0f;
}

// Snapshots are written once and read once, thousands of steps apart: their
// stores and loads are non-temporal, so they do not displace the fields,
// tables and image that every step re-reads from the caches.
//
// Leapfrog update and damping of one output cell (src/rtm/forward.cpp:45,
// :56-63; src/rtm/backward.cpp:57-64):
//   p^{n+1} = w ((2 p^n - w p^{n-1}) + (v dt)^2 lap), w = wx[x] wz[z]
// (p^{n-1} is damped on read: the reference damps it in place after the
// previous step, which gives the same value; injection cells lie where
// w == 1, so injecting after the damping multiply is the reference's
// inject-then-damp).
__device__ __forceinline__ float leapfrog(float pc, float lap, const ColumnOps& o, float wzv) {
    const float w = o.wx * wzv;
    const float prev = o.pv * w;
    const float out = (2.0f * pc - prev) + o.vv * lap;
    return out * w;
}

// p^{n+1} of output cell (x, z), after injection, to the field, and the
// snapshot / imaging epilogue of the cell (src/rtm/forward.cpp:112-118;
// src/rtm/backward.cpp:66-75).
template <Dir D, bool S16>
__device__ __forceinline__ void store_cell(const StepArgs& a, const SnapCtx& sc, int x, int z, bool z_int, int jz,
                                           float out, const ColumnOps& o, float& vmax) {
    const ptrdiff_t i = (ptrdiff_t)x * a.ldz + z;
    gbl(a.p_next)[i] = out;

    const int jx = x - a.nb;
    if (z_int && jx >= 0 && jx < a.nx) {
        const ptrdiff_t j = (ptrdiff_t)jx * a.nz + jz;
        if constexpr (D == Dir::Forward) {
            if constexpr (S16) {
                if (sc.store_max) vmax = fmaxf(vmax, fabsf(out));
                if (sc.quantise && !a.dry)
                    __builtin_nontemporal_store(encode_snapshot(o.pv, sc.k_q, a.snap_fmt), gbl(a.quant_out) + j);
            } else {
                if (sc.store32 && !a.dry) __builtin_nontemporal_store(out, gbl(a.snap_out) + j);
            }
        } else {
            if (sc.imaging && !a.dry) {
                float img, ill;
                image_cell<S16>(a, sc.k_in2, o, snapshot_value<S16>(a, sc, o.fb, j), out, img, ill);
                gbl(a.image)[j] = img;
                gbl(a.illum)[j] = ill;
            }
        }
    }
}

// One cell of column x with its injection entries [e0, e1), applied per
// cell with the whole wave active; only the store is masked to output lanes.
template <Dir D, bool S16>
__device__ __forceinline__ void finish_cell(const StepArgs& a, const SnapCtx& sc, int x, int z, bool x_out,
                                            bool z_int, int jz, float pc, float lap, const ColumnOps& o, float wzv,
                                            const TileEntries& te, int e0, int e1, float& vmax) {
    float out = leapfrog(pc, lap, o, wzv);
    out = apply_entries<D>(a, te, x_out ? x * a.ldz + z : -1, out, o.vv, e0, e1);
    if (x_out) store_cell<D, S16>(a, sc, x, z, z_int, jz, out, o, vmax);
}

// Entry (en_cell, en_code, amp) on the TX output cells of a lane (cell[r]:
// the cell index of column r, -1 where the lane outputs none): an entry hits
// at most one of them; every compare and add is a select, no branch.
template <Dir D, int TX>
__device__ __forceinline__ void entry_on_cells(const StepArgs& a, const int (&cell)[TX], float (&out)[TX],
                                               const ColumnOps (&o)[TX], int en_cell, int en_code, float amp) {
    if (D == Dir::Forward && en_code >= 0) {
        // trace recording: the cell's value at this point of its entry order
        bool any = false;
        float v = 0.0f;
#pragma unroll
        for (int r = 0; r < TX; ++r) {
            const bool hit = cell[r] == en_cell;
            v = hit ? out[r] : v;
            any = any || hit;
        }
        if (any) gbl(a.rec_out)[(ptrdiff_t)en_code * a.rec_stride + a.it] = v;
        return;
    }
#pragma unroll
    for (int r = 0; r < TX; ++r) {
        const float inj = out[r] + o[r].vv * amp;
        out[r] = cell[r] == en_cell ? inj : out[r];
    }
}

// Source / receiver injection and trace recording of a whole tile, in one
// pass over its entries [e0, e1) (sorted by cell, then code), so every cell
// sees its own entries in entry order, as the per-cell loop of
// apply_entries does; a per-cell loop costs each tile that holds entries
// TX loops with their branches, and those tiles set the end of the launch.
template <Dir D, int TX>
__device__ __forceinline__ void apply_tile_entries(const StepArgs& a, const TileEntries& te, const int (&cell)[TX],
                                                   float (&out)[TX], const ColumnOps (&o)[TX], int e0, int e1) {
    const int e64 = min(e1, te.base + kTileZ);
    for (int e = e0; e < e64; ++e) {
        const int k = e - te.base;
        const int en_cell = __builtin_amdgcn_readlane(te.cell, k);
        const int en_code = D == Dir::Forward ? __builtin_amdgcn_readlane(te.code, k) : 0;
        const float amp = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(te.amp), k));
        entry_on_cells<D, TX>(a, cell, out, o, en_cell, en_code, amp);
    }
    // entries past the wave's 64 lanes: scalar loads (e is uniform)
    for (int e = e64; e < e1; ++e) {
        const int en_cell = sld(&a.entries[e].cell);
        const int en_code = sld(&a.entries[e].code);
        float amp;
        if constexpr (D == Dir::Backward) amp = sld(a.amps + (ptrdiff_t)a.it * a.amp_count + e);
        else amp = en_code < 0 ? sld(a.wavelet + a.it) : 0.0f;
        entry_on_cells<D, TX>(a, cell, out, o, en_cell, en_code, amp);
    }
}

// The wave's max |p^{n+1}| of a stored 16-bit step, to partial `part` of
// the launch's `nparts`.
template <bool S16>
__device__ __forceinline__ void finish_tile(const StepArgs& a, const SnapCtx& sc, float vmax, int lane, int part,
                                            int nparts) {
    if (!S16) return;
    if (sc.store_max && !a.dry) {
        vmax = wave_max(vmax);
        if (lane == 0) {
            gbl(a.wave_max)[part] = vmax;
            if (part == 0) *gbl(a.wave_count) = nparts;
        }
    }
}

// Wave w of a launch of nwaves: out[w] = max of slice w of in[0 .. *count)
// (count clamped to cap), for w < r = min(kReduceWaves, nwaves), in r slices;
// wave 0 writes r to *out_count.
__device__ __forceinline__ void reduce_slice(const float* in, const int* count, int cap, float* out, int* out_count,
                                             int w, int nwaves, int lane) {
    const int r = min(kReduceWaves, nwaves);
    if (w >= r) return;
    const int n = min(sld(count), cap);
    const int chunk = (n + r - 1) / r;
    const int lo = w * chunk, hi = min(n, lo + chunk);
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
    int i = lo + lane;
    for (; i + 3 * kTileZ < hi; i += 4 * kTileZ) {
        const float x0 = gbl(in)[i], x1 = gbl(in)[i + kTileZ];
        const float x2 = gbl(in)[i + 2 * kTileZ], x3 = gbl(in)[i + 3 * kTileZ];
        m0 = fmaxf(m0, x0);
        m1 = fmaxf(m1, x1);
        m2 = fmaxf(m2, x2);
        m3 = fmaxf(m3, x3);
    }
    for (; i < hi; i += kTileZ) m0 = fmaxf(m0, gbl(in)[i]);
    const float m = wave_max(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)));
    if (lane == 0) {
        gbl(out)[w] = m;
        if (w == 0) *gbl(out_count) = r;
    }
}

// The scale bookkeeping of the launch's wave `w` of `nwaves`: one slice of
// the previous stored step's per-wave maxima, and the quantised snapshot's
// max (wave 0).
template <bool S16>
__device__ __forceinline__ void settle_scales(const StepArgs& a, const SnapCtx& sc, int w, int nwaves, int lane) {
    if (!S16 || a.dry) return;
    if (sc.quantise && w == 0 && lane == 0) *gbl(a.quant_max_out) = sc.q_max;
    if (a.red_in != nullptr)
        reduce_slice(a.red_in, a.red_count, a.parts_cap, a.red_out, a.red_out_count, w, nwaves, lane);
}

// ---------------------------------------------------------------------------
// Strip addressing of the march mapping.  A wave's 64 lanes hold depth cells
// z0 .. z0+63 of the columns of its tile.  A column's address is
// wave-uniform (scalar arithmetic) and each lane adds its own offset, fixed
// for the wave: the vector unit, the step's binding resource, spends one
// address add per access and nothing on bounds.  Lane offsets are clamped to
// the last row of their array and interior columns to the last interior
// column; field columns need no clamping, since every field, (v dt)^2 and the
// x taper end with kGuardCols columns of zeros, past which no tile's window
// reaches.  A clamped lane or column, and a guard column, only ever feed
// cells that are not outputs.  (Buffer instructions, with the column in a
// scalar offset and no address arithmetic at all, measured slower on this
// chip than global loads and stores.)
// ---------------------------------------------------------------------------
struct Strip {
    int z0, x0;             // first depth cell of the strip, first output column
    int z, jz;              // this lane's depth cell and interior depth index
    int lz;                 // lane offset in a field column: min(lane, nze-1-z0)
    int jzc;                // interior depth index clamped to [0, nz)
    bool z_out, z_int;      // output lane; output lane inside the interior
};

// t must be wave-uniform (a read-first-lane value or one derived from it).
template <int H>
__device__ __forceinline__ Strip make_strip(const StepArgs& a, int t, int lane, int tx_cols) {
    constexpr int TZ = kTileZ - 2 * H;
    Strip s;
    // scalar arithmetic: t / ntz by the host's reciprocal (exact for every
    // tile of the launch)
    const int tx = a.ntz == 1 ? t : (int)__umulhi((unsigned)t, a.ntz_magic);
    const int tz = t - tx * a.ntz;
    s.z0 = tz * TZ;
    s.x0 = H + tx * tx_cols;
    s.z = s.z0 + lane;
    s.jz = s.z - a.nb;
    s.z_out = lane >= H && lane < kTileZ - H && s.z < a.nze - H;
    s.z_int = s.z_out && s.jz >= 0 && s.jz < a.nz;
    s.lz = min(lane, a.nze - 1 - s.z0);
    s.jzc = min(max(s.jz, 0), a.nz - 1);
    return s;
}

// Column x (x < nxe + kGuardCols) of a field at the strip's first row.
template <typename T>
__device__ __forceinline__ global_t<T>* field_col(T* p, const StepArgs& a, const Strip& s, int x) {
    return gbl(p) + ((ptrdiff_t)x * a.ldz + s.z0);
}
// Interior column jx, clamped to [0, nx), of an interior array.
template <typename T>
__device__ __forceinline__ global_t<T>* interior_col(T* p, const StepArgs& a, int jx) {
    return gbl(p) + (ptrdiff_t)min(max(jx, 0), a.nx - 1) * a.nz;
}

// The snapshot value of stored bits fb (a 16-bit code, an FP32 value, or a
// pending field's FP32 value encoded and decoded here, which gives the value
// the 16-bit store would hold).
template <bool S16>
__device__ __forceinline__ float strip_snapshot(const StepArgs& a, const SnapCtx& c, uint32_t fb) {
    if constexpr (!S16) return __uint_as_float(fb);
    const uint32_t code = a.pending_field ? encode_snapshot(__uint_as_float(fb), c.k_in, a.snap_fmt) : fb;
    return decode_code(code, c.k_in, a.snap_fmt);
}

// The imaging operands of output columns x0 .. x0+TX-1 of a backward step
// that images: the snapshot bits, the image and the illumination.  One
// uniform branch around all of them (a branch per column between the tile's
// loads lets the compiler delay the loads behind it).  Columns outside the
// interior read a clamped column and are never imaged.
template <int TX, bool S16>
__device__ __forceinline__ void strip_imaging_ops(const StepArgs& a, const Strip& s, ColumnOps (&o)[TX]) {
#pragma unroll
    for (int r = 0; r < TX; ++r) {
        const int jx = s.x0 + r - a.nb;
        if constexpr (S16) {
            // a 16-bit code, or (first backward step of a batched job) the
            // FP32 field still pending its encoding: both loads are from
            // valid memory (the unused one from the image), one is kept
            const uint16_t* codes = a.pending_field ? (const uint16_t*)a.image : a.snap16_in;
            const float* pend = a.pending_field ? a.pending_field : a.p_cur;
            const uint32_t code = __builtin_nontemporal_load(interior_col(codes, a, jx) + s.jzc);
            const uint32_t pv = __float_as_uint(field_col(pend, a, s, s.x0 + r)[s.lz]);
            o[r].fb = a.pending_field ? pv : code;
        } else {
            o[r].fb = __float_as_uint(__builtin_nontemporal_load(interior_col(a.snap_in, a, jx) + s.jzc));
        }
        o[r].iv = interior_col(a.image, a, jx)[s.jzc];
        o[r].sv = interior_col(a.illum, a, jx)[s.jzc];
    }
    if (a.hold_field) {                                       // wave-uniform: a paired step
#pragma unroll
        for (int r = 0; r < TX; ++r) {
            const int jx = s.x0 + r - a.nb;
            if constexpr (S16) o[r].fb2 = __builtin_nontemporal_load(interior_col(a.snap16_in2, a, jx) + s.jzc);
            else o[r].fb2 = __float_as_uint(__builtin_nontemporal_load(interior_col(a.snap_in2, a, jx) + s.jzc));
            o[r].hv = field_col(a.hold_field, a, s, s.x0 + r)[s.lz];
        }
    }
}

// p^{n+1} of output cell (x, lane) to the field (NTS: a non-temporal store),
// and its snapshot / imaging epilogue (src/rtm/forward.cpp:112-118;
// src/rtm/backward.cpp:66-75).  The caller has checked that the lane and the
// column are outputs.
template <Dir D, bool S16, bool NTS = false>
__device__ __forceinline__ void strip_store(const StepArgs& a, const SnapCtx& sc, const Strip& s, int x, float out,
                                            const ColumnOps& o, float& vmax) {
    if constexpr (NTS) __builtin_nontemporal_store(out, field_col(a.p_next, a, s, x) + s.lz);
    else field_col(a.p_next, a, s, x)[s.lz] = out;
    const int jx = x - a.nb;
    if constexpr (D == Dir::Forward && S16) {
        // one lane predicate for the interior instead of early exits: the
        // maximum is a select, the code store the only masked instruction
        const bool in = (unsigned)jx < (unsigned)a.nx && s.z_int;
        if (sc.store_max) vmax = fmaxf(vmax, in ? fabsf(out) : 0.0f);
        if (sc.quantise && !a.dry) {
            const uint16_t code = encode_snapshot(o.pv, sc.k_q, a.snap_fmt);
            if (in) __builtin_nontemporal_store(code, gbl(a.quant_out) + ((ptrdiff_t)jx * a.nz + s.jzc));
        }
        return;
    }
    if (jx < 0 || jx >= a.nx) return;                       // wave-uniform
    if (!s.z_int) return;
    const ptrdiff_t j = (ptrdiff_t)jx * a.nz + s.jzc;
    if constexpr (D == Dir::Forward) {
        if constexpr (!S16) {
            if (sc.store32 && !a.dry) __builtin_nontemporal_store(out, gbl(a.snap_out) + j);
        }
    } else {
        if (sc.imaging && !a.dry) {
            float img, ill;
            image_cell<S16>(a, sc.k_in2, o, strip_snapshot<S16>(a, sc, o.fb), out, img, ill);
            gbl(a.image)[j] = img;
            gbl(a.illum)[j] = ill;
        }
    }
}

// ---------------------------------------------------------------------------
// March mapping: one wave64 = one tile of (64 - 2H) x TX outputs, every load
// of the tile issued before any use.
// ---------------------------------------------------------------------------
// LEAN: a step with no snapshot work of its own cells (step_is_lean: a
// backward step that does not image, a forward step with no snapshot store,
// maxima or codes), built without that epilogue; a 16-bit forward step that
// only reduces the previous stored step's maxima runs lean.
// PH (16-bit forward steps): the snapshot work compiled in, 1 the maxima of
// the step's own field, 2 the codes of the field two steps back, 3 both; a
// build with only one of them leaves out the other's code.
// t must be wave-uniform (a value read-first-lane'd or derived from one).
template <int H, int TX, Dir D, bool DPP, bool S16, bool LEAN = false, int PH = 3>
__device__ __forceinline__ void march_tile(const StepArgs& a, const Coeffs& cf, int t, int lane) {
    constexpr int NQ = TX + 2 * H;           // p^n columns a lane needs
    // a lean step has no snapshot operand of its own: the FP32 build's code
    constexpr bool S = S16 && !LEAN;
    const Strip s = make_strip<H>(a, t, lane, TX);
    const int x0 = s.x0;
    SnapCtx sc = snap_ctx<S>(a, D);
    if constexpr (LEAN) sc.imaging = sc.store32 = sc.store_max = sc.quantise = false;
    if constexpr (!(PH & 1)) sc.store_max = false;
    if constexpr (!(PH & 2)) sc.quantise = false;
    snap_prefetch<S>(a, sc, lane);
    // The entry bounds are scalar loads, requested here, ahead of the tile's
    // loads (the compiler may still schedule them after those, as it does in
    // the lean builds; either way they are waited for only before the entry
    // loads); the 16-bit build has its slice-max load here already and does
    // better with the bounds after the tile's loads.
    int e0 = 0, e1 = 0;
    if constexpr (!S) e0 = entry_bound(a, t), e1 = entry_bound(a, t + 1);

    // Every load of the tile, issued before any use, in the order the
    // columns use them: the depth taper, p^n columns 0 .. 2H-1, then per
    // output column r the p^n column r + 2H that completes its window and
    // its own p^{n-1} and (v dt)^2.  The vector memory counter retires in
    // order, so the arithmetic of column r waits only for the loads up to
    // its own, and overlaps the latency of the later columns' loads.
    const float wzv = gbl(a.wz)[s.z0 + s.lz];
    const global_t<const float>* pc = field_col(a.p_cur, a, s, x0 - H);
    const global_t<const float>* pp = field_col(a.p_prev, a, s, x0);
    const global_t<const float>* pv2 = field_col(a.vdt2, a, s, x0);
    float q[NQ];
    ColumnOps ops[TX];
#pragma unroll
    for (int j = 0; j < 2 * H; ++j) q[j] = pc[(ptrdiff_t)j * a.ldz + s.lz];
#pragma unroll
    for (int r = 0; r < TX; ++r) {
        q[r + 2 * H] = pc[(ptrdiff_t)(r + 2 * H) * a.ldz + s.lz];
        ops[r].pv = pp[(ptrdiff_t)r * a.ldz + s.lz];
        ops[r].vv = pv2[(ptrdiff_t)r * a.ldz + s.lz];
        ops[r].wx = sld(a.wx + x0 + r);
        ops[r].iv = ops[r].sv = ops[r].hv = 0.0f;
        ops[r].fb = ops[r].fb2 = 0;
    }
    if constexpr (D == Dir::Backward && !LEAN)
        if (sc.imaging) strip_imaging_ops<TX, S>(a, s, ops);
    if constexpr (S) e0 = entry_bound(a, t), e1 = entry_bound(a, t + 1);
    const TileEntries te = load_entries<D>(a, e0, e1, lane);
    snap_scales<S>(a, sc, lane);
    float vmax = 0.0f;

    // One pass over the columns computes every output once and stores each
    // cell as soon as it is computed, which keeps the stores streaming.  A
    // tile with injection entries applies them in one pass after the
    // columns: a lean build (a field store and nothing else per cell) has
    // stored every column already and stores again, from the same lane to
    // the same address, only the cells whose bits the entries changed; a
    // build with a snapshot or imaging epilogue, which must see the value
    // after injection, keeps its outputs and stores them after the pass.
    // (Two separate column loops, one per case, made the compiler hoist
    // their common sums above the branch and hold every column live at
    // once.)  No early exit past the last column: only the stores are
    // guarded, so no branch stands between the tile's loads.
    const bool plain = e0 == e1;                            // wave-uniform
    const bool early = LEAN || plain;
    // Forward steps store p^{n+1} non-temporally (5.94 against 6.22 us per
    // 2301 x 751 step); backward steps gain nothing from it, and batched
    // backward ticks lose.
    constexpr bool NTS = D == Dir::Forward;
    float out[TX];
#pragma unroll
    for (int r = 0; r < TX; ++r) {
        out[r] = settle(leapfrog(q[r + H], laplacian<H, DPP>(cf, q, r + H), ops[r], wzv));
        const int x = x0 + r;
        if (early && x < a.nxe - H && s.z_out) strip_store<D, S, NTS>(a, sc, s, x, out[r], ops[r], vmax);
    }
    if (!plain) {
        // a tile with entries waits for its entry registers once, here
        entries_ready(te);
        int cell[TX];
        uint32_t stored[TX];
#pragma unroll
        for (int r = 0; r < TX; ++r) {
            cell[r] = (s.z_out && x0 + r < a.nxe - H) ? (x0 + r) * a.ldz + s.z : -1;
            stored[r] = __float_as_uint(out[r]);
        }
        apply_tile_entries<D, TX>(a, te, cell, out, ops, e0, e1);
#pragma unroll
        for (int r = 0; r < TX; ++r)
            if (cell[r] >= 0 && (!LEAN || __float_as_uint(out[r]) != stored[r]))
                strip_store<D, S, NTS>(a, sc, s, x0 + r, out[r], ops[r], vmax);
    }
    finish_tile<S>(a, sc, vmax, lane, t, a.ntiles);
    // the reduction of the previous stored step's maxima runs in lean
    // 16-bit forward steps too (it is their only snapshot work)
    settle_scales<S16>(a, sc, t, a.ntiles, lane);
}

// ---------------------------------------------------------------------------
// Stream mapping: one wave64 = one strip of (64 - 2H) depth cells marching
// along XS columns in chunks of C.  p^n sits in a sliding register window of
// C + 2H columns, so each column is loaded once per strip (the march mapping
// loads TX + 2H columns for TX outputs); the loads of chunk k+1 are issued
// before chunk k is computed and stored.  Injection entries of the strip are
// sorted by cell (column-major), so a uniform cursor hands each column its
// own entries.
// ---------------------------------------------------------------------------
template <int H, int XS, Dir D, bool S16>
__device__ __forceinline__ void stream_tile(const StepArgs& a, const Coeffs& cf, int t, int lane) {
    constexpr int C = 4;                     // columns per chunk
    constexpr int NCH = XS / C;
    constexpr int W = C + 2 * H;             // p^n window of one chunk
    constexpr int TZ = kTileZ - 2 * H;
    static_assert(XS % C == 0, "strip width must be a multiple of the chunk");

    const int tz = t % a.ntz, tx = t / a.ntz;
    const int z = tz * TZ + lane;
    const int x0 = H + tx * XS;
    const bool z_load = z < a.nze;
    const bool z_out = lane >= H && lane < kTileZ - H && z < a.nze - H;
    const int jz = z - a.nb;
    const bool z_int = z_out && jz >= 0 && jz < a.nz;
    SnapCtx sc = snap_ctx<S16>(a, D);
    snap_prefetch<S16>(a, sc, lane);
    int ec = entry_bound(a, t);                                      // in flight with the tile's loads
    const int e1 = entry_bound(a, t + 1);

    float q[W];
#pragma unroll
    for (int j = 0; j < W; ++j) q[j] = load_cur(a, x0 - H + j, z, z_load);
    ColumnOps ops[C];
#pragma unroll
    for (int r = 0; r < C; ++r) ops[r] = column_ops<H, D, S16>(a, sc, x0 + r, z, z_out, z_int, jz);
    const float wzv = z_load ? gbl(a.wz)[z] : 0.0f;
    const TileEntries te = load_entries<D>(a, ec, e1, lane);
    snap_scales<S16>(a, sc, lane);
    entries_ready(te);
    float vmax = 0.0f;

#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        const int xc = x0 + k * C;
        // next chunk's loads, in flight while this chunk computes
        float nq[C];
        ColumnOps nops[C];
        if (k + 1 < NCH) {
#pragma unroll
            for (int j = 0; j < C; ++j) nq[j] = load_cur(a, xc + C + H + j, z, z_load);
#pragma unroll
            for (int r = 0; r < C; ++r) nops[r] = column_ops<H, D, S16>(a, sc, xc + C + r, z, z_out, z_int, jz);
        }
#pragma unroll
        for (int r = 0; r < C; ++r) {
            const int x = xc + r;
            // entries of column x: cells [x ldz, (x+1) ldz)
            int ee = ec;
            const long long col_end = (long long)(x + 1) * a.ldz;
            while (ee < e1 && entry_cell(a, te, ee) < col_end) ++ee;
            const float lap = laplacian<H, true>(cf, q, r + H);
            finish_cell<D, S16>(a, sc, x, z, z_out && x < a.nxe - H, z_int, jz, q[r + H], lap, ops[r], wzv, te, ec,
                                ee, vmax);
            ec = ee;
        }
        if (k + 1 < NCH) {
#pragma unroll
            for (int j = 0; j < 2 * H; ++j) q[j] = q[j + C];
#pragma unroll
            for (int j = 0; j < C; ++j) q[2 * H + j] = nq[j];
#pragma unroll
            for (int r = 0; r < C; ++r) ops[r] = nops[r];
        }
    }
    finish_tile<S16>(a, sc, vmax, lane, t, a.ntiles);
    settle_scales<S16>(a, sc, t, a.ntiles, lane);
}

}  // namespace device
}  // namespace rtm_hip
