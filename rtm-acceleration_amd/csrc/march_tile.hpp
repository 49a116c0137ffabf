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
__device__ __forceinline__ float from_lane_above(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_lane_below(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, kTileZ));
    return m;
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

template <Dir D>
__device__ __forceinline__ TileEntries load_entries(const StepArgs& a, int e0, int e1, int lane) {
    TileEntries te;
    te.base = e0;
    te.cell = -1;
    te.code = 0;
    te.amp = 0.0f;
    if (e0 + lane < e1) {
        te.cell = gbl(a.entries)[e0 + lane].cell;
        te.code = gbl(a.entries)[e0 + lane].code;
        if constexpr (D == Dir::Backward) te.amp = gbl(a.traces)[(ptrdiff_t)te.code * a.trace_stride + a.it];
        else if (te.code < 0) te.amp = gbl(a.wavelet)[a.it];
    }
    return te;
}

__device__ __forceinline__ int entry_cell(const StepArgs& a, const TileEntries& te, int e) {
    const int k = e - te.base;
    return k < kTileZ ? __builtin_amdgcn_readlane(te.cell, k) : gbl(a.entries)[e].cell;
}

// Source / receiver injection and trace recording of one output cell, in
// entry order (the tile's entries are sorted by cell, then code): entries
// [lo, hi) of the tile are the ones that may hit this cell.
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
            en_cell = gbl(a.entries)[e].cell;
            en_code = gbl(a.entries)[e].code;
            if constexpr (D == Dir::Backward) amp = gbl(a.traces)[(ptrdiff_t)en_code * a.trace_stride + a.it];
            else amp = en_code < 0 ? gbl(a.wavelet)[a.it] : 0.0f;
        }
        if (en_cell != cell) continue;
        if constexpr (D == Dir::Forward) {
            if (en_code < 0) out = out + v2 * amp;
            else gbl(a.rec_out)[(ptrdiff_t)en_code * a.rec_stride + a.it] = out;
        } else {
            out = out + v2 * amp;
        }
    }
    return out;
}

// What the snapshot/imaging epilogue does on this launch.
struct SnapCtx {
    bool imaging;      // backward stored step
    bool store32;      // forward stored step, FP32 snapshot
    bool store_max;    // forward stored step, 16-bit snapshot (max reduction)
    bool quantise;     // forward, previous step was stored (16-bit)
    int k_in;          // backward decode exponent
    int k_q;           // forward encode exponent
};

__device__ __forceinline__ SnapCtx snap_ctx(const StepArgs& a, Dir d) {
    SnapCtx c;
    const bool wide = a.snap_fmt == (int)SnapFmt::F32;
    c.imaging = d == Dir::Backward &&
                (wide ? a.snap_in != nullptr : (a.snap16_in != nullptr || a.pending_field != nullptr));
    c.store32 = d == Dir::Forward && wide && a.snap_out != nullptr;
    c.store_max = d == Dir::Forward && !wide && a.wave_max != nullptr;
    c.quantise = d == Dir::Forward && !wide && a.quant_out != nullptr;
    c.k_in = (c.imaging && !wide) ? snapshot_scale_exponent(*gbl(a.snap16_max)) : 0;
    c.k_q = c.quantise ? snapshot_scale_exponent(*gbl(a.quant_max)) : 0;
    return c;
}

// Snapshot value f for imaging at interior index j / extended index i: FP32,
// a decoded 16-bit code, or a 16-bit snapshot still pending in its field
// (encoded and decoded here, which gives the value the store would hold).
__device__ __forceinline__ float snapshot_value(const StepArgs& a, const SnapCtx& c, ptrdiff_t j, ptrdiff_t i) {
    if (a.snap_fmt == (int)SnapFmt::F32) return gbl(a.snap_in)[j];
    const uint16_t code =
        a.pending_field ? encode_snapshot(gbl(a.pending_field)[i], c.k_in, a.snap_fmt) : gbl(a.snap16_in)[j];
    return decode_snapshot(code, c.k_in, (SnapFmt)a.snap_fmt);
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
    float pv, vv, wx, fv, iv, sv;
};

template <int H, Dir D>
__device__ __forceinline__ ColumnOps column_ops(const StepArgs& a, const SnapCtx& sc, int x, int z, bool z_out,
                                                bool z_int, int jz) {
    ColumnOps o;
    const bool ok = z_out && x < a.nxe - H;
    const ptrdiff_t i = (ptrdiff_t)x * a.ldz + z;
    o.pv = (ok && !a.zero_prev) ? gbl(a.p_prev)[i] : 0.0f;
    o.vv = ok ? gbl(a.vdt2)[i] : 0.0f;
    o.wx = x < a.nxe ? gbl(a.wx)[x] : 0.0f;
    o.fv = o.iv = o.sv = 0.0f;
    if constexpr (D == Dir::Backward) {
        const int jx = x - a.nb;
        const bool in = sc.imaging && ok && z_int && jx >= 0 && jx < a.nx;
        const ptrdiff_t j = (ptrdiff_t)jx * a.nz + jz;
        o.fv = in ? snapshot_value(a, sc, j, i) : 0.0f;
        o.iv = in ? gbl(a.image)[j] : 0.0f;
        o.sv = in ? gbl(a.illum)[j] : 0.0f;
    }
    return o;
}

__device__ __forceinline__ float load_cur(const StepArgs& a, int x, int z, bool z_load) {
    return (z_load && x < a.nxe && !a.zero_cur) ? gbl(a.p_cur)[(ptrdiff_t)x * a.ldz + z] : 0.0f;
}

// Leapfrog update, damping, injection and the snapshot / imaging epilogue of
// one output cell (x, z) (src/rtm/forward.cpp:45, :56-63, :112-118;
// src/rtm/backward.cpp:57-75):
//   p^{n+1} = w ((2 p^n - w p^{n-1}) + (v dt)^2 lap), w = wx[x] wz[z]
// (p^{n-1} is damped on read: the reference damps it in place after the
// previous step, which gives the same value; injection cells lie where
// w == 1, so injecting after the damping multiply is the reference's
// inject-then-damp).  Entries [e0, e1) may hit this cell.
template <Dir D>
__device__ __forceinline__ void finish_cell(const StepArgs& a, const SnapCtx& sc, int x, int z, bool z_int, int jz,
                                            float pc, float lap, const ColumnOps& o, float wzv,
                                            const TileEntries& te, int e0, int e1, float& vmax) {
    const ptrdiff_t i = (ptrdiff_t)x * a.ldz + z;
    const float w = o.wx * wzv;
    const float prev = o.pv * w;
    float out = (2.0f * pc - prev) + o.vv * lap;
    out = out * w;
    out = apply_entries<D>(a, te, (int)i, out, o.vv, e0, e1);
    gbl(a.p_next)[i] = out;

    const int jx = x - a.nb;
    if (z_int && jx >= 0 && jx < a.nx) {
        const ptrdiff_t j = (ptrdiff_t)jx * a.nz + jz;
        if constexpr (D == Dir::Forward) {
            if (sc.store32 && !a.dry) gbl(a.snap_out)[j] = out;
            if (sc.store_max) vmax = fmaxf(vmax, fabsf(out));
            if (sc.quantise && !a.dry) gbl(a.quant_out)[j] = encode_snapshot(pc, sc.k_q, a.snap_fmt);
        } else {
            if (sc.imaging && !a.dry) {
                gbl(a.image)[j] = o.iv + o.fv * out;
                gbl(a.illum)[j] = o.sv + o.fv * o.fv;
            }
        }
    }
}

// The wave's max |p^{n+1}| of a stored 16-bit step, to partial `part`.
__device__ __forceinline__ void finish_tile(const StepArgs& a, const SnapCtx& sc, float vmax, int lane, int part) {
    if (sc.store_max && !a.dry) {
        vmax = wave_max(vmax);
        if (lane == 0) gbl(a.wave_max)[part] = vmax;
    }
}

// ---------------------------------------------------------------------------
// March mapping: one wave64 = one tile of (64 - 2H) x TX outputs, every load
// of the tile issued before any use.
// ---------------------------------------------------------------------------
template <int H, int TX, Dir D, bool DPP>
__device__ __forceinline__ void march_tile(const StepArgs& a, const Coeffs& cf, int t, int lane) {
    constexpr int TZ = kTileZ - 2 * H;       // output depth cells per strip
    constexpr int NQ = TX + 2 * H;           // p^n columns a lane needs

    const int tz = t % a.ntz, tx = t / a.ntz;
    const int z = tz * TZ + lane;            // depth cell held by this lane
    const int x0 = H + tx * TX;              // first output column
    const bool z_load = z < a.nze;
    const bool z_out = lane >= H && lane < kTileZ - H && z < a.nze - H;
    const int jz = z - a.nb;                 // interior depth index
    const bool z_int = z_out && jz >= 0 && jz < a.nz;
    const SnapCtx sc = snap_ctx(a, D);

    float q[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) q[j] = load_cur(a, x0 - H + j, z, z_load);
    ColumnOps ops[TX];
#pragma unroll
    for (int r = 0; r < TX; ++r) ops[r] = column_ops<H, D>(a, sc, x0 + r, z, z_out, z_int, jz);
    const float wzv = z_load ? gbl(a.wz)[z] : 0.0f;
    const int e0 = gbl(a.tile_ptr)[t], e1 = gbl(a.tile_ptr)[t + 1];
    const TileEntries te = load_entries<D>(a, e0, e1, lane);
    float vmax = 0.0f;

#pragma unroll
    for (int r = 0; r < TX; ++r) {
        const float lap = laplacian<H, DPP>(cf, q, r + H);
        const int x = x0 + r;
        if (!z_out || x >= a.nxe - H) continue;
        finish_cell<D>(a, sc, x, z, z_int, jz, q[r + H], lap, ops[r], wzv, te, e0, e1, vmax);
    }
    finish_tile(a, sc, vmax, lane, t);
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
template <int H, int XS, Dir D>
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
    const SnapCtx sc = snap_ctx(a, D);

    float q[W];
#pragma unroll
    for (int j = 0; j < W; ++j) q[j] = load_cur(a, x0 - H + j, z, z_load);
    ColumnOps ops[C];
#pragma unroll
    for (int r = 0; r < C; ++r) ops[r] = column_ops<H, D>(a, sc, x0 + r, z, z_out, z_int, jz);
    const float wzv = z_load ? gbl(a.wz)[z] : 0.0f;
    int ec = gbl(a.tile_ptr)[t];
    const int e1 = gbl(a.tile_ptr)[t + 1];
    const TileEntries te = load_entries<D>(a, ec, e1, lane);
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
            for (int r = 0; r < C; ++r) nops[r] = column_ops<H, D>(a, sc, xc + C + r, z, z_out, z_int, jz);
        }
#pragma unroll
        for (int r = 0; r < C; ++r) {
            const int x = xc + r;
            // entries of column x: cells [x ldz, (x+1) ldz)
            int ee = ec;
            const long long col_end = (long long)(x + 1) * a.ldz;
            while (ee < e1 && entry_cell(a, te, ee) < col_end) ++ee;
            const float lap = laplacian<H, true>(cf, q, r + H);
            if (z_out && x < a.nxe - H)
                finish_cell<D>(a, sc, x, z, z_int, jz, q[r + H], lap, ops[r], wzv, te, ec, ee, vmax);
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
    finish_tile(a, sc, vmax, lane, t);
}

}  // namespace device
}  // namespace rtm_hip
