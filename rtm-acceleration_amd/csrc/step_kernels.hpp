// step_kernels.hpp -- launch interface of the fused leapfrog step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace rtm_hip {

constexpr int kMaxHalf = 6;
constexpr int kTileZ = 64;        // one wave64 spans 64 consecutive depth cells
constexpr int kWaves = 4;         // waves per workgroup
constexpr int kBlockThreads = kTileZ * kWaves;
constexpr int kReduceWaves = 64;  // waves of a launch that share a snapshot-max reduction
// Columns of zeros after the last column of every field, of (v dt)^2 and of
// the x taper: a march tile's window ends at most tile_x - 1 <= 15 columns past
// the last column, so its column offsets need no clamping.
constexpr int kGuardCols = 16;


// Stencil coefficients pre-divided by the spacing (src/rtm/rtm_cpu.cpp:75-79).
struct Coeffs {
    float c0;
    float cx[kMaxHalf + 1];
    float cz[kMaxHalf + 1];
};

// An injection/recording entry of the per-tile table: the extended-grid cell
// (device index ix*ldz+iz) and a code: -1 = add vdt2*wavelet[it] (source),
// >= 0 = receiver ir (forward: record the cell into rec_out[ir*rec_stride+it];
// backward: add vdt2 times receiver ir's sample at it).
struct Entry {
    int cell;
    int code;
};

enum class Dir : int { Forward = 0, Backward = 1 };

// Snapshot storage: FP32 (the reference's), or 16-bit with one FP32 value per
// snapshot, max|p| over the interior, from which an exact power-of-two scale
// 2^k is derived (snapshot_scale_exponent).
enum class SnapFmt : int { F32 = 0, F16 = 1, BF16 = 2 };

struct StepArgs {
    // fields: p_prev holds p^{n-1} damped once; p_cur p^n; p_next p^{n+1}
    const float* p_prev;
    const float* p_cur;
    float*       p_next;
    const float* vdt2;        // [nxe][ldz]
    const float* wx;          // [nxe]  Cerjan taper along x
    const float* wz;          // [ldz]  Cerjan taper along z (0 in the pitch pad)
    Coeffs       coef;
    int nxe, nze, ldz, nb, nx, nz;
    int ntz, ntx, ntiles;     // tile grid of the launch (StepGeometry)
    unsigned ntz_magic;       // tile_div_magic(ntz)
    int it;
    // per-tile injection / recording table (CSR over tiles)
    const int*   tile_ptr;    // [ntiles+1]
    const Entry* entries;
    const float* wavelet;     // [nt]   forward source amplitude
    // backward: receiver amplitudes in entry order, time-major:
    // amps[it * amp_count + e] is the amplitude of entry e at time it
    // (launch_gather_amps), so a wave loads its entries' amplitudes with
    // the entries themselves, not after them
    const float* amps;
    int          amp_count;
    float*       rec_out;     // forward modelling: trace-major, stride rec_stride
    long long    rec_stride;
    int          snap_fmt;    // SnapFmt
    // forward, stored step, FP32: the new field's interior into snap_out
    float*       snap_out;
    // F16/BF16 snapshots.  The scale of the snapshot of stored forward step
    // `it` (max|p| over the interior) is settled inside the next two step
    // launches, with no launch of its own:
    //   step it:   wave w of the launch writes max|p^{n+1}| over its interior
    //              cells to wave_max[w]; wave 0 writes the launch's wave
    //              count to *wave_count
    //   step it+1: waves w < min(kReduceWaves, waves of the launch) each
    //              reduce a slice of red_in[0 .. *red_count) (clamped to
    //              parts_cap) to red_out[w]; wave 0 writes the slice count
    //              to *red_out_count
    //   step it+2: every wave reduces quant_parts[0 .. *quant_count) to the
    //              snapshot's max, quantises p^{n-1} (the field of step it,
    //              which it loads anyway) into quant_out, and wave 0 writes
    //              the max to *quant_max_out for the backward decode
    float*       wave_max;
    int*         wave_count;
    const float* red_in;
    const int*   red_count;
    float*       red_out;
    int*         red_out_count;
    const float* quant_parts;
    const int*   quant_count;
    uint16_t*    quant_out;
    float*       quant_max_out;
    int          parts_cap;
    // backward, stored step: image the new field against the snapshot
    // (snap_in for FP32; snap16_in + *snap16_max for F16/BF16), or no imaging
    const float*    snap_in;
    const uint16_t* snap16_in;
    const float*    snap16_max;
    // paired imaging (pair_step): this step images two stored steps in one
    // pass over the image, first the earlier one in backward order, whose
    // field was kept in hold_field and whose snapshot is snap_in2 (FP32) or
    // snap16_in2 with max *snap16_max2, then its own; hold_field is null on
    // every other step
    const float*    hold_field;
    const float*    snap_in2;
    const uint16_t* snap16_in2;
    const float*    snap16_max2;
    float*       image;
    float*       illum;
    // batched phases: the first steps of a phase read p_prev/p_cur as zero
    // instead of their buffers (no reset pass between phases); a 16-bit
    // snapshot of the forward phase's last step is imaged, at the backward
    // phase's first step, straight from pending_field (its extended field),
    // with the max reduced from quant_parts[0 .. *quant_count); dry runs
    // write only p_next: no image, illumination or snapshot update (timing
    // candidates on live state)
    int          zero_prev, zero_cur;
    const float* pending_field;
    int          dry;
};

// Paired imaging of the backward phase.  Every stored step (it % store == 0)
// images I += f b, S += f^2 with its snapshot f and its field b, in
// descending it.  Imaging two stored steps in one pass reads and writes I and
// S once instead of twice: from the stored step `first` down, consecutive
// stored steps pair up; the first of a pair (Hold) writes its field into hold
// buffer `hold` and images nothing, the second (Pair), store steps later,
// images the held field with the earlier snapshot and then its own, in the
// reference's order, so the sums are unchanged.  A stored step above `first`
// (a batched job's pending 16-bit snapshot) or a last one without a partner
// images alone; first < 0 turns pairing off.  Two hold buffers alternate by
// pair, so a Hold never overwrites a field still read (store = 1 and 2
// included).
enum class PairRole : int { None = 0, Alone = 1, Hold = 2, Pair = 3 };
struct PairStep {
    PairRole role;
    int hold;      // hold buffer written (Hold) or read (Pair)
};
__host__ __device__ inline PairStep pair_step(int it, int store, int first) {
    PairStep p{PairRole::None, 0};
    if (it % store) return p;
    p.role = PairRole::Alone;
    if (first < 0 || it > first) return p;
    const int c = (first - it) / store;              // stored steps since `first`
    if (c & 1) {
        p.role = PairRole::Pair;
        p.hold = ((c - 1) >> 1) & 1;
    } else if (it - store >= 0) {
        p.role = PairRole::Hold;
        p.hold = (c >> 1) & 1;
    }
    return p;
}
// The first stored step of the backward phase that can be held: the top
// stored step, or the one below it when the top one is imaged from its
// still-pending field (pending_top).
__host__ __device__ inline int pair_first(int nt, int store, bool pending_top) {
    const int top = (nt - 1) / store * store;
    return pending_top && top == nt - 1 ? top - store : top;
}

// Whether a step has no snapshot work, so a lean build of the step (one
// without the snapshot epilogue of its cells) can run it: a backward step
// that does not image, a forward step that neither stores an FP32 snapshot
// nor the maxima or codes of a 16-bit one (the reduction of the maxima, red_in,
// runs in the lean build too).  The same tests as the kernels' own
// (snap_ctx), on the host.
__host__ __device__ inline bool step_is_lean(Dir d, const StepArgs& a, bool s16) {
    if (d == Dir::Forward) return a.snap_out == nullptr && a.wave_max == nullptr && a.quant_out == nullptr;
    return s16 ? (a.snap16_in == nullptr && a.pending_field == nullptr) : a.snap_in == nullptr;
}

// Kernel families of the fused step.
//   kLdsTile: a 256-thread workgroup stages a 64 x tile_x tile of p^n (+halo)
//             in LDS; tiles start at cell (0, 0) and cover the padded grid.
//   kMarch:   each wave64 owns a depth strip of 64 cells whose 64 - 2H middle
//             lanes are outputs (strips overlap by 2H), and tile_x columns;
//             it loads every operand of its tile up front (no LDS, no
//             barrier) and takes depth neighbours from other lanes through
//             ds_bpermute (kMarch) or chained DPP lane shifts (kMarchDpp).
//   kStream:  the march strip marching along tile_x columns with p^n in a
//             sliding register window (DPP depth neighbours).
enum class Family : int { kLdsTile = 0, kMarch = 1, kMarchDpp = 2, kStream = 3 };

// Tile grid of one launch.  Tiles are numbered x-major over (ntz x ntx):
// tile = tx * ntz + tz; the injection table is indexed the same way.
struct StepGeometry {
    Family family = Family::kMarchDpp;
    int tile_x = 4;      // output columns per tile
    int tile_z = 64;     // output depth cells per tile
    int x_origin = 0;    // first column of tile 0
    int z_origin = 0;    // first depth cell of tile 0
    int ntz = 0, ntx = 0, ntiles = 0;
    unsigned ntz_magic = 0;   // tile_div_magic(ntz)
    int blocks = 0;      // workgroups launched
    // occupancy cap: workgroups per CU (0 = none), enforced by reserving
    // lds_bytes of dynamic LDS per workgroup (the march kernels use none)
    int blocks_per_cu = 0;
    int lds_bytes = 0;
};

// Dynamic LDS per workgroup that lets at most `blocks_per_cu` workgroups of a
// kernel without LDS reside on one CU of `device`; 0 when no cap is asked or
// the device cannot give a workgroup that much.
int occupancy_lds_bytes(int device, int blocks_per_cu);

StepGeometry step_geometry(Family family, int half, int tile_x, int nxe, int nze, int ldz);

// t / ntz for a wave-uniform tile index t, as the march kernels compute it:
// umulhi(t, magic) with magic = ceil(2^32 / ntz), exact whenever t * ntz < 2^32
// (t < ntz * ntx); ntz == 1 needs no division and has magic 0.
inline unsigned tile_div_magic(int ntz) {
    return ntz <= 1 ? 0u : (unsigned)((0x100000000ull + (unsigned)ntz - 1) / (unsigned)ntz);
}

// Whether the step kernels can address an extended grid of nxe x ldz cells
// and an interior of nx x nz: every field and interior array fits a 31-bit
// byte range (buffer resources), and the tile index times the strip count
// stays below 2^32 for the narrowest tiles (tile_div_magic).
inline bool step_grid_ok(int nxe, int nze, int ldz, int nx, int nz) {
    const unsigned long long lim = 1ull << 31;
    if ((unsigned long long)(nxe + kGuardCols) * ldz * 4 >= lim || (unsigned long long)nx * nz * 4 >= lim) return false;
    const unsigned long long ntz = (unsigned long long)nze / 50 + 1;   // strips of >= 52 cells
    return ntz * ntz * (unsigned long long)nxe < (1ull << 32);
}

// Tile widths each family is instantiated for.  The odd widths of the DPP
// march let the tile count of a grid fall just under a multiple of the
// chip's resident-wave capacity instead of just over it.
inline std::vector<int> tile_widths(Family f) {
    switch (f) {
        case Family::kMarchDpp: return {4, 3, 2, 5, 6, 8, 16};
        case Family::kStream: return {8, 16, 32};
        default: return {4, 8, 16};
    }
}

inline int tile_of_cell(const StepGeometry& g, int ixe, int ize) {
    return ((ixe - g.x_origin) / g.tile_x) * g.ntz + (ize - g.z_origin) / g.tile_z;
}

// The two builds of the fused step kernels (step_launch.hpp): FP32 and
// 16-bit snapshot arguments; launch_step picks one by a.snap_fmt.
hipError_t launch_step_f32(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                           hipEvent_t start, hipEvent_t stop);
hipError_t launch_step_s16(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                           hipEvent_t start, hipEvent_t stop);

// Launch one fused step on `stream` with the geometry from step_geometry.
// With start/stop events, they record the kernel's own start and end.
hipError_t launch_step(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                       hipEvent_t start = nullptr, hipEvent_t stop = nullptr);

// amps[it * nent + e] = traces[ent[e].code * trace_stride + it] for it < nt,
// e < nent: the backward entry table's receiver amplitudes in entry order.
hipError_t launch_gather_amps(const Entry* ent, int nent, const float* traces, long long trace_stride, int nt,
                              float* amps, hipStream_t stream);

// For a 16-bit stored step among the last two of the forward loop (no step
// it+2 settles it): *out = max(parts[0..n)) over its per-wave maxima or
// slice maxima (one workgroup), then launch_quantise below.
hipError_t launch_reduce_max(const float* parts, int n, float* out, hipStream_t stream);
// Quantise the interior of field p (extended layout) into a 16-bit snapshot
// with the scale of *max.
hipError_t launch_quantise(const float* p, int ldz, int nb, int nx, int nz, SnapFmt fmt, const float* max,
                           uint16_t* out, hipStream_t stream);

// image/illum (FP32, n values) added into the FP64 stacks, then reset to zero.
hipError_t launch_fold_fp64(float* image, float* illum, double* stack_i, double* stack_s, size_t n,
                            hipStream_t stream);

// Snapshot scale exponent k: |v| * 2^k < 2^15 for every |v| <= maxabs (0 when
// maxabs is 0), so 16-bit codes hold v * 2^k and decode exactly as code * 2^-k.
__host__ __device__ inline int snapshot_scale_exponent(float maxabs) {
    if (!(maxabs > 0.0f)) return 0;
    int e = 0;
    (void)frexpf(maxabs, &e);      // maxabs = m 2^e, m in [0.5, 1)
    return 15 - e;
}

// Decoding of a 16-bit snapshot code to the stored float (exact).
__host__ __device__ inline float decode_snapshot(uint16_t code, int k, SnapFmt fmt) {
    float f;
    if (fmt == SnapFmt::BF16) {
        const uint32_t bits = (uint32_t)code << 16;
        __builtin_memcpy(&f, &bits, 4);
    } else {
        // IEEE binary16 -> binary32
        const uint32_t s = (uint32_t)(code >> 15) << 31;
        const uint32_t e = (code >> 10) & 0x1f;
        const uint32_t m = code & 0x3ff;
        if (e == 0) {
            f = ldexpf((float)m, -24);           // zero or subnormal: m * 2^-24
            if (s) f = -f;
        } else {
            const uint32_t bits = s | ((e == 31 ? 255u : e - 15 + 127) << 23) | (m << 13);
            __builtin_memcpy(&f, &bits, 4);
        }
    }
    return ldexpf(f, -k);
}

}  // namespace rtm_hip
