// step_launch.hpp -- the fused step kernels (mappings in step_kernels.hip's
// header comment) and their launch dispatch, templated on the snapshot
// width: step_kernels_f32.hip and step_kernels_s16.hip each build one half
// (FP32 and 16-bit snapshots), so the two halves compile in parallel and
// neither kernel carries the other's snapshot code.
#pragma once

#include <hip/hip_ext.h>

#include "march_tile.hpp"
#include "step_kernels.hpp"

namespace rtm_hip {

namespace {

using namespace device;

// ---------------------------------------------------------------------------
// LDS-tile mapping: 256 threads, 64 depth cells x TX = 4*RX columns.
// ---------------------------------------------------------------------------
template <int H, int RX, Dir D, bool S16>
__global__ __launch_bounds__(kBlockThreads) void k_step(StepArgs a) {
    constexpr int TX = RX * kWaves;          // tile columns
    constexpr int LZ = kTileZ + 2 * H;       // LDS row length (depth + halo)
    constexpr int LX = TX + 2 * H;           // LDS rows (columns + halo)
    __shared__ float tile[LX * LZ];

    request_args(a);
    const int t = xcd_tile((int)blockIdx.x, a.ntiles);
    const int tz = t % a.ntz, tx = t / a.ntz;
    const int z0 = tz * kTileZ, x0 = tx * TX;
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int wave = (int)threadIdx.x / kTileZ;
    SnapCtx sc = snap_ctx<S16>(a, D);
    snap_prefetch<S16>(a, sc, lane);
    const int e0 = entry_bound(a, t), e1 = entry_bound(a, t + 1);   // in flight with the tile's loads

    const int iz = z0 + lane;
    const bool z_ok = iz >= H && iz < a.nze - H;
    const int jz = iz - a.nb;
    const bool z_int = z_ok && jz >= 0 && jz < a.nz;
    const int lx0 = wave * RX;               // first local column of this wave

    // ---- the streamed operands first, so their latency overlaps the staging
    float pv[RX], vv[RX], wxv[RX];
    ColumnOps im[RX];                       // imaging operands (fb, fb2, hv, iv, sv)
#pragma unroll
    for (int r = 0; r < RX; ++r) {
        const int ix = x0 + lx0 + r;
        const bool ok = z_ok && ix >= H && ix < a.nxe - H;
        const ptrdiff_t i = (ptrdiff_t)(ok ? ix : 0) * a.ldz + iz;
        pv[r] = (ok && !a.zero_prev) ? a.p_prev[i] : 0.0f;
        vv[r] = ok ? a.vdt2[i] : 0.0f;
        wxv[r] = ok ? a.wx[ix] : 0.0f;
        const int jx = ix - a.nb;
        const bool in = sc.imaging && ok && z_int && jx >= 0 && jx < a.nx;
        const ptrdiff_t j = (ptrdiff_t)jx * a.nz + jz;
        im[r].fb = in ? snapshot_bits<S16>(a, j, i) : 0u;
        im[r].iv = in ? a.image[j] : 0.0f;
        im[r].sv = in ? a.illum[j] : 0.0f;
        const bool pair = in && a.hold_field;
        if constexpr (S16) im[r].fb2 = pair ? (uint32_t)a.snap16_in2[j] : 0u;
        else im[r].fb2 = pair ? __float_as_uint(__builtin_nontemporal_load(a.snap_in2 + j)) : 0u;
        im[r].hv = pair ? a.hold_field[i] : 0.0f;
    }

    // ---- stage p^n (tile + halo) in LDS; out-of-grid cells read as 0 -------
    for (int r = wave; r < LX; r += kWaves) {
        const int gx = x0 - H + r;
        const bool xin = gx >= 0 && gx < a.nxe && !a.zero_cur;
        const float* row = a.p_cur + (ptrdiff_t)(xin ? gx : 0) * a.ldz;
        tile[r * LZ + H + lane] = xin ? row[z0 + lane] : 0.0f;
        if (lane < 2 * H) {
            const int hz = lane < H ? z0 - H + lane : z0 + kTileZ + (lane - H);
            const int col = lane < H ? lane : kTileZ + lane;
            const bool zin = hz >= 0 && hz < a.ldz;
            tile[r * LZ + col] = (xin && zin) ? row[zin ? hz : 0] : 0.0f;
        }
    }
    __syncthreads();

    // this lane's depth row of p^n across its columns and the x-halo
    float q[RX + 2 * H];
#pragma unroll
    for (int j = 0; j < RX + 2 * H; ++j) q[j] = tile[(lx0 + j) * LZ + H + lane];

    const float wzv = a.wz[iz];
    const TileEntries te = load_entries<D>(a, e0, e1, lane);
    snap_scales<S16>(a, sc, lane);
    entries_ready(te);
    float vmax = 0.0f;

#pragma unroll
    for (int r = 0; r < RX; ++r) {
        const int ix = x0 + lx0 + r;
        const bool ok = z_ok && ix >= H && ix < a.nxe - H;

        // every lane computes (the operands of non-output lanes are zeros
        // or in-tile LDS values) and takes part in the entry pass, whose
        // entries are read across lanes; only the stores are masked
        const float* T = tile + (lx0 + r + H) * LZ + H + lane;
        const float pc = q[r + H];
        float lap = a.coef.c0 * pc;
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            const float tx_k = a.coef.cx[k] * (q[r + H + k] + q[r + H - k]);
            const float tz_k = a.coef.cz[k] * (T[k] + T[-k]);
            lap = lap + (tx_k + tz_k);
        }

        const ptrdiff_t i = (ptrdiff_t)ix * a.ldz + iz;
        const float w = wxv[r] * wzv;
        const float prev = pv[r] * w;
        float out = (2.0f * pc - prev) + vv[r] * lap;
        out = out * w;
        out = apply_entries<D>(a, te, ok ? (int)i : -1, out, vv[r], e0, e1);
        if (!ok) continue;
        a.p_next[i] = out;

        const int jx = ix - a.nb;
        if (z_int && jx >= 0 && jx < a.nx) {
            const ptrdiff_t j = (ptrdiff_t)jx * a.nz + jz;
            if constexpr (D == Dir::Forward) {
                if constexpr (S16) {
                    if (sc.store_max) vmax = fmaxf(vmax, fabsf(out));
                    if (sc.quantise && !a.dry)
                        __builtin_nontemporal_store(encode_snapshot(pv[r], sc.k_q, a.snap_fmt), a.quant_out + j);
                } else {
                    if (sc.store32 && !a.dry) __builtin_nontemporal_store(out, a.snap_out + j);
                }
            } else {
                if (sc.imaging && !a.dry) {
                    float img, ill;
                    image_cell<S16>(a, sc.k_in2, im[r], snapshot_value<S16>(a, sc, im[r].fb, j), out, img, ill);
                    a.image[j] = img;
                    a.illum[j] = ill;
                }
            }
        }
    }
    finish_tile<S16>(a, sc, vmax, lane, t * kWaves + wave, a.ntiles * kWaves);
    settle_scales<S16>(a, sc, t * kWaves + wave, a.ntiles * kWaves, lane);
}

// March mapping: one wave per tile (march_tile.hpp).
template <int H, int TX, Dir D, bool DPP, bool S16, bool LEAN = false, int PH = 3>
__global__ __launch_bounds__(kBlockThreads) void k_march(StepArgs a) {
    request_args(a);
    // the wave index is wave-uniform: the tile's coordinates and column
    // offsets are then scalar arithmetic
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTileZ);
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (a.ntiles + kWaves - 1) / kWaves;
    const int t = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (t >= a.ntiles) return;               // whole wave leaves together
    march_tile<H, TX, D, DPP, S16, LEAN, PH>(a, a.coef, t, lane);
}


// Stream mapping: one wave per strip of XS columns (march_tile.hpp).
template <int H, int XS, Dir D, bool S16>
__global__ __launch_bounds__(kBlockThreads) void k_stream(StepArgs a) {
    request_args(a);
    const int wave = (int)threadIdx.x / kTileZ;
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (a.ntiles + kWaves - 1) / kWaves;
    const int t = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (t >= a.ntiles) return;               // whole wave leaves together
    stream_tile<H, XS, D, S16>(a, a.coef, t, lane);
}

// Launches a kernel; with events, through hipExtLaunchKernelGGL so that the
// events carry the kernel's own start and end timestamps.
#define RTM_LAUNCH(KERNEL)                                                                   \
    do {                                                                                     \
        if (t0) hipExtLaunchKernelGGL((KERNEL), grid, block, lds, s, t0, t1, 0, a);          \
        else hipLaunchKernelGGL((KERNEL), grid, block, lds, s, a);                           \
    } while (0)

template <int H, Dir D, bool S16>
hipError_t launch_h(const StepGeometry& g, const StepArgs& a, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    const size_t lds = g.family == Family::kLdsTile ? 0 : (size_t)g.lds_bytes;
    const dim3 block(kBlockThreads);
    const dim3 grid((unsigned)g.blocks);
    if (g.family == Family::kLdsTile) {
        switch (g.tile_x) {
            case 4:  RTM_LAUNCH((k_step<H, 1, D, S16>)); break;
            case 8:  RTM_LAUNCH((k_step<H, 2, D, S16>)); break;
            case 16: RTM_LAUNCH((k_step<H, 4, D, S16>)); break;
            default: return hipErrorInvalidValue;
        }
    } else if (g.family == Family::kStream) {
        switch (g.tile_x) {
            case 8:  RTM_LAUNCH((k_stream<H, 8, D, S16>)); break;
            case 16: RTM_LAUNCH((k_stream<H, 16, D, S16>)); break;
            case 32: RTM_LAUNCH((k_stream<H, 32, D, S16>)); break;
            default: return hipErrorInvalidValue;
        }
    } else if (g.family == Family::kMarch) {
        switch (g.tile_x) {
            case 4:  RTM_LAUNCH((k_march<H, 4, D, false, S16>)); break;
            case 8:  RTM_LAUNCH((k_march<H, 8, D, false, S16>)); break;
            case 16: RTM_LAUNCH((k_march<H, 16, D, false, S16>)); break;
            default: return hipErrorInvalidValue;
        }
    } else if (step_is_lean(D, a, S16)) {
        // the DPP march family has lean builds for steps without snapshot work
        switch (g.tile_x) {
            case 2:  RTM_LAUNCH((k_march<H, 2, D, true, S16, true>)); break;
            case 3:  RTM_LAUNCH((k_march<H, 3, D, true, S16, true>)); break;
            case 4:  RTM_LAUNCH((k_march<H, 4, D, true, S16, true>)); break;
            case 5:  RTM_LAUNCH((k_march<H, 5, D, true, S16, true>)); break;
            case 6:  RTM_LAUNCH((k_march<H, 6, D, true, S16, true>)); break;
            case 8:  RTM_LAUNCH((k_march<H, 8, D, true, S16, true>)); break;
            case 16: RTM_LAUNCH((k_march<H, 16, D, true, S16, true>)); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        // a 16-bit forward step with one of its two kinds of snapshot work
        // runs the build with only that one
#define RTM_MARCH_TILES(PHV)                                                        \
        switch (g.tile_x) {                                                         \
            case 2:  RTM_LAUNCH((k_march<H, 2, D, true, S16, false, PHV>)); break;  \
            case 3:  RTM_LAUNCH((k_march<H, 3, D, true, S16, false, PHV>)); break;  \
            case 4:  RTM_LAUNCH((k_march<H, 4, D, true, S16, false, PHV>)); break;  \
            case 5:  RTM_LAUNCH((k_march<H, 5, D, true, S16, false, PHV>)); break;  \
            case 6:  RTM_LAUNCH((k_march<H, 6, D, true, S16, false, PHV>)); break;  \
            case 8:  RTM_LAUNCH((k_march<H, 8, D, true, S16, false, PHV>)); break;  \
            case 16: RTM_LAUNCH((k_march<H, 16, D, true, S16, false, PHV>)); break; \
            default: return hipErrorInvalidValue;                                   \
        }
        if constexpr (D == Dir::Forward && S16) {
            const int ph = (a.wave_max != nullptr ? 1 : 0) | (a.quant_out != nullptr ? 2 : 0);
            if (ph == 1) {
                RTM_MARCH_TILES(1)
            } else if (ph == 2) {
                RTM_MARCH_TILES(2)
            } else {
                RTM_MARCH_TILES(3)
            }
        } else {
            RTM_MARCH_TILES(3)
        }
#undef RTM_MARCH_TILES
    }
    return hipGetLastError();
}

template <Dir D, bool S16>
hipError_t launch_dir(int half, const StepGeometry& g, const StepArgs& a, hipStream_t s, hipEvent_t t0,
                      hipEvent_t t1) {
    switch (half) {
        case 1: return launch_h<1, D, S16>(g, a, s, t0, t1);
        case 2: return launch_h<2, D, S16>(g, a, s, t0, t1);
        case 3: return launch_h<3, D, S16>(g, a, s, t0, t1);
        case 4: return launch_h<4, D, S16>(g, a, s, t0, t1);
        case 5: return launch_h<5, D, S16>(g, a, s, t0, t1);
        case 6: return launch_h<6, D, S16>(g, a, s, t0, t1);
        default: return hipErrorInvalidValue;
    }
}


template <bool S16>
hipError_t launch_step_t(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                         hipEvent_t start, hipEvent_t stop) {
    return dir == Dir::Forward ? launch_dir<Dir::Forward, S16>(half, g, a, stream, start, stop)
                               : launch_dir<Dir::Backward, S16>(half, g, a, stream, start, stop);
}

}  // namespace
}  // namespace rtm_hip
