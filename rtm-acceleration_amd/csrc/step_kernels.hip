// step_kernels.hip -- the fused leapfrog step for gfx950 (CDNA4, wave64).
//
// One launch advances one shot by one time step and does, in a single pass
// over the extended grid, what the reference does in up to five passes:
//
//   fd_time_step      src/rtm/forward.cpp:29-50    order-2H Laplacian + leapfrog
//   apply_sponge      src/rtm/forward.cpp:56-63    Cerjan damping of p^n and p^{n+1}
//   inject_source     src/rtm/forward.cpp:70-72    forward
//   inject_receivers  src/rtm/backward.cpp:20-27   backward
//   record_traces     src/rtm/forward.cpp:75-79    forward modelling
//   save_snapshot     src/rtm/forward.cpp:83-88    forward, stored steps
//   imaging           src/rtm/imaging.cpp:25-40    backward, stored steps
//
// Exactness.  Every float operation of the reference is performed once, in
// the reference's order, with no contraction (built -ffp-contract=off) and
// with FP32 denormals preserved, so the result is bit-identical:
//   lap = c0*p; lap = lap + (cx_k*(x+ + x-) + cz_k*(z+ + z-)),  k = 1..H
//   p^{n+1} = ((2 p^n - p^{n-1}*w) + vdt2*lap) * w,   w = wx[ix]*wz[iz]
// The reference damps p^n in place after the stencil (forward.cpp:56-63); that
// multiply is applied here when the buffer is read back as p^{n-1} on the next
// step (p_prev*w), so no thread ever writes a cell another workgroup reads in
// the same launch.  Injection and recording cells are interior cells, where
// w == 1.0f exactly, so injecting after the damped write is the reference's
// inject-then-damp.  Duplicate receiver cells add in ascending receiver order
// through a per-tile entry table, which is the reference's sequential order.
//
// 16-bit snapshots (no reference counterpart).  On a stored step each wave
// writes the max of |p^{n+1}| over its interior cells to a per-wave slot and a
// one-workgroup kernel reduces the slots to the snapshot's max (plain stores
// and one small launch; one atomic max per wave on a shared address took
// ~60 us per stored step on a 5.7 M-cell grid); the next step's kernel, which
// reads that field anyway as p^n, writes code = round(p^n * 2^k) to the
// 16-bit store, with
// 2^k the power of two that keeps |code| < 2^15.  Scaling by 2^k is exact, so
// the only error is the FP16/BF16 rounding; decoding is code * 2^-k, exact.
//
// Three mappings (Family in step_kernels.hpp) share the arithmetic above.
//
// kMarch / kMarchDpp (default).  Each wave64 owns one tile: a depth strip of
// 64 consecutive cells, of which the middle 64 - 2H lanes are outputs (strips
// overlap by 2H cells), times tile_x columns.  A lane loads its depth row of
// p^n over the tile_x + 2H columns the tile needs, and every other operand of
// its tile_x output columns, before any arithmetic: up to 3*tile_x + 2H
// independent 256-byte row loads in flight per wave, no LDS, no barrier.
// x-neighbours are in the lane's own registers; z-neighbours come from the
// lanes k above/below through ds_bpermute or chained DPP lane shifts.  The
// step is bound by the bytes in flight, not by instruction issue, so this
// layout trades a 2H/64 depth overlap and a 2H/tile_x column overlap (both
// re-read from L2) for one memory round trip per wave.
//
// kStream.  The march strip, marching along tile_x (8, 16 or 32) columns in
// chunks of 4 with p^n in a sliding register window of 4 + 2H columns: each
// column of p^n is loaded once per strip instead of (tile_x + 2H)/tile_x
// times, and the next chunk's loads are in flight while a chunk computes.
//
// kLdsTile.  A 256-thread workgroup (4 x wave64) owns 64 depth cells x tile_x
// columns; the tile of p^n plus its halo is staged in LDS with 256-byte row
// loads (ldz is a multiple of 64 floats) and read back for both directions.
//
// In both, workgroups are numbered so that each of the 8 XCDs gets a
// contiguous band of tiles, keeping halo re-reads inside one XCD's L2.
#include "step_kernels.hpp"

#include <hip/hip_ext.h>

#include "march_tile.hpp"

namespace rtm_hip {

namespace {

using namespace device;

// ---------------------------------------------------------------------------
// LDS-tile mapping: 256 threads, 64 depth cells x TX = 4*RX columns.
// ---------------------------------------------------------------------------
template <int H, int RX, Dir D>
__global__ __launch_bounds__(kBlockThreads) void k_step(StepArgs a) {
    constexpr int TX = RX * kWaves;          // tile columns
    constexpr int LZ = kTileZ + 2 * H;       // LDS row length (depth + halo)
    constexpr int LX = TX + 2 * H;           // LDS rows (columns + halo)
    __shared__ float tile[LX * LZ];

    const int t = xcd_tile((int)blockIdx.x, a.ntiles);
    const int tz = t % a.ntz, tx = t / a.ntz;
    const int z0 = tz * kTileZ, x0 = tx * TX;
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int wave = (int)threadIdx.x / kTileZ;
    const SnapCtx sc = snap_ctx(a, D);

    const int iz = z0 + lane;
    const bool z_ok = iz >= H && iz < a.nze - H;
    const int jz = iz - a.nb;
    const bool z_int = z_ok && jz >= 0 && jz < a.nz;
    const int lx0 = wave * RX;               // first local column of this wave

    // ---- the streamed operands first, so their latency overlaps the staging
    float pv[RX], vv[RX], wxv[RX], fv[RX], iv[RX], sv[RX];
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
        fv[r] = in ? snapshot_value(a, sc, j, i) : 0.0f;
        iv[r] = in ? a.image[j] : 0.0f;
        sv[r] = in ? a.illum[j] : 0.0f;
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
    const int e0 = a.tile_ptr[t], e1 = a.tile_ptr[t + 1];
    const TileEntries te = load_entries<D>(a, e0, e1, lane);
    float vmax = 0.0f;

#pragma unroll
    for (int r = 0; r < RX; ++r) {
        const int ix = x0 + lx0 + r;
        if (!z_ok || ix < H || ix >= a.nxe - H) continue;

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
        out = apply_entries<D>(a, te, (int)i, out, vv[r], e0, e1);
        a.p_next[i] = out;

        const int jx = ix - a.nb;
        if (z_int && jx >= 0 && jx < a.nx) {
            const ptrdiff_t j = (ptrdiff_t)jx * a.nz + jz;
            if constexpr (D == Dir::Forward) {
                if (sc.store32) a.snap_out[j] = out;
                if (sc.store_max) vmax = fmaxf(vmax, fabsf(out));
                if (sc.quantise) a.quant_out[j] = encode_snapshot(pc, sc.k_q, a.snap_fmt);
            } else {
                if (sc.imaging && !a.dry) {
                    a.image[j] = iv[r] + fv[r] * out;
                    a.illum[j] = sv[r] + fv[r] * fv[r];
                }
            }
        }
    }
    if (sc.store_max) {
        vmax = wave_max(vmax);
        if (lane == 0) a.wave_max[t * kWaves + wave] = vmax;
    }
}

// March mapping: one wave per tile (march_tile.hpp).
template <int H, int TX, Dir D, bool DPP>
__global__ __launch_bounds__(kBlockThreads) void k_march(StepArgs a) {
    const int wave = (int)threadIdx.x / kTileZ;
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (a.ntiles + kWaves - 1) / kWaves;
    const int t = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (t >= a.ntiles) return;               // whole wave leaves together
    march_tile<H, TX, D, DPP>(a, a.coef, t, lane);
}

// Stream mapping: one wave per strip of XS columns (march_tile.hpp).
template <int H, int XS, Dir D>
__global__ __launch_bounds__(kBlockThreads) void k_stream(StepArgs a) {
    const int wave = (int)threadIdx.x / kTileZ;
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (a.ntiles + kWaves - 1) / kWaves;
    const int t = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (t >= a.ntiles) return;               // whole wave leaves together
    stream_tile<H, XS, D>(a, a.coef, t, lane);
}

// Quantise the interior of an extended field into a 16-bit snapshot.
__global__ __launch_bounds__(256) void k_quantise(const float* p, int ldz, int nb, int nx, int nz, int fmt,
                                                  const float* maxv, uint16_t* out) {
    const size_t n = (size_t)nx * nz;
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int jx = (int)(j / (size_t)nz), jz = (int)(j % (size_t)nz);
    const int k = snapshot_scale_exponent(*maxv);
    out[j] = encode_snapshot(p[(ptrdiff_t)(jx + nb) * ldz + jz + nb], k, fmt);
}

// *out = max(parts[0..n)), one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_reduce_max(const float* parts, int n, float* out) {
    float m = 0.0f;
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) m = fmaxf(m, parts[i]);
    m = wave_max(m);
    __shared__ float wm[1024 / kTileZ];
    const int w = (int)threadIdx.x / kTileZ;
    if ((threadIdx.x & (kTileZ - 1)) == 0) wm[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (int k = 0; k < (int)blockDim.x / kTileZ; ++k) r = fmaxf(r, wm[k]);
        *out = r;
    }
}

// FP64 cross-shot stack: stack += (double)image; image = 0.
__global__ __launch_bounds__(256) void k_fold(float* image, float* illum, double* si, double* ss, size_t n) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    si[j] += (double)image[j];
    ss[j] += (double)illum[j];
    image[j] = 0.0f;
    illum[j] = 0.0f;
}

// Launches a kernel; with events, through hipExtLaunchKernelGGL so that the
// events carry the kernel's own start and end timestamps.
#define RTM_LAUNCH(KERNEL)                                                                   \
    do {                                                                                     \
        if (t0) hipExtLaunchKernelGGL((KERNEL), grid, block, lds, s, t0, t1, 0, a);          \
        else hipLaunchKernelGGL((KERNEL), grid, block, lds, s, a);                           \
    } while (0)

template <int H, Dir D>
hipError_t launch_h(const StepGeometry& g, const StepArgs& a, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    const size_t lds = g.family == Family::kLdsTile ? 0 : (size_t)g.lds_bytes;
    const dim3 block(kBlockThreads);
    const dim3 grid((unsigned)g.blocks);
    if (g.family == Family::kLdsTile) {
        switch (g.tile_x) {
            case 4:  RTM_LAUNCH((k_step<H, 1, D>)); break;
            case 8:  RTM_LAUNCH((k_step<H, 2, D>)); break;
            case 16: RTM_LAUNCH((k_step<H, 4, D>)); break;
            default: return hipErrorInvalidValue;
        }
    } else if (g.family == Family::kStream) {
        switch (g.tile_x) {
            case 8:  RTM_LAUNCH((k_stream<H, 8, D>)); break;
            case 16: RTM_LAUNCH((k_stream<H, 16, D>)); break;
            case 32: RTM_LAUNCH((k_stream<H, 32, D>)); break;
            default: return hipErrorInvalidValue;
        }
    } else if (g.family == Family::kMarch) {
        switch (g.tile_x) {
            case 4:  RTM_LAUNCH((k_march<H, 4, D, false>)); break;
            case 8:  RTM_LAUNCH((k_march<H, 8, D, false>)); break;
            case 16: RTM_LAUNCH((k_march<H, 16, D, false>)); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (g.tile_x) {
            case 2:  RTM_LAUNCH((k_march<H, 2, D, true>)); break;
            case 3:  RTM_LAUNCH((k_march<H, 3, D, true>)); break;
            case 4:  RTM_LAUNCH((k_march<H, 4, D, true>)); break;
            case 5:  RTM_LAUNCH((k_march<H, 5, D, true>)); break;
            case 6:  RTM_LAUNCH((k_march<H, 6, D, true>)); break;
            case 8:  RTM_LAUNCH((k_march<H, 8, D, true>)); break;
            case 16: RTM_LAUNCH((k_march<H, 16, D, true>)); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

template <Dir D>
hipError_t launch_dir(int half, const StepGeometry& g, const StepArgs& a, hipStream_t s, hipEvent_t t0,
                      hipEvent_t t1) {
    switch (half) {
        case 1: return launch_h<1, D>(g, a, s, t0, t1);
        case 2: return launch_h<2, D>(g, a, s, t0, t1);
        case 3: return launch_h<3, D>(g, a, s, t0, t1);
        case 4: return launch_h<4, D>(g, a, s, t0, t1);
        case 5: return launch_h<5, D>(g, a, s, t0, t1);
        case 6: return launch_h<6, D>(g, a, s, t0, t1);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

int occupancy_lds_bytes(int device, int blocks_per_cu) {
    if (blocks_per_cu <= 0) return 0;
    int per_cu = 0, per_block = 0;
    if (hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return 0;
    const int lds = per_cu / blocks_per_cu;     // blocks_per_cu fit, one more does not
    return lds <= per_block ? lds : 0;
}

StepGeometry step_geometry(Family family, int half, int tile_x, int nxe, int nze, int ldz) {
    StepGeometry g;
    g.family = family;
    g.tile_x = tile_x;
    if (family == Family::kLdsTile) {
        g.tile_z = kTileZ;
        g.x_origin = 0;
        g.z_origin = 0;
        g.ntz = ldz / kTileZ;
        g.ntx = (nxe + tile_x - 1) / tile_x;
        g.ntiles = g.ntz * g.ntx;
        g.blocks = g.ntiles;
    } else {
        // output cells [H, nze-H) x [H, nxe-H); strip tz covers depth
        // [H + tz*(64-2H), ...) and loads [tz*(64-2H), tz*(64-2H) + 64)
        g.tile_z = kTileZ - 2 * half;
        g.x_origin = half;
        g.z_origin = half;
        g.ntz = (nze - 2 * half + g.tile_z - 1) / g.tile_z;
        g.ntx = (nxe - 2 * half + tile_x - 1) / tile_x;
        g.ntiles = g.ntz * g.ntx;
        g.blocks = (g.ntiles + kWaves - 1) / kWaves;
    }
    return g;
}

hipError_t launch_step(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                       hipEvent_t start, hipEvent_t stop) {
    return dir == Dir::Forward ? launch_dir<Dir::Forward>(half, g, a, stream, start, stop)
                               : launch_dir<Dir::Backward>(half, g, a, stream, start, stop);
}

hipError_t launch_reduce_max(const float* parts, int n, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, stream, parts, n, out);
    return hipGetLastError();
}

hipError_t launch_quantise(const float* p, int ldz, int nb, int nx, int nz, SnapFmt fmt, const float* max,
                           uint16_t* out, hipStream_t stream) {
    const size_t n = (size_t)nx * nz;
    hipLaunchKernelGGL(k_quantise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, ldz, nb, nx, nz,
                       (int)fmt, max, out);
    return hipGetLastError();
}

hipError_t launch_fold_fp64(float* image, float* illum, double* stack_i, double* stack_s, size_t n,
                            hipStream_t stream) {
    hipLaunchKernelGGL(k_fold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, image, illum, stack_i,
                       stack_s, n);
    return hipGetLastError();
}

}  // namespace rtm_hip
