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
// step (p_prev*w), so no step writes a cell of p^n, which other workgroups
// read as their halo.  p^{n-1} is used only at a step's own output cells, in
// the lane that writes p^{n+1} there, so p_next may be p_prev (an in-place
// update): march strips also load p^{n-1} in their halo rows, but never use
// those values.  Injection and recording cells are interior cells, where
// w == 1.0f exactly, so injecting after the damped write is the reference's
// inject-then-damp.  Duplicate receiver cells add in ascending receiver order
// through a per-tile entry table, which is the reference's sequential order.
//
// 16-bit snapshots (no reference counterpart).  The snapshot of stored step
// `it` is scaled by its max |p| over the interior, which needs every wave of
// the launch; the max is settled inside the next two step launches, with no
// launch of its own (a kernel boundary costs several microseconds on this
// chip, a tenth of a step): on step it each wave writes its max to a per-wave
// slot, on step it+1 up to 64 waves reduce the slots in slices, and on step
// it+2 every wave reduces the slices, then writes code = round(p^{n-1} 2^k)
// (p^{n-1} of that step is the field of step it, loaded anyway) to the
// 16-bit store, with
// 2^k the power of two that keeps |code| < 2^15.  Scaling by 2^k is exact, so
// the only error is the FP16/BF16 rounding; decoding is code * 2^-k, exact.
// A stored step among the last two of the forward phase has no step it+2: the
// per-shot engine settles it with a reduce and a quantise launch after the
// loop, the batch with one bookkeeping tick (batch_kernels.hip).
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

#include "march_tile.hpp"

namespace rtm_hip {

namespace {

using namespace device;

// Receiver amplitudes in backward-entry order, time-major (StepArgs::amps).
// Threads run along entries, so the writes are contiguous.
__global__ __launch_bounds__(256) void k_gather_amps(const Entry* ent, int nent, const float* traces,
                                                     long long stride, int nt, float* amps) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)nt * nent) return;
    const int it = (int)(k / nent), e = (int)(k % nent);
    amps[k] = traces[(long long)ent[e].code * stride + it];
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
    const float m = block_max<1024>(parts, n);
    if (threadIdx.x == 0) *out = m;
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
    g.ntz_magic = tile_div_magic(g.ntz);
    return g;
}

hipError_t launch_step(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                       hipEvent_t start, hipEvent_t stop) {
    return a.snap_fmt == (int)SnapFmt::F32 ? launch_step_f32(dir, half, g, a, stream, start, stop)
                                           : launch_step_s16(dir, half, g, a, stream, start, stop);
}

hipError_t launch_gather_amps(const Entry* ent, int nent, const float* traces, long long trace_stride, int nt,
                              float* amps, hipStream_t stream) {
    const long long n = (long long)nt * nent;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_amps, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ent, nent, traces,
                       trace_stride, nt, amps);
    return hipGetLastError();
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
