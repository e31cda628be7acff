// K3/K4: the O(N) kernels of one implicit step, all resident on the device.
//   residual       Simulation::calculateG (simulation.cpp:189-192)
//   newton update  Simulation::integrate (simulation.cpp:78-81)
//   error control  Simulation::calculateXError + PrecisionHandler::updateError
//                  / getNewStepSize / reset + max-stepsize clamp
//                  (simulation.cpp:398-405, 516, 643-649;
//                  precision_handler.cpp:62-119)
//   sums           the log reductions of simulation.cpp:460-461, 508-509,
//                  551, 557 and utility.h:92-100: in the reference's
//                  sequential order with the exact and ordered field (the
//                  logged values match bit for bit), as a fixed-order tree
//                  with the fast field.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "context.h"
#include "field.h"
#include "step_core.h"

namespace sdamd {

namespace {

// g and, for the Krylov solve that follows, ||g||^2: each 256-lane
// workgroup adds its lanes' squares by a wavefront butterfly and its four
// wavefronts in order into part[block]; the last workgroup to take the
// ticket adds the partials in block order (lanes strided by 64, then a
// butterfly) -- the order of the solver's own reduction over 256-lane
// blocks, so ||g||^2 is the value that reduction would give.
__global__ __launch_bounds__(256) void k_residual(int n, double h, const double *__restrict__ xn,
                                                  const double *__restrict__ xo, const double *__restrict__ csp,
                                                  const double *__restrict__ isp, const double *__restrict__ d,
                                                  double *__restrict__ g, double *__restrict__ part,
                                                  unsigned *__restrict__ ticket, double *__restrict__ norm2)
{
    __shared__ double wsum[4];
    __shared__ bool last;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    double sq = 0.0;
    if (i < n) {
        const double v = stepf::residual(xn[i], xo[i], csp[i], isp[i], d[i], h);
        g[i] = v;
        sq = 0.0 + v * v;
    }
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
    if (lane == 0) wsum[wave] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last || wave != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double t = 0.0;
    for (int b = lane; b < (int)gridDim.x; b += 64) t += __builtin_nontemporal_load(&part[b]);
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) {
        *norm2 = t;
        *ticket = 0u;
    }
}

__global__ void k_newton_update(int n, double *__restrict__ x, const double *__restrict__ delta)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    x[i] -= delta[i];
}

// precision_handler.cpp:103-119 + simulation.cpp:646-649
__global__ __launch_bounds__(1024) void k_error_control(int n, const double *__restrict__ xbig,
                                                        const double *__restrict__ xsmall2, double *__restrict__ tol,
                                                        double minPrecSqr, double h, double hmax,
                                                        DevScalars *__restrict__ out)
{
    __shared__ double smax[1024];
    __shared__ int sidx[1024];
    const int t = (int)threadIdx.x;
    double best = 0.0;
    int bidx = 0x7fffffff;
    for (int i = t; i < n; i += 1024) {
        const double e = __builtin_fabs(xbig[i] - xsmall2[i]);
        if (0.0 < e) {
            const double r = e * e / tol[i];
            if (r > best) {   // NaN never wins, as in updateError
                best = r;
                bidx = i;
            }
        }
        tol[i] = minPrecSqr;
    }
    smax[t] = best;
    sidx[t] = bidx;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off) {
            const double o = smax[t + off];
            const int oi = sidx[t + off];
            if (o > smax[t] || (o == smax[t] && oi < sidx[t])) {
                smax[t] = o;
                sidx[t] = oi;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        const double m = smax[0];
        out->max_ratio = m;
        out->selected = (m > 0.0) ? sidx[0] : 0;
        out->accept = m < 1.0 ? 1 : 0;
        out->h_next = stepf::new_step_size(m, h, hmax);
    }
}

// Sequential sums in index order (std::accumulate semantics).  The
// workgroup stages 2048 values at a time in LDS with coalesced loads; one
// lane then runs the three dependent chains from LDS.

constexpr int kSumChunk = 2048;

// Sequential sums in index order (the bits of the reference's loops): the
// workgroup stages 2048 values at a time in LDS, one lane adds them.  Used
// with the exact and ordered field, whose speeds carry the reference's bits.
__global__ __launch_bounds__(256) void k_speed_sums(int n, const double *__restrict__ v,
                                                    const double *__restrict__ b, DevScalars *__restrict__ out)
{
    __shared__ double sv[kSumChunk], sb[kSumChunk];
    double sa = 0.0, sq = 0.0, sbv = 0.0;
    for (int base = 0; base < n; base += kSumChunk) {
        const int cnt = min(kSumChunk, n - base);
        for (int k = (int)threadIdx.x; k < cnt; k += 256) {
            sv[k] = v[base + k];
            sb[k] = b[base + k];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll 8
            for (int k = 0; k < cnt; ++k) {
                const double vi = sv[k];
                sa = sa + __builtin_fabs(vi);
                sq = sq + vi * vi;
                sbv += sb[k] * vi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out->sum_abs = sa;
        out->sum_sq = sq;
        out->sum_bv = sbv;
    }
}

// utility.h:92-100: sum of b_i * (x_new - x_old) in index order.
__global__ __launch_bounds__(256) void k_strain(int n, const double *__restrict__ xo, const double *__restrict__ xn,
                                                const double *__restrict__ b, DevScalars *__restrict__ out)
{
    __shared__ double sd[kSumChunk], sb[kSumChunk];
    double r = 0;
    for (int base = 0; base < n; base += kSumChunk) {
        const int cnt = min(kSumChunk, n - base);
        for (int k = (int)threadIdx.x; k < cnt; k += 256) {
            sd[k] = xn[base + k] - xo[base + k];
            sb[k] = b[base + k];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll 8
            for (int k = 0; k < cnt; ++k) r += sb[k] * sd[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out->strain = r;
}

// The same sums as a fixed-order tree (one workgroup: lane t takes elements
// t, t + 1024, ..., then wave sums and wave partials in order) for the fast
// field, whose speeds
// differ from the reference's in the last bits anyway.
__device__ __forceinline__ double tree_finish(double acc, double *red, int t)
{
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    double s = 0.0;
    if (t == 0)
        for (int k = 0; k < 16; ++k) s += red[k];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(1024) void k_speed_sums_tree(int n, const double *__restrict__ v,
                                                          const double *__restrict__ b,
                                                          DevScalars *__restrict__ out)
{
    __shared__ double red[16];
    const int t = (int)threadIdx.x;
    double sa = 0.0, sq = 0.0, sbv = 0.0;
    // lane t takes t, t + 1024, ...: coalesced loads, a fixed order
    for (int k = t; k < n; k += 1024) {
        const double vi = v[k];
        sa += __builtin_fabs(vi);
        sq += vi * vi;
        sbv += b[k] * vi;
    }
    sa = tree_finish(sa, red, t);
    sq = tree_finish(sq, red, t);
    sbv = tree_finish(sbv, red, t);
    if (t == 0) {
        out->sum_abs = sa;
        out->sum_sq = sq;
        out->sum_bv = sbv;
    }
}

__global__ __launch_bounds__(1024) void k_strain_tree(int n, const double *__restrict__ xo,
                                                      const double *__restrict__ xn, const double *__restrict__ b,
                                                      DevScalars *__restrict__ out)
{
    __shared__ double red[16];
    const int t = (int)threadIdx.x;
    double r = 0.0;
    for (int k = t; k < n; k += 1024) r += b[k] * (xn[k] - xo[k]);
    r = tree_finish(r, red, t);
    if (t == 0) out->strain = r;
}

// One lane: copy the payload, then publish seq at system scope (the copies
// and every earlier kernel's results become visible to the host first).
__global__ void k_signal(const uint64_t *__restrict__ src, int words, HostSignal *__restrict__ sig, uint64_t seq)
{
    if (threadIdx.x != 0) return;
    for (int k = 0; k < words; ++k) sig->payload[k] = src[k];
    __hip_atomic_store(&sig->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The stage-III record of sdddst_stage_end as seven words: the squared-speed
// sums of V_INIT and V_INIT2, the controller's max ratio and next step, its
// accept flag and selected index (two ints in one word), the two strain
// increments.
__global__ void k_stage_pack(DevScalars *ds)
{
    if (threadIdx.x != 0) return;
    double *out = reinterpret_cast<double *>(ds + kStageOut);
    out[0] = ds[kStageSumA].sum_sq;
    out[1] = ds[kStageSumB].sum_sq;
    out[2] = ds[0].max_ratio;
    out[3] = ds[0].h_next;
    int *ints = reinterpret_cast<int *>(out + 4);
    ints[0] = ds[0].accept;
    ints[1] = ds[0].selected;
    out[5] = ds[kStageStrainA].strain;
    out[6] = ds[kStageStrainB].strain;
}

__global__ void k_fill(double *p, double v, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}

}  // namespace

static inline int nblocks(int n, int t) { return (n + t - 1) / t; }

int launch_residual(Context &c, double h, int new_slot, int old_slot, int csp, int isp)
{
    hipLaunchKernelGGL(k_residual, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, h, c.x[new_slot],
                       c.x[old_slot], c.v[csp], c.v[isp], c.dvec, c.g, c.gpart, c.gticket, c.gnorm2);
    SD_HIP(hipGetLastError());
    c.gnorm2_valid = true;
    return SDDDST_OK;
}

int launch_newton_update(Context &c, int slot)
{
    hipLaunchKernelGGL(k_newton_update, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, c.x[slot], c.delta);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

int launch_error_control(Context &c, double h, double hmax)
{
    const double minPrecSqr = c.p.precision * c.p.precision;
    hipLaunchKernelGGL(k_error_control, dim3(1), dim3(1024), 0, c.stream, c.n, c.x[SDDDST_SLOT_BIG_STEP],
                       c.x[SDDDST_SLOT_SECOND_SMALL], c.tol, minPrecSqr, h, hmax, c.ds);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

int launch_speed_sums(Context &c, int buf, DevScalars *dst)
{
    DevScalars *ds = dst ? dst : c.ds;
    if (c.p.field_mode == SDDDST_FIELD_FAST)
        hipLaunchKernelGGL(k_speed_sums_tree, dim3(1), dim3(1024), 0, c.stream, c.n, c.v[buf], c.b, ds);
    else
        hipLaunchKernelGGL(k_speed_sums, dim3(1), dim3(256), 0, c.stream, c.n, c.v[buf], c.b, ds);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

int launch_strain(Context &c, int old_slot, int new_slot, DevScalars *dst)
{
    DevScalars *ds = dst ? dst : c.ds;
    if (c.p.field_mode == SDDDST_FIELD_FAST)
        hipLaunchKernelGGL(k_strain_tree, dim3(1), dim3(1024), 0, c.stream, c.n, c.x[old_slot], c.x[new_slot], c.b,
                           ds);
    else
        hipLaunchKernelGGL(k_strain, dim3(1), dim3(256), 0, c.stream, c.n, c.x[old_slot], c.x[new_slot], c.b, ds);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

int launch_stage_pack(Context &c)
{
    hipLaunchKernelGGL(k_stage_pack, dim3(1), dim3(64), 0, c.stream, c.ds);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

int host_wait(Context &c, const void *src, int words, void *dst)
{
    SD_CHECK(words >= 0 && words <= 14 && (words == 0 || (src && dst)), SDDDST_ERR_ARG, "bad host_wait payload");
    const uint64_t seq = ++c.sig_seq;
    hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, c.stream, static_cast<const uint64_t *>(src), words, c.sig_dev,
                       seq);
    SD_HIP(hipGetLastError());
    // poll for up to 50 ms (a step's waits are microseconds to a few ms), then
    // let the runtime block: that also surfaces a failed kernel as an error
    const auto t0 = std::chrono::steady_clock::now();
    unsigned spins = 0;
    while (__atomic_load_n(&c.sig->seq, __ATOMIC_ACQUIRE) < seq) {
        __builtin_ia32_pause();
        if ((++spins & 4095u) == 0 &&
            std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
            SD_HIP(hipStreamSynchronize(c.stream));
            SD_CHECK(__atomic_load_n(&c.sig->seq, __ATOMIC_ACQUIRE) >= seq, SDDDST_ERR_HIP,
                     "device completion signal %llu not seen", (unsigned long long)seq);
            break;
        }
    }
    if (words) std::memcpy(dst, const_cast<const uint64_t *>(c.sig->payload), sizeof(uint64_t) * words);
    return SDDDST_OK;
}

int launch_reset_tolerance(Context &c)
{
    hipLaunchKernelGGL(k_fill, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.tol,
                       c.p.precision * c.p.precision, c.n);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

}  // namespace sdamd
