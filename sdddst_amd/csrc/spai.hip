// Learned sparse approximate inverse for the FGMRES Newton solves.  The
// reference solves A delta = g directly with UMFPACK (simulation.cpp:386-396)
// and has no preconditioner; this has no counterpart there.
//
// M = B^-1 + C.  B^-1 is the pair-block inverse of solver_fgmres.hip (bd =
// its diagonal, bo / bcol the partner entry).  C lives on the core graph of
// A: the off-diagonal entries whose pair lies within the cutoff radius (the
// untapered part of the Jacobian, told from the positions the pattern was
// built from), minus the pair partner, plus the diagonal:
//     C_ij = -gamma_ij * bd_i * A_ij * bd_j / (1 + |x_ij| + |w_ij|)
//     C_ii =  gamma_i  * bd_i * sum_j x_ij w_ij / (1 + |x_ij w_ij|)
// with x_ij = bd_i A_ij and w_ij = bd_j A_ji: the first-order (Neumann)
// terms of (B + E)^-1, damped where they are meaningless (|x|, |w| >> 1).
// gamma comes from a two-layer message-passing network (hidden width 32,
// FP32 weights) over the core graph:
//     h0 = relu(W0 u_i + b0)
//     a_ij = relu(P_l h_j + Q_l e_ij + c_l),  s_i = mean_j a_ij,
//     h_i = relu(U_l h_i + V_l s_i + d_l)                     (l = 1, 2)
//     gamma_ij = wa.h_i + wb.h_j + wab.(h_i*h_j) + we.e_ij + beta
//     gamma_i  = wd.h_i + beta_d
// gamma = 0 is the pair-block method itself.
//
// Training (sdddst_spai_train_step) runs on the context's current Newton
// matrix: K Rademacher probes W, R = A M W - W, q_i = mean_k R_ik^2,
// L = mean_i log(1 + q_i / eps); the backward pass runs on the device
// (A^T dR by the CSC view, dC on the core graph, the network in reverse)
// with deterministic fixed-order reductions, then Adam.
//
// Layouts.  Core graph in CSR by rows: Cp[n+1], Ci (column), Cr (row), Ck
// (position of the entry in A's CSR view), Cmir (edge index of (j, i)).
// Node arrays are [n][32] floats, edge features [ne][4].
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "context.h"
#include "field.h"
#include "solvers.h"

namespace sdamd {

namespace lay {
constexpr int H = 32, FN = 4, FE = 4;
constexpr int W0 = 0, B0 = W0 + H * FN;
constexpr int L1 = B0 + H;
constexpr int LP = 0, LQ = LP + H * H, LC = LQ + H * FE, LU = LC + H, LV = LU + H * H, LD = LV + H * H,
              LSZ = LD + H;
constexpr int layer(int l) { return L1 + (l - 1) * LSZ; }
constexpr int RO = L1 + 2 * LSZ;
constexpr int WA = RO, WB = WA + H, WAB = WB + H, WE = WAB + H, BETA = WE + FE, WD = BETA + 1, BETAD = WD + H;
constexpr int NPARAM = BETAD + 1;
constexpr int NRO = NPARAM - RO;   // readout parameters
}  // namespace lay

int spai_param_count() { return lay::NPARAM; }

struct Spai {
    int n = 0;
    int64_t ecap = 0, ne = 0;
    int kcap = 0;
    bool weights = false;
    bool graph_valid = false;
    float *theta = nullptr;
    double *grad = nullptr, *am = nullptr, *av = nullptr;
    int adam_t = 0;
    // core graph
    int *cnt = nullptr, *Cp = nullptr;
    int *Ci = nullptr, *Cr = nullptr, *Ck = nullptr, *Cmir = nullptr;
    float *U = nullptr, *E = nullptr;
    double *se = nullptr, *sd = nullptr;
    // activations: a0 h0 z1 S1 ph1 h1 z2 S2 ph2 h2, each [n][32]
    float *act = nullptr;
    float *ge = nullptr, *gd = nullptr;
    double *Ce = nullptr, *Cd = nullptr;
    // training
    double *W = nullptr, *Y = nullptr, *R = nullptr;
    double *lrow = nullptr, *loss = nullptr;
    float *dge = nullptr, *dgd = nullptr;
    float *dh = nullptr, *dhin = nullptr, *dph = nullptr, *dz = nullptr;
    float *dpe = nullptr;
    double *rowg = nullptr;      // [n][NRO] readout gradient rows
    double *part = nullptr;      // [kPartBlocks][<= 1056] partial sums
    long long *d_total = nullptr;
    float *act_(int k) const { return act + (size_t)k * n * lay::H; }
};

namespace {

constexpr int kPartBlocks = 64;
constexpr int kMaxOuter = lay::H * lay::H + lay::H;

template <class T>
int grow(T **p, int64_t count)
{
    if (*p) SD_HIP(hipFree(*p));
    *p = nullptr;
    if (count < 1) count = 1;
    SD_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)count));
    return SDDDST_OK;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// -- core graph -------------------------------------------------------------

// The entry (i, j) is a core edge when its pair's r^2 at the pattern's build,
// formed as the assembly forms it (higher index minus lower), is not above
// the squared cutoff.  For a pattern built with the within-cutoff bitmask the
// graph comes from its bits (k_core_count_mask, k_core_fill_mask); else this
// test runs on the positions of the build.
struct CoreTest {
    const double *x, *y;
    double rc2;
};

__device__ __forceinline__ bool core_entry(int i, int k, const int *__restrict__ Ai, const CoreTest &ct,
                                           int partner)
{
    const int j = Ai[k];
    if (j == i || j == partner) return false;
    const int lo = j < i ? j : i, hi = j < i ? i : j;
    const double dx = sdfield::normalize(ct.x[hi] - ct.x[lo]);
    const double dy = sdfield::normalize(ct.y[hi] - ct.y[lo]);
    return !(dx * dx + dy * dy > ct.rc2);
}

__global__ __launch_bounds__(256) void k_core_count(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                    CoreTest ct, const int *__restrict__ bcol,
                                                    int *__restrict__ cnt)
{
    const int lane = (int)(threadIdx.x & 63);
    const int i = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    const int end = Ap[i + 1], p = bcol[i];
    int c = 0;
    for (int base = Ap[i]; base < end; base += 64) {
        const int k = base + lane;
        const bool in = k < end && core_entry(i, k, Ai, ct, p);
        c += __popcll(__ballot(in));
    }
    if (lane == 0) cnt[i] = c;
}

__global__ __launch_bounds__(256) void k_core_fill(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                   CoreTest ct, const int *__restrict__ bcol,
                                                   const int *__restrict__ Cp,
                                                   int *__restrict__ Ci, int *__restrict__ Cr, int *__restrict__ Ck)
{
    const int lane = (int)(threadIdx.x & 63);
    const int i = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    const int end = Ap[i + 1], p = bcol[i];
    int pos = Cp[i];
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int base = Ap[i]; base < end; base += 64) {
        const int k = base + lane;
        const bool in = k < end && core_entry(i, k, Ai, ct, p);
        const unsigned long long m = __ballot(in);
        if (in) {
            const int e = pos + __popcll(m & below);
            Ci[e] = Ai[k];
            Cr[e] = i;
            Ck[e] = k;
        }
        pos += __popcll(m);
    }
}

// Cmir[e] for e = (i, j): the edge (j, i), by binary search of row j.
// The same graph from the count pass's within-cutoff bitmask (row i's bits
// are its core rows; the diagonal is never set): the count is the row's
// total less its partner, and the rows come out of the mask words in
// ascending order, each entry's position in A from the pattern's mask word
// and word start.
__global__ void k_core_count_mask(int n, int words, const unsigned long long *__restrict__ cmask,
                                  const int *__restrict__ ccnt, const int *__restrict__ bcol, int *__restrict__ cnt)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int p = bcol[i];
    int c = ccnt[i];
    if (p >= 0 && ((cmask[(size_t)i * words + (p >> 6)] >> (p & 63)) & 1ull)) --c;
    cnt[i] = c;
}

__global__ __launch_bounds__(256) void k_core_fill_mask(int n, int words, const unsigned long long *__restrict__ cmask,
                                                        const unsigned long long *__restrict__ mask,
                                                        const int *__restrict__ wstart, const int *__restrict__ bcol,
                                                        const int *__restrict__ Cp, int *__restrict__ Ci,
                                                        int *__restrict__ Cr, int *__restrict__ Ck)
{
    const int lane = (int)(threadIdx.x & 63);
    const int i = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    const int p = bcol[i];
    const size_t row = (size_t)i * words;
    int pos = Cp[i];
    for (int w0 = 0; w0 < words; w0 += 64) {
        const int w = w0 + lane;
        unsigned long long bits = w < words ? cmask[row + w] : 0ull;
        if (p >= 0 && (p >> 6) == w) bits &= ~(1ull << (p & 63));
        const int pc = __popcll(bits);
        int incl = pc;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        int e = pos + incl - pc;
        if (bits) {
            const unsigned long long pm = mask[row + w];
            const int ws = wstart[row + w];
            while (bits) {
                const int b = __builtin_ctzll(bits);
                Ci[e] = w * 64 + b;
                Cr[e] = i;
                Ck[e] = ws + __popcll(pm & ((1ull << b) - 1ull));
                ++e;
                bits &= bits - 1ull;
            }
        }
        pos += __shfl(incl, 63, 64);
    }
}

__global__ void k_core_mirror(int64_t ne, const int *__restrict__ Cp, const int *__restrict__ Ci,
                              const int *__restrict__ Cr, int *__restrict__ Cmir)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int i = Cr[e], j = Ci[e];
    int lo = Cp[j], hi = Cp[j + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Ci[mid] < i) lo = mid + 1;
        else hi = mid;
    }
    Cmir[e] = lo;
}

// Sum over the G lanes of a lane group (G a power of two dividing 64).
template <int G>
__device__ __forceinline__ double group_sum(double s)
{
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    return s;
}

// Node and edge features and the FP64 scales of C.  A group of G lanes
// shares a row: lane g takes its edges g, g + G, ... (coalesced loads and
// 16-byte feature stores), the row sums close with a butterfly in the group.
template <int G>
__global__ __launch_bounds__(256) void k_features(int n, const int *__restrict__ Cp, const int *__restrict__ Ci,
                                                  const int *__restrict__ Ck, const int *__restrict__ Cmir,
                                                  const int *__restrict__ diag, const double *__restrict__ AxRow,
                                                  const double *__restrict__ binv, const int *__restrict__ bcol,
                                                  float *__restrict__ U, float *__restrict__ E,
                                                  double *__restrict__ se, double *__restrict__ sd)
{
    const int g = (int)(threadIdx.x & (G - 1));
    const int i = (int)((blockIdx.x * 256 + threadIdx.x) / G);
    const bool ok = i < n;
    double sr = 0, sc = 0, t = 0, bdi = 0;
    if (ok) {
        bdi = binv[2 * i];
        const int end = Cp[i + 1];
        for (int e = Cp[i] + g; e < end; e += G) {
            // A_ji: row j's entry at column i, the mirror edge's
            const int k = Ck[e], j = Ci[e];
            const double aij = AxRow[k], aji = AxRow[Ck[Cmir[e]]], bdj = binv[2 * j];
            const double xe = bdi * aij, ye = aij * bdj, we = bdj * aji;
            se[e] = -bdi * aij * bdj / (1.0 + fabs(xe) + fabs(we));
            t += xe * we / (1.0 + fabs(xe * we));
            sr += fabs(xe);
            sc += fabs(aji * bdi);
            const double xw = fabs(xe * we);
            *reinterpret_cast<float4 *>(E + 4 * (size_t)e) =
                make_float4((float)(xe / (1.0 + fabs(xe))), (float)(ye / (1.0 + fabs(ye))),
                            (float)(we / (1.0 + fabs(we))), (float)(xw / (1.0 + xw)));
        }
    }
    sr = group_sum<G>(sr);
    sc = group_sum<G>(sc);
    t = group_sum<G>(t);
    if (!ok || g != 0) return;
    sd[i] = bdi * t;
    const double rho = fabs(fabs(AxRow[diag[i]]) * bdi);
    *reinterpret_cast<float4 *>(U + 4 * (size_t)i) =
        make_float4((float)(sr / (1.0 + sr)), (float)(sc / (1.0 + sc)), bcol[i] >= 0 ? 1.0f : 0.0f,
                    (float)(rho / (1.0 + rho)));
}

// -- forward ----------------------------------------------------------------
// Node kernels: a workgroup of 256 lanes holds 8 nodes x 32 channels; the
// 32x32 matrices are staged transposed in LDS (lane = output channel reads
// consecutive words).

constexpr int kLdT = 33;   // padded row of a transposed 32x32 matrix in LDS (no bank conflicts)
__device__ __forceinline__ void stage_t(float *dst, const float *__restrict__ src)
{
    // dst[c*kLdT + r] = src[r*32 + c]
    for (int t = (int)threadIdx.x; t < lay::H * lay::H; t += 256) {
        const int r = t >> 5, c = t & 31;
        dst[c * kLdT + r] = src[t];
    }
}

__global__ __launch_bounds__(256) void k_gnn_in(int n, const float *__restrict__ th, const float *__restrict__ U,
                                                float *__restrict__ a0, float *__restrict__ h0,
                                                float *__restrict__ z1)
{
    __shared__ float Pt[32 * kLdT];
    __shared__ float hs[8][32];
    stage_t(Pt, th + lay::layer(1) + lay::LP);
    const int nl = (int)(threadIdx.x >> 5), ch = (int)(threadIdx.x & 31);
    const int i = (int)blockIdx.x * 8 + nl;
    const bool ok = i < n;
    float h = 0.f;
    if (ok) {
        float a = th[lay::B0 + ch];
        for (int f = 0; f < 4; ++f) a += th[lay::W0 + ch * 4 + f] * U[4 * i + f];
        h = fmaxf(a, 0.f);
        a0[(size_t)i * 32 + ch] = a;
        h0[(size_t)i * 32 + ch] = h;
    }
    hs[nl][ch] = h;
    __syncthreads();
    if (ok) {
        float z = 0.f;
        for (int c = 0; c < 32; ++c) z += Pt[c * kLdT + ch] * hs[nl][c];
        z1[(size_t)i * 32 + ch] = z;
    }
}

// One message-passing layer; znext = P_{l+1} h_out when l == 1.
__global__ __launch_bounds__(256) void k_gnn_mp(int n, int l, const float *__restrict__ th,
                                                const int *__restrict__ Cp, const int *__restrict__ Ci,
                                                const float *__restrict__ E, const float *__restrict__ z,
                                                const float *__restrict__ hin, float *__restrict__ S,
                                                float *__restrict__ ph, float *__restrict__ hout,
                                                float *__restrict__ znext)
{
    __shared__ float Ut[32 * kLdT], Vt[32 * kLdT], Pn[32 * kLdT];
    __shared__ float ss[8][32], hs[8][32];
    const float *L = th + lay::layer(l);
    stage_t(Ut, L + lay::LU);
    stage_t(Vt, L + lay::LV);
    if (znext) stage_t(Pn, th + lay::layer(l + 1) + lay::LP);
    const int nl = (int)(threadIdx.x >> 5), ch = (int)(threadIdx.x & 31);
    const int i = (int)blockIdx.x * 8 + nl;
    const bool ok = i < n;
    float s = 0.f, hv = 0.f;
    if (ok) {
        const float q0 = L[lay::LQ + ch * 4 + 0], q1 = L[lay::LQ + ch * 4 + 1];
        const float q2 = L[lay::LQ + ch * 4 + 2], q3 = L[lay::LQ + ch * 4 + 3];
        const float cc = L[lay::LC + ch];
        const int e0 = Cp[i], e1 = Cp[i + 1];
        float acc = 0.f;
        // four edges' loads issued before their arithmetic (same order of
        // additions)
        int e = e0;
        for (; e + 4 <= e1; e += 4) {
            int jj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) jj[u] = Ci[e + u];
            float zv[4];
            float4 ef[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                zv[u] = z[(size_t)jj[u] * 32 + ch];
                ef[u] = *reinterpret_cast<const float4 *>(E + 4 * (size_t)(e + u));
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float pe = zv[u] + cc + q0 * ef[u].x + q1 * ef[u].y + q2 * ef[u].z + q3 * ef[u].w;
                acc += fmaxf(pe, 0.f);
            }
        }
        for (; e < e1; ++e) {
            const float *ef = E + 4 * (size_t)e;
            const float pe = z[(size_t)Ci[e] * 32 + ch] + cc + q0 * ef[0] + q1 * ef[1] + q2 * ef[2] + q3 * ef[3];
            acc += fmaxf(pe, 0.f);
        }
        const int deg = e1 - e0;
        s = acc / (float)(deg > 0 ? deg : 1);
        S[(size_t)i * 32 + ch] = s;
        hv = hin[(size_t)i * 32 + ch];
    }
    ss[nl][ch] = s;
    hs[nl][ch] = hv;
    __syncthreads();
    float h = 0.f;
    if (ok) {
        float p = L[lay::LD + ch];
        for (int c = 0; c < 32; ++c) p += Ut[c * kLdT + ch] * hs[nl][c] + Vt[c * kLdT + ch] * ss[nl][c];
        ph[(size_t)i * 32 + ch] = p;
        h = fmaxf(p, 0.f);
        hout[(size_t)i * 32 + ch] = h;
    }
    if (!znext) return;
    __syncthreads();
    hs[nl][ch] = h;
    __syncthreads();
    if (ok) {
        float zz = 0.f;
        for (int c = 0; c < 32; ++c) zz += Pn[c * kLdT + ch] * hs[nl][c];
        znext[(size_t)i * 32 + ch] = zz;
    }
}

__global__ void k_gnn_edges(int64_t ne, const float *__restrict__ th, const int *__restrict__ Cr,
                            const int *__restrict__ Ci, const float *__restrict__ E, const float *__restrict__ h,
                            const double *__restrict__ se, float *__restrict__ ge, double *__restrict__ Ce)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const float *hi = h + (size_t)Cr[e] * 32;
    const float *hj = h + (size_t)Ci[e] * 32;
    float g = th[lay::BETA];
    for (int c = 0; c < 32; ++c) {
        const float a = hi[c], b = hj[c];
        g += th[lay::WA + c] * a + th[lay::WB + c] * b + th[lay::WAB + c] * (a * b);
    }
    for (int f = 0; f < 4; ++f) g += th[lay::WE + f] * E[4 * (size_t)e + f];
    ge[e] = g;
    Ce[e] = (double)g * se[e];
}

__global__ void k_gnn_nodes(int n, const float *__restrict__ th, const float *__restrict__ h,
                            const double *__restrict__ sd, float *__restrict__ gd, double *__restrict__ Cd)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    float g = th[lay::BETAD];
    for (int c = 0; c < 32; ++c) g += th[lay::WD + c] * h[(size_t)i * 32 + c];
    gd[i] = g;
    Cd[i] = (double)g * sd[i];
}

// z_p = M v_p for probe p = blockIdx.y; G lanes per row share its core
// entries, as in k_features.
template <int G>
__global__ __launch_bounds__(256) void k_spai_apply(int n, const double *__restrict__ binv,
                                                    const int *__restrict__ bcol, const int *__restrict__ Cp,
                                                    const int *__restrict__ Ci, const double *__restrict__ Ce,
                                                    const double *__restrict__ Cd, const double *__restrict__ v,
                                                    double *__restrict__ z)
{
    const int g = (int)(threadIdx.x & (G - 1));
    const int i = (int)((blockIdx.x * 256 + threadIdx.x) / G);
    const bool ok = i < n;
    const size_t off = (size_t)blockIdx.y * n;
    const double *vp = v + off;
    double s = 0.0;
    if (ok) {
        const int end = Cp[i + 1];
        for (int e = Cp[i] + g; e < end; e += G) s += Ce[e] * vp[Ci[e]];
    }
    s = group_sum<G>(s);
    if (!ok || g != 0) return;
    const int j = bcol[i];
    double r = binv[2 * i] * vp[i];
    if (j >= 0) r += binv[2 * i + 1] * vp[j];
    r += Cd[i] * vp[i];
    z[off + i] = r + s;
}

// -- training ---------------------------------------------------------------

__global__ void k_probes(int k, int n, unsigned long long seed, double *__restrict__ W)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)k * n) return;
    const unsigned long long z = mix64(seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(t + 1));
    W[t] = (z >> 63) ? -1.0 : 1.0;
}

// y_p = A x_p (or A^T x_p with the CSC values), one wavefront per row.
__global__ __launch_bounds__(256) void k_spmv_batch(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                    const double *__restrict__ Av, const double *__restrict__ x,
                                                    double *__restrict__ y)
{
    const int lane = (int)(threadIdx.x & 63);
    const int r = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (r >= n) return;
    const size_t off = (size_t)blockIdx.y * n;
    double s = 0.0;
    for (int k = Ap[r] + lane; k < Ap[r + 1]; k += 64) s += Av[k] * x[off + Ai[k]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) y[off + r] = s;
}

// R (holding A M W on entry) -> dL/dR; lrow_i = log(1 + q_i / eps).
__global__ void k_loss_rows(int k, int n, double eps, const double *__restrict__ W, double *__restrict__ R,
                            double *__restrict__ lrow)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double q = 0;
    for (int p = 0; p < k; ++p) {
        const double r = R[(size_t)p * n + i] - W[(size_t)p * n + i];
        R[(size_t)p * n + i] = r;
        q += r * r;
    }
    q /= k;
    lrow[i] = log1p(q / eps);
    const double f = 2.0 / ((double)k * n) / (eps + q);
    for (int p = 0; p < k; ++p) R[(size_t)p * n + i] *= f;
}

// Deterministic sum of v[0..n) by one workgroup: out = (sum) * scale.
__global__ __launch_bounds__(1024) void k_sum(int n, const double *__restrict__ v, double scale,
                                              double *__restrict__ out)
{
    __shared__ double red[1024];
    double acc = 0;
    for (int i = (int)threadIdx.x; i < n; i += 1024) acc += v[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

// dL/dgamma from dY = A^T dR on the core graph.
__global__ void k_dgamma(int n, int k, const int *__restrict__ Cp, const int *__restrict__ Ci,
                         const double *__restrict__ dY, const double *__restrict__ W, const double *__restrict__ se,
                         const double *__restrict__ sd, float *__restrict__ dge, float *__restrict__ dgd)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    for (int e = Cp[i]; e < Cp[i + 1]; ++e) {
        const int j = Ci[e];
        double acc = 0;
        for (int p = 0; p < k; ++p) acc += dY[(size_t)p * n + i] * W[(size_t)p * n + j];
        dge[e] = (float)(acc * se[e]);
    }
    double acc = 0;
    for (int p = 0; p < k; ++p) acc += dY[(size_t)p * n + i] * W[(size_t)p * n + i];
    dgd[i] = (float)(acc * sd[i]);
}

// Readout backward: dL/dh2 per node and the per-node rows of the readout
// gradient (wa wb wab we beta wd betad), summed over nodes afterwards.
__global__ __launch_bounds__(256) void k_bwd_readout(int n, const float *__restrict__ th,
                                                     const int *__restrict__ Cp, const int *__restrict__ Ci,
                                                     const int *__restrict__ Cmir, const float *__restrict__ E,
                                                     const float *__restrict__ h, const float *__restrict__ dge,
                                                     const float *__restrict__ dgd, float *__restrict__ dh,
                                                     double *__restrict__ rowg)
{
    const int nl = (int)(threadIdx.x >> 5), ch = (int)(threadIdx.x & 31);
    const int i = (int)blockIdx.x * 8 + nl;
    if (i >= n) return;
    (void)nl;
    const float wa = th[lay::WA + ch], wb = th[lay::WB + ch], wab = th[lay::WAB + ch], wd = th[lay::WD + ch];
    const float hi = h[(size_t)i * 32 + ch];
    const float gdi = dgd[i];
    float d = gdi * wd;
    double ra = 0, rb = 0, vab = 0, we = 0;
    for (int e = Cp[i]; e < Cp[i + 1]; ++e) {
        const float hj = h[(size_t)Ci[e] * 32 + ch];
        const float g1 = dge[e], g2 = dge[Cmir[e]];
        d += g1 * (wa + wab * hj) + g2 * (wb + wab * hj);
        ra += g1;
        rb += g2;
        vab += (double)g1 * hj;
        if (ch < 4) we += (double)g1 * E[4 * (size_t)e + ch];
    }
    dh[(size_t)i * 32 + ch] = d;
    double *rg = rowg + (size_t)i * lay::NRO;
    rg[lay::WA - lay::RO + ch] = ra * hi;
    rg[lay::WB - lay::RO + ch] = rb * hi;
    rg[lay::WAB - lay::RO + ch] = vab * hi;
    rg[lay::WD - lay::RO + ch] = (double)gdi * hi;
    if (ch < 4) rg[lay::WE - lay::RO + ch] = we;
    if (ch == 0) {
        rg[lay::BETA - lay::RO] = ra;
        rg[lay::BETAD - lay::RO] = gdi;
    }
}

// Layer backward, node part: dph = dh * [ph > 0]; dhin = U^T dph;
// dS = V^T dph / deg; per edge of the row dpe = dS * [pe > 0].
__global__ __launch_bounds__(256) void k_bwd_mp(int n, int l, const float *__restrict__ th,
                                                const int *__restrict__ Cp, const int *__restrict__ Ci,
                                                const float *__restrict__ E, const float *__restrict__ z,
                                                const float *__restrict__ ph, const float *__restrict__ dh,
                                                float *__restrict__ dph, float *__restrict__ dhin,
                                                float *__restrict__ dpe)
{
    __shared__ float Um[1024], Vm[1024];
    __shared__ float ds[8][32];
    const float *L = th + lay::layer(l);
    // untransposed: lane ch reads column ch, i.e. Um[r*32 + ch]
    for (int t = (int)threadIdx.x; t < 1024; t += 256) {
        Um[t] = L[lay::LU + t];
        Vm[t] = L[lay::LV + t];
    }
    const int nl = (int)(threadIdx.x >> 5), ch = (int)(threadIdx.x & 31);
    const int i = (int)blockIdx.x * 8 + nl;
    const bool ok = i < n;
    float g = 0.f;
    if (ok) {
        g = ph[(size_t)i * 32 + ch] > 0.f ? dh[(size_t)i * 32 + ch] : 0.f;
        dph[(size_t)i * 32 + ch] = g;
    }
    ds[nl][ch] = g;
    __syncthreads();
    if (!ok) return;
    float a = 0.f, b = 0.f;
    for (int r = 0; r < 32; ++r) {
        a += Um[r * 32 + ch] * ds[nl][r];
        b += Vm[r * 32 + ch] * ds[nl][r];
    }
    dhin[(size_t)i * 32 + ch] = a;
    const int e0 = Cp[i], e1 = Cp[i + 1];
    const float dS = b / (float)(e1 - e0 > 0 ? e1 - e0 : 1);
    const float q0 = L[lay::LQ + ch * 4 + 0], q1 = L[lay::LQ + ch * 4 + 1];
    const float q2 = L[lay::LQ + ch * 4 + 2], q3 = L[lay::LQ + ch * 4 + 3];
    const float cc = L[lay::LC + ch];
    for (int e = e0; e < e1; ++e) {
        const float *ef = E + 4 * (size_t)e;
        const float pe = z[(size_t)Ci[e] * 32 + ch] + cc + q0 * ef[0] + q1 * ef[1] + q2 * ef[2] + q3 * ef[3];
        dpe[(size_t)e * 32 + ch] = pe > 0.f ? dS : 0.f;
    }
}

// Layer backward, gather part: dz_j = sum of dpe over the edges into j
// (the mirrors of row j's edges); dhin += P^T dz.
__global__ __launch_bounds__(256) void k_bwd_gather(int n, int l, const float *__restrict__ th,
                                                    const int *__restrict__ Cp, const int *__restrict__ Cmir,
                                                    const float *__restrict__ dpe, float *__restrict__ dz,
                                                    float *__restrict__ dhin)
{
    __shared__ float Pm[1024];
    __shared__ float zs[8][32];
    const float *L = th + lay::layer(l);
    for (int t = (int)threadIdx.x; t < 1024; t += 256) Pm[t] = L[lay::LP + t];
    const int nl = (int)(threadIdx.x >> 5), ch = (int)(threadIdx.x & 31);
    const int j = (int)blockIdx.x * 8 + nl;
    const bool ok = j < n;
    float acc = 0.f;
    if (ok) {
        for (int e = Cp[j]; e < Cp[j + 1]; ++e) acc += dpe[(size_t)Cmir[e] * 32 + ch];
        dz[(size_t)j * 32 + ch] = acc;
    }
    zs[nl][ch] = acc;
    __syncthreads();
    if (!ok) return;
    float a = 0.f;
    for (int r = 0; r < 32; ++r) a += Pm[r * 32 + ch] * zs[nl][r];
    dhin[(size_t)j * 32 + ch] += a;
}

__global__ void k_relu_back(int64_t m, const float *__restrict__ pre, float *__restrict__ g)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m && !(pre[t] > 0.f)) g[t] = 0.f;
}

// partial[b][r*nb + c] = sum over this block's rows of A[row][r] * B[row][c]
// (c < nb), and with bias partial[b][32*nb + r] = sum of A[row][r].
__global__ __launch_bounds__(256) void k_outer(int64_t rows, int64_t per, const float *__restrict__ A,
                                               const float *__restrict__ B, int nb, int bias,
                                               double *__restrict__ partial)
{
    __shared__ float As[32][33], Bs[32][33];
    const int nout = 32 * nb + (bias ? 32 : 0);
    const int64_t r0 = (int64_t)blockIdx.x * per;
    const int64_t r1 = r0 + per < rows ? r0 + per : rows;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t t0 = r0; t0 < r1; t0 += 32) {
        __syncthreads();
        for (int t = (int)threadIdx.x; t < 32 * 32; t += 256) {
            const int rr = t >> 5, cc = t & 31;
            const int64_t row = t0 + rr;
            const bool in = row < r1;
            As[rr][cc] = in ? A[row * 32 + cc] : 0.f;
            Bs[rr][cc] = in && cc < nb ? B[row * nb + cc] : 0.f;
        }
        __syncthreads();
        for (int s = 0; s < 5; ++s) {
            const int o = (int)threadIdx.x + 256 * s;
            if (o >= nout) break;
            double a = 0;
            if (o < 32 * nb) {
                const int r = o / nb, c = o % nb;
                for (int rr = 0; rr < 32; ++rr) a += (double)As[rr][r] * Bs[rr][c];
            } else {
                const int r = o - 32 * nb;
                for (int rr = 0; rr < 32; ++rr) a += (double)As[rr][r];
            }
            acc[s] += a;
        }
    }
    for (int s = 0; s < 5; ++s) {
        const int o = (int)threadIdx.x + 256 * s;
        if (o < nout) partial[(size_t)blockIdx.x * kMaxOuter + o] = acc[s];
    }
}

// out[o] = sum over blocks (in order) of partial[b][o].
__global__ void k_reduce_part(int nblocks, int nout, int stride, const double *__restrict__ partial,
                              double *__restrict__ out)
{
    const int o = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (o >= nout) return;
    double s = 0;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * stride + o];
    out[o] = s;
}

// partial[b][c] = sum of X[row][c] over the block's rows (X is [rows][cols]).
__global__ void k_colsum(int rows, int per, int cols, const double *__restrict__ X, double *__restrict__ partial)
{
    const int c = (int)threadIdx.x;
    if (c >= cols) return;
    const int r0 = (int)blockIdx.x * per;
    const int r1 = r0 + per < rows ? r0 + per : rows;
    double s = 0;
    for (int r = r0; r < r1; ++r) s += X[(size_t)r * cols + c];
    partial[(size_t)blockIdx.x * kMaxOuter + c] = s;
}

__global__ void k_adam(int np, const double *__restrict__ g, float *__restrict__ th, double *__restrict__ m,
                       double *__restrict__ v, double lr, double c1, double c2)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= np) return;
    const double gt = g[t];
    const double mt = 0.9 * m[t] + 0.1 * gt;
    const double vt = 0.999 * v[t] + 0.001 * gt * gt;
    m[t] = mt;
    v[t] = vt;
    th[t] = (float)((double)th[t] - lr * (mt / c1) / (sqrt(vt / c2) + 1e-8));
}

}  // namespace

static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Lanes per row of the row-group kernels: about the mean core degree, so a
// row takes one or two passes.
static int row_group(const Spai &s)
{
    const double avg = s.n > 0 ? (double)s.ne / s.n : 0.0;
    return avg <= 3 ? 4 : avg <= 7 ? 8 : avg <= 15 ? 16 : 32;
}

template <int G>
static void launch_features(Context &c, const Spai &s, const Fgmres &f)
{
    hipLaunchKernelGGL(k_features<G>, dim3(nblk((int64_t)c.n * G)), dim3(256), 0, c.stream, c.n, s.Cp, s.Ci, s.Ck,
                       s.Cmir, c.diag, c.AxRow, f.binv, f.bcol, s.U, s.E, s.se, s.sd);
}

template <int G>
static void launch_apply_g(Context &c, const Spai &s, const Fgmres &f, int k, const double *v, double *z)
{
    hipLaunchKernelGGL(k_spai_apply<G>, dim3(nblk((int64_t)c.n * G), k), dim3(256), 0, c.stream, c.n, f.binv,
                       f.bcol, s.Cp, s.Ci, s.Ce, s.Cd, v, z);
}

// z_p = M v_p for k column vectors of length n.
static void launch_apply(Context &c, const Spai &s, const Fgmres &f, int k, const double *v, double *z)
{
    switch (row_group(s)) {
    case 4: launch_apply_g<4>(c, s, f, k, v, z); break;
    case 8: launch_apply_g<8>(c, s, f, k, v, z); break;
    case 16: launch_apply_g<16>(c, s, f, k, v, z); break;
    default: launch_apply_g<32>(c, s, f, k, v, z); break;
    }
}

static int spai_alloc_nodes(Context &c, Spai &s)
{
    if (s.n == c.n && s.act) return SDDDST_OK;
    const int n = c.n;
    s.n = n;
    int rc;
    if ((rc = grow(&s.cnt, n)) || (rc = grow(&s.Cp, n + 1)) || (rc = grow(&s.U, 4 * (int64_t)n)) ||
        (rc = grow(&s.sd, n)) || (rc = grow(&s.act, 10 * (int64_t)n * 32)) || (rc = grow(&s.gd, n)) ||
        (rc = grow(&s.Cd, n)) || (rc = grow(&s.d_total, 1)))
        return rc;
    return SDDDST_OK;
}

static int spai_alloc_edges(Spai &s, int64_t ne)
{
    if (ne <= s.ecap) return SDDDST_OK;
    const int64_t cap = ne + ne / 4 + 1024;
    int rc;
    if ((rc = grow(&s.Ci, cap)) || (rc = grow(&s.Cr, cap)) || (rc = grow(&s.Ck, cap)) || (rc = grow(&s.Cmir, cap)) ||
        (rc = grow(&s.E, 4 * cap)) || (rc = grow(&s.se, cap)) || (rc = grow(&s.ge, cap)) || (rc = grow(&s.Ce, cap)))
        return rc;
    if (s.dge) {
        if ((rc = grow(&s.dge, cap)) || (rc = grow(&s.dpe, 32 * cap))) return rc;
    }
    s.ecap = cap;
    return SDDDST_OK;
}

static int spai_alloc_training(Context &c, Spai &s, int k)
{
    const int n = c.n;
    int rc;
    if (!s.dh) {
        if ((rc = grow(&s.lrow, n)) || (rc = grow(&s.loss, 1)) || (rc = grow(&s.dgd, n)) ||
            (rc = grow(&s.dh, 32 * (int64_t)n)) || (rc = grow(&s.dhin, 32 * (int64_t)n)) ||
            (rc = grow(&s.dph, 32 * (int64_t)n)) || (rc = grow(&s.dz, 32 * (int64_t)n)) ||
            (rc = grow(&s.rowg, (int64_t)lay::NRO * n)) || (rc = grow(&s.part, (int64_t)kPartBlocks * kMaxOuter)) ||
            (rc = grow(&s.dge, s.ecap)) || (rc = grow(&s.dpe, 32 * s.ecap)))
            return rc;
    }
    if (k > s.kcap) {
        if ((rc = grow(&s.W, (int64_t)k * n)) || (rc = grow(&s.Y, (int64_t)k * n)) || (rc = grow(&s.R, (int64_t)k * n)))
            return rc;
        s.kcap = k;
    }
    if (!s.grad) {
        if ((rc = grow(&s.grad, lay::NPARAM)) || (rc = grow(&s.am, lay::NPARAM)) || (rc = grow(&s.av, lay::NPARAM)))
            return rc;
        SD_HIP(hipMemsetAsync(s.am, 0, sizeof(double) * lay::NPARAM, c.stream));
        SD_HIP(hipMemsetAsync(s.av, 0, sizeof(double) * lay::NPARAM, c.stream));
    }
    return SDDDST_OK;
}

static Spai &spai_state(Context &c)
{
    if (!c.spai) c.spai = new Spai();
    return *c.spai;
}

// Core graph, features and the network's forward pass on the current
// Newton matrix; binv / bcol is the pair-block inverse of that matrix.
static int spai_forward(Context &c, const Fgmres &f)
{
    Spai &s = *c.spai;
    SD_CHECK(c.jac_symmetric, SDDDST_ERR_STATE, "learned preconditioner needs a symmetric Jacobian pattern");
    const int n = c.n;
    int rc = spai_alloc_nodes(c, s);
    if (rc) return rc;
    const int words = (n + 63) / 64;
    const bool from_mask = c.jcmask_valid;
    const CoreTest ct{c.jx, c.y, c.jrc2};
    if (from_mask)
        hipLaunchKernelGGL(k_core_count_mask, dim3(nblk(n)), dim3(256), 0, c.stream, n, words, c.jcmask, c.jccnt,
                           f.bcol, s.cnt);
    else
        hipLaunchKernelGGL(k_core_count, dim3(nblk(n, 4)), dim3(256), 0, c.stream, n, c.Ap, c.Ai, ct, f.bcol, s.cnt);
    SD_HIP(hipGetLastError());
    if ((rc = launch_scan(c, n, s.cnt, s.Cp, s.d_total))) return rc;
    long long ne = 0;
    if ((rc = host_wait(c, s.d_total, 1, &ne))) return rc;
    if ((rc = spai_alloc_edges(s, ne))) return rc;
    s.ne = ne;
    if (from_mask)
        hipLaunchKernelGGL(k_core_fill_mask, dim3(nblk(n, 4)), dim3(256), 0, c.stream, n, words, c.jcmask, c.jmask,
                           c.jwstart, f.bcol, s.Cp, s.Ci, s.Cr, s.Ck);
    else
        hipLaunchKernelGGL(k_core_fill, dim3(nblk(n, 4)), dim3(256), 0, c.stream, n, c.Ap, c.Ai, ct, f.bcol, s.Cp,
                           s.Ci, s.Cr, s.Ck);
    if (ne > 0)
        hipLaunchKernelGGL(k_core_mirror, dim3(nblk(ne)), dim3(256), 0, c.stream, (int64_t)ne, s.Cp, s.Ci, s.Cr,
                           s.Cmir);
    switch (row_group(s)) {
    case 4: launch_features<4>(c, s, f); break;
    case 8: launch_features<8>(c, s, f); break;
    case 16: launch_features<16>(c, s, f); break;
    default: launch_features<32>(c, s, f); break;
    }
    float *a0 = s.act_(0), *h0 = s.act_(1), *z1 = s.act_(2), *S1 = s.act_(3), *ph1 = s.act_(4), *h1 = s.act_(5);
    float *z2 = s.act_(6), *S2 = s.act_(7), *ph2 = s.act_(8), *h2 = s.act_(9);
    hipLaunchKernelGGL(k_gnn_in, dim3(nblk(n, 8)), dim3(256), 0, c.stream, n, s.theta, s.U, a0, h0, z1);
    hipLaunchKernelGGL(k_gnn_mp, dim3(nblk(n, 8)), dim3(256), 0, c.stream, n, 1, s.theta, s.Cp, s.Ci, s.E, z1, h0, S1,
                       ph1, h1, z2);
    hipLaunchKernelGGL(k_gnn_mp, dim3(nblk(n, 8)), dim3(256), 0, c.stream, n, 2, s.theta, s.Cp, s.Ci, s.E, z2, h1, S2,
                       ph2, h2, (float *)nullptr);
    if (ne > 0)
        hipLaunchKernelGGL(k_gnn_edges, dim3(nblk(ne)), dim3(256), 0, c.stream, (int64_t)ne, s.theta, s.Cr, s.Ci, s.E,
                           h2, s.se, s.ge, s.Ce);
    hipLaunchKernelGGL(k_gnn_nodes, dim3(nblk(n)), dim3(256), 0, c.stream, n, s.theta, h2, s.sd, s.gd, s.Cd);
    SD_HIP(hipGetLastError());
    s.graph_valid = true;
    return SDDDST_OK;
}

int spai_factor(Context &c, const Fgmres &f)
{
    SD_CHECK(c.spai && c.spai->weights, SDDDST_ERR_STATE,
             "learned preconditioner selected but no weights are installed");
    return spai_forward(c, f);
}

int spai_apply(Context &c, const Fgmres &f, const double *v, double *z)
{
    Spai &s = *c.spai;
    launch_apply(c, s, f, 1, v, z);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

// Weight gradient of one outer-product block into grad[off ...].
static int outer(Context &c, Spai &s, int64_t rows, const float *A, const float *B, int nb, int bias, int off)
{
    const int nout = 32 * nb + (bias ? 32 : 0);
    int64_t per = (rows + kPartBlocks - 1) / kPartBlocks;
    per = per < 32 ? 32 : per;
    const int nblocks = rows > 0 ? (int)((rows + per - 1) / per) : 0;
    if (nblocks == 0) {
        SD_HIP(hipMemsetAsync(s.grad + off, 0, sizeof(double) * nout, c.stream));
        return SDDDST_OK;
    }
    hipLaunchKernelGGL(k_outer, dim3(nblocks), dim3(256), 0, c.stream, rows, per, A, B, nb, bias, s.part);
    hipLaunchKernelGGL(k_reduce_part, dim3(nblk(nout)), dim3(256), 0, c.stream, nblocks, nout, kMaxOuter, s.part,
                       s.grad + off);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

int spai_train_step(Context &c, Fgmres &f, int k, unsigned long long seed, double eps, double lr, double *loss_out,
                    double *grad_out)
{
    Spai &s = *c.spai;
    const int n = c.n;
    int rc = spai_forward(c, f);
    if (rc) return rc;
    if ((rc = spai_alloc_training(c, s, k))) return rc;
    if (s.ecap > 0 && !s.dge) return SDDDST_ERR_STATE;
    const int64_t ne = s.ne;
    // loss and dL/dR
    hipLaunchKernelGGL(k_probes, dim3(nblk((int64_t)k * n)), dim3(256), 0, c.stream, k, n, seed, s.W);
    launch_apply(c, s, f, k, s.W, s.Y);
    hipLaunchKernelGGL(k_spmv_batch, dim3(nblk(n, 4), k), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.AxRow, s.Y, s.R);
    hipLaunchKernelGGL(k_loss_rows, dim3(nblk(n)), dim3(256), 0, c.stream, k, n, eps, s.W, s.R, s.lrow);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, c.stream, n, s.lrow, 1.0 / n, s.loss);
    // dY = A^T dR (CSC values as the CSR of A^T)
    if ((rc = ensure_ax(c))) return rc;
    hipLaunchKernelGGL(k_spmv_batch, dim3(nblk(n, 4), k), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.Ax, s.R, s.Y);
    hipLaunchKernelGGL(k_dgamma, dim3(nblk(n)), dim3(256), 0, c.stream, n, k, s.Cp, s.Ci, s.Y, s.W, s.se, s.sd, s.dge,
                       s.dgd);
    SD_HIP(hipGetLastError());
    float *a0 = s.act_(0), *h0 = s.act_(1), *z1 = s.act_(2), *S1 = s.act_(3), *ph1 = s.act_(4), *h1 = s.act_(5);
    float *z2 = s.act_(6), *S2 = s.act_(7), *ph2 = s.act_(8), *h2 = s.act_(9);
    // readout
    hipLaunchKernelGGL(k_bwd_readout, dim3(nblk(n, 8)), dim3(256), 0, c.stream, n, s.theta, s.Cp, s.Ci, s.Cmir, s.E, h2,
                       s.dge, s.dgd, s.dh, s.rowg);
    {
        const int per = (n + kPartBlocks - 1) / kPartBlocks;
        const int nblocks = (n + per - 1) / per;
        hipLaunchKernelGGL(k_colsum, dim3(nblocks), dim3(256), 0, c.stream, n, per, lay::NRO, s.rowg, s.part);
        hipLaunchKernelGGL(k_reduce_part, dim3(1), dim3(256), 0, c.stream, nblocks, lay::NRO, kMaxOuter, s.part,
                           s.grad + lay::RO);
        SD_HIP(hipGetLastError());
    }
    // layers 2, 1: dh holds dL/dh_l on entry, dhin receives dL/dh_{l-1}
    const float *zs[3] = {nullptr, z1, z2}, *phs[3] = {nullptr, ph1, ph2}, *Ss[3] = {nullptr, S1, S2};
    const float *hins[3] = {nullptr, h0, h1};
    float *dcur = s.dh, *dnext = s.dhin;
    for (int l = 2; l >= 1; --l) {
        const int base = lay::layer(l);
        hipLaunchKernelGGL(k_bwd_mp, dim3(nblk(n, 8)), dim3(256), 0, c.stream, n, l, s.theta, s.Cp, s.Ci, s.E, zs[l],
                           phs[l], dcur, s.dph, dnext, s.dpe);
        hipLaunchKernelGGL(k_bwd_gather, dim3(nblk(n, 8)), dim3(256), 0, c.stream, n, l, s.theta, s.Cp, s.Cmir, s.dpe,
                           s.dz, dnext);
        SD_HIP(hipGetLastError());
        if ((rc = outer(c, s, n, s.dph, hins[l], 32, 0, base + lay::LU))) return rc;
        if ((rc = outer(c, s, n, s.dph, Ss[l], 32, 1, base + lay::LV))) return rc;
        if ((rc = outer(c, s, n, s.dz, hins[l], 32, 0, base + lay::LP))) return rc;
        if ((rc = outer(c, s, ne, s.dpe, s.E, 4, 1, base + lay::LQ))) return rc;
        float *t = dcur;
        dcur = dnext;
        dnext = t;
    }
    // input layer: dcur = dL/dh0
    hipLaunchKernelGGL(k_relu_back, dim3(nblk((int64_t)n * 32)), dim3(256), 0, c.stream, (int64_t)n * 32, a0, dcur);
    SD_HIP(hipGetLastError());
    if ((rc = outer(c, s, n, dcur, s.U, 4, 1, lay::W0))) return rc;
    if (lr > 0) {
        s.adam_t += 1;
        const double c1 = 1.0 - pow(0.9, s.adam_t), c2 = 1.0 - pow(0.999, s.adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nblk(lay::NPARAM)), dim3(256), 0, c.stream, lay::NPARAM, s.grad, s.theta, s.am,
                           s.av, lr, c1, c2);
        SD_HIP(hipGetLastError());
        s.graph_valid = false;   // outputs no longer match the weights
    }
    if (loss_out) SD_HIP(hipMemcpyAsync(loss_out, s.loss, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (grad_out)
        SD_HIP(hipMemcpyAsync(grad_out, s.grad, sizeof(double) * lay::NPARAM, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    (void)h1;
    (void)h2;
    return SDDDST_OK;
}

int spai_set_weights(Context &c, const float *w)
{
    Spai &s = spai_state(c);
    if (!s.theta) {
        int rc = grow(&s.theta, lay::NPARAM);
        if (rc) return rc;
    }
    SD_HIP(hipMemcpyAsync(s.theta, w, sizeof(float) * lay::NPARAM, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    s.weights = true;
    s.graph_valid = false;
    return SDDDST_OK;
}

int spai_get_weights(Context &c, float *w)
{
    SD_CHECK(c.spai && c.spai->weights, SDDDST_ERR_STATE, "no weights installed");
    SD_HIP(hipMemcpyAsync(w, c.spai->theta, sizeof(float) * lay::NPARAM, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

// Initial weights: matrices uniform in +-1/sqrt(fan_in) from the SplitMix64
// stream of seed (element t of the flat layout), biases and the readout 0.
void spai_init_weights(unsigned long long seed, float *w)
{
    auto mixh = [](unsigned long long z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    struct Blk {
        int off, rows, cols;
    };
    std::vector<Blk> mats = {{lay::W0, lay::H, lay::FN}};
    for (int l = 1; l <= 2; ++l) {
        const int b = lay::layer(l);
        mats.push_back({b + lay::LP, 32, 32});
        mats.push_back({b + lay::LQ, 32, 4});
        mats.push_back({b + lay::LU, 32, 32});
        mats.push_back({b + lay::LV, 32, 32});
    }
    memset(w, 0, sizeof(float) * lay::NPARAM);
    for (const Blk &m : mats) {
        for (int t = m.off; t < m.off + m.rows * m.cols; ++t) {
            const unsigned long long z = mixh(seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(t + 1));
            const double u = (double)(z >> 11) * 0x1.0p-53;
            w[t] = (float)((2.0 * u - 1.0) / sqrt((double)m.cols));
        }
    }
}

int spai_train_init(Context &c, unsigned long long seed)
{
    std::vector<float> w(lay::NPARAM);
    spai_init_weights(seed, w.data());
    int rc = spai_set_weights(c, w.data());
    if (rc) return rc;
    Spai &s = *c.spai;
    s.adam_t = 0;
    if (s.am) {
        SD_HIP(hipMemsetAsync(s.am, 0, sizeof(double) * lay::NPARAM, c.stream));
        SD_HIP(hipMemsetAsync(s.av, 0, sizeof(double) * lay::NPARAM, c.stream));
    }
    return SDDDST_OK;
}

int spai_graph(Context &c, const Fgmres &f, int64_t *ne, int *Cp, int *Ci, double *Ce, double *Cd, float *ge,
               float *gd, double *bd, double *bo, int *bcol)
{
    SD_CHECK(c.spai && c.spai->graph_valid, SDDDST_ERR_STATE, "no learned preconditioner evaluated");
    Spai &s = *c.spai;
    const int n = c.n;
    *ne = s.ne;
    auto d2h = [&](void *dst, const void *src, size_t bytes) -> int {
        if (dst && bytes) SD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
        return SDDDST_OK;
    };
    int rc;
    if ((rc = d2h(Cp, s.Cp, sizeof(int) * (n + 1))) || (rc = d2h(Ci, s.Ci, sizeof(int) * s.ne)) ||
        (rc = d2h(Ce, s.Ce, sizeof(double) * s.ne)) || (rc = d2h(Cd, s.Cd, sizeof(double) * n)) ||
        (rc = d2h(ge, s.ge, sizeof(float) * s.ne)) || (rc = d2h(gd, s.gd, sizeof(float) * n)) ||
        (rc = d2h(bcol, f.bcol, sizeof(int) * n)))
        return rc;
    if (bd || bo) {
        std::vector<double> b(2 * (size_t)n);
        SD_HIP(hipMemcpyAsync(b.data(), f.binv, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c.stream));
        SD_HIP(hipStreamSynchronize(c.stream));
        for (int i = 0; i < n; ++i) {
            if (bd) bd[i] = b[2 * (size_t)i];
            if (bo) bo[i] = b[2 * (size_t)i + 1];
        }
    }
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

void spai_free(Context &c)
{
    if (!c.spai) return;
    Spai &s = *c.spai;
    void *ptrs[] = {s.theta, s.grad, s.am,  s.av,  s.cnt,  s.Cp,   s.Ci,  s.Cr,  s.Ck,   s.Cmir, s.U,    s.E,
                    s.se,    s.sd,   s.act, s.ge,  s.gd,   s.Ce,   s.Cd,  s.W,   s.Y,    s.R,    s.lrow, s.loss,
                    s.dge,   s.dgd,  s.dh,  s.dhin, s.dph, s.dz,   s.dpe, s.rowg, s.part, s.d_total};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete c.spai;
    c.spai = nullptr;
}

}  // namespace sdamd
